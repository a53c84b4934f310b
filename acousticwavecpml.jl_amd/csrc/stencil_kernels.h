// stencil_kernels.h -- gfx950 kernels of one CPML acoustic time step.
//
// One launch of stencil2d/stencil3d performs, for every shot of a plan, what
// the reference does with update_psi_{x,y,z}! followed by update_p!
// (src/shared/acoustic_2D_xPU.jl:7-75, src/shared/acoustic_3D_xPU.jl:8-104):
//
//  * every thread owns one (x, y) column (3D) or one x position (2D) and
//    marches along z (3D) or y (2D) through a chunk of the grid; a wave
//    fetches each pcur line from HBM once (as the node ahead along the
//    march) and re-reads it from L1/L2 as the node itself, the node behind
//    and the in-plane neighbours of the neighbouring lanes and waves;
//  * the march is software-pipelined through a ring of D register slots
//    holding the next D nodes: every load a node needs -- pressures, pold,
//    fact and, inside the CPML slabs, psi and xi -- is issued into the slot
//    the current node has just consumed (so its registers are reused without
//    copies) and stays in flight during the next D-1 nodes' arithmetic;
//  * every loop body issues a fixed number of loads, so the compiler's
//    memory-counter waits are exact and never drain the loads in flight:
//    slab membership across the march is a per-wave template choice (a wave
//    with any lane in an x slab runs the x-slab body, lanes outside the slab
//    read a valid dummy location and do not store), the march is split into
//    low-slab / interior / high-slab segments, the last D nodes of a segment
//    run without refills, and the CPML coefficients of the march axis come
//    from LDS;
//  * addresses are 32-bit byte offsets from per-shot base pointers that are
//    uniform across the block (scalar registers), when a shot's arrays are
//    below 4 GiB; a 64-bit variant covers larger grids;
//  * the psi recursions are fused into the pressure update.  psi is
//    double-buffered (psi_rd holds step n-1, psi_wr receives step n): the
//    thread at node m recomputes psi[m] and psi[m-1] from psi_rd and the
//    three pressures it holds, then writes psi[m] (and psi[m-1] at the
//    side's first layer, psi[0] unless the two sides' slabs overlap), so no
//    second launch and no grid-wide ordering are needed and each psi value
//    is produced with exactly the reference's expression;
//  * only threads inside a CPML slab touch psi/xi, so their traffic is paid
//    in the slabs only; a side whose coefficients are a = 0, b_K = 1 (the
//    free top) is skipped entirely;
//  * the arithmetic is written in the reference's evaluation order and the
//    file is compiled with -ffp-contract=off, so the FP64 instantiation
//    reproduces the reference operation for operation.
//
// Blocks are mapped XCD-aware: the hardware hands block b to XCD b % 8, and
// logical_block() renumbers so that each XCD walks one contiguous range of
// logical tiles (x tile fastest, then y tile, march chunk, shot group).
// Neighbouring tiles, which share halo lines, then run on the same XCD at
// about the same time and meet in its L2.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace awcpml {

constexpr int kTileX = 64;     // one wavefront across x
constexpr int kNumXcd = 8;     // MI355X: 8 XCDs, round-robin block dispatch

template <typename T>
struct CpmlSide {
    const T* psi_rd;  // psi of the previous step
    T* psi_wr;        // psi of this step
    T* xi;            // xi, updated in place
    const T* a;       // "ongrd" a   (halo values)
    const T* bK;      // "ongrd" b/K (halo values)
    const T* ah;      // "halfgrd" a   (halo+1 values)
    const T* bKh;     // "halfgrd" b/K (halo+1 values)
};

template <typename T>
struct StencilArgs {
    int64_t nx, ny, nz;       // nz = 1 in 2D
    int64_t px;               // x pitch (elements) of fields and y/z slabs
    int64_t halo;
    int64_t field_stride;     // elements between two shots' fields
    const T* pold;
    const T* pcur;
    T* pnew;
    const T* fact;
    T inv_d[3];
    T inv_d2[3];
    CpmlSide<T> side[3][2];   // [axis][0 = low/left, 1 = high/right]
    int active[3][2];
    int64_t psi_stride[3];    // elements between two shots' psi slabs
    int64_t xi_stride[3];
    int tiles_x, tiles_y, nchunks, march_len, nshots;
    // shot_group waves of a block march the same rows of different shots
    // (so the shared fact rows are fetched once per block); a block covers
    // blockDim.y / shot_group rows (2D: march chunks) of nsgroups groups
    int shot_group, nsgroups;
    // 1: the x slabs use the lane-shuffle path (each active side's psi
    // layers 0..halo lie in one 64-wide x tile, nx >= 2*halo+3)
    int x_fast;
    int logical_blocks;
    int grid_blocks;          // launched blocks, a multiple of kNumXcd
    // optional: per node the derivative of pnew with respect to fact,
    // q = sum d2 + sum over slabs (dpsi + xi), written for the adjoint's
    // gradient (same layout as the fields, one array per shot); null = off
    T* qout;
};

__device__ __forceinline__ int logical_block(int b, int grid) {
    const int per = grid / kNumXcd;
    return (b % kNumXcd) * per + b / kNumXcd;
}

// Loads/stores of global memory at a byte offset from a base pointer.
// With a uniform base and a 32-bit offset the compiler emits the
// scalar-base addressing form.  The access goes through a global
// (address space 1) pointer: a base read from a device-side argument table
// is otherwise generic to the compiler, whose flat_* accesses also count
// against the LDS wait counter and serialise the march's waits.
template <typename T>
using GlobalPtr = __attribute__((address_space(1))) T*;

template <typename T, typename Off>
__device__ __forceinline__ T ldg(const T* base, Off byte_off) {
    return *(GlobalPtr<const T>)((const char*)base + byte_off);
}
template <typename T, typename Off>
__device__ __forceinline__ void stg(T* base, Off byte_off, T v) {
    *(GlobalPtr<T>)((char*)base + byte_off) = v;
}
// Streaming variants for data touched once per launch (pold is read once,
// pnew is written once and next read in the following launch): the
// nontemporal hint keeps them from displacing the reused pcur and fact
// lines in L2.
template <typename T, typename Off>
__device__ __forceinline__ T ldg_stream(const T* base, Off byte_off) {
    return __builtin_nontemporal_load((GlobalPtr<const T>)((const char*)base + byte_off));
}
template <typename T, typename Off>
__device__ __forceinline__ void stg_stream(T* base, Off byte_off, T v) {
    __builtin_nontemporal_store(v, (GlobalPtr<T>)((char*)base + byte_off));
}

// Classify a node along one axis: side 0/1 and layer m, or side -1.
// Mirrors `if i <= halo+1 ... elseif i >= nx-halo` (1-based) with the
// 0-based node index q: left if q <= halo, right if q >= n-halo-1, where
// the right layer is m = q - (n-halo-2).  An inactive side still takes its
// branch (so the other side is not tested), as in the reference.  On a
// grid with n < 2*halo+2 the two slabs overlap and the left side wins
// (acoustic_2D_xPU.jl:45-56): the right side's first node is then q =
// halo+1, at layer right_first(n, halo) > 1.
__device__ __forceinline__ int classify(int64_t q, int64_t n, int64_t halo,
                                        const int* active, int* m) {
    if (q <= halo) {
        *m = (int)q;
        return active[0] ? 0 : -1;
    } else if (q >= n - halo - 1) {
        *m = (int)(q - (n - halo - 2));
        return active[1] ? 1 : -1;
    }
    return -1;
}

// The six CPML coefficients one layer m needs: b_K_h[m], a_h[m],
// b_K_h[m-1], a_h[m-1] (staggered) and b_K[m-1], a[m-1] (nodal).
__device__ __forceinline__ int right_first(int64_t n, int64_t halo) {
    const int64_t m = 2 * halo + 3 - n;
    return m > 1 ? (int)m : 1;
}

template <typename T>
struct LayerCoef {
    T bKh, ah, bKh1, ah1, bK1, a1;
};

template <typename T>
__device__ __forceinline__ LayerCoef<T> layer_coef(const CpmlSide<T>& s, int m) {
    LayerCoef<T> c;
    c.bKh = s.bKh[m];
    c.ah = s.ah[m];
    c.bKh1 = s.bKh[m - 1];
    c.ah1 = s.ah[m - 1];
    c.bK1 = s.bK[m - 1];
    c.a1 = s.a[m - 1];
    return c;
}

// CPML update of one side of one axis at layer m, from the previous psi
// values psi_old = psi[m], psi_old1 = psi[m-1] and xi_old = xi[m-1]
// (already loaded).  pm, pc, pp: pressure before, at and after the node
// along the axis; d2: the node's second derivative along the axis; f: fact.
// Stores the new psi[m] (and psi[m-1] when `first`: the node is its side's
// first slab node, whose psi[m-1] no other node writes) and xi[m-1]; returns
// dpsi + xi, which the caller scales by fact into the reference's damping
// term damp += fact*(dpsi + xi) (acoustic_2D_xPU.jl:7-18 and 45-56 for
// x-left; the other sides are identical up to index shifts).
template <typename T, typename Off>
__device__ __forceinline__ T cpml_apply(const LayerCoef<T>& c, T* psi_wr, T* xi_arr, Off psi_m,
                                        Off psi_m1, Off xi_m1, bool first, T psi_old,
                                        T psi_old1, T xi_old, T pm, T pc, T pp, T d2,
                                        T inv_d, T f, bool st = true) {
    // update_psi_*!: psi[m] = b_K_h[m]*psi[m] + a_h[m]*(p[m+1]-p[m])*_d
    const T psi_here = c.bKh * psi_old + c.ah * (pp - pc) * inv_d;
    const T psi_prev = c.bKh1 * psi_old1 + c.ah1 * (pc - pm) * inv_d;
    if (st) stg(psi_wr, psi_m, psi_here);
    if (st && first) stg(psi_wr, psi_m1, psi_prev);
    // update_p!: dpsi, in-place xi, damping
    const T dpsi = (psi_here - psi_prev) * inv_d;
    const T xin = c.bK1 * xi_old + c.a1 * (d2 + dpsi);
    if (st) stg(xi_arr, xi_m1, xin);
    return dpsi + xin;
}

// Pointers of one slab side for one shot, selected without memory loads.
template <typename T>
struct SlabPtrs {
    const T* psi_rd;
    T* psi_wr;
    T* xi;
};

template <typename T>
__device__ __forceinline__ SlabPtrs<T> slab_ptrs(const StencilArgs<T>& A, int axis, int side,
                                                 int shot) {
    const CpmlSide<T>& L = A.side[axis][0];
    const CpmlSide<T>& R = A.side[axis][1];
    SlabPtrs<T> p;
    p.psi_rd = (side ? R.psi_rd : L.psi_rd) + shot * A.psi_stride[axis];
    p.psi_wr = (side ? R.psi_wr : L.psi_wr) + shot * A.psi_stride[axis];
    p.xi = (side ? R.xi : L.xi) + shot * A.xi_stride[axis];
    return p;
}

// An in-plane CPML slab of one column (x in 3D and 2D, y in 3D): side,
// layer, pointers, coefficients and byte offsets of psi[m], psi[m-1],
// xi[m-1] at march position 0 plus their increments per march step.
template <typename T, typename Off>
struct ColSlab {
    bool in;            // this lane's column is in the slab (1 <= m <= halo)
    bool own;           // lane-shuffle x path: this lane computes psi[m], 0 <= m <= halo
    bool first;         // in, and the first slab node of its side (also writes psi[m-1])
    int m;
    SlabPtrs<T> p;
    LayerCoef<T> c;
    Off o_psi, o_psi1, o_xi;
    Off st_psi, st_xi;
};

// March-axis CPML coefficients staged in LDS: for side s, ah, bKh (halo+1
// values each) then a, bK (halo values each).
template <typename T>
__device__ __forceinline__ void stage_march_coef(const StencilArgs<T>& A, int axis, T* lds) {
    const int h1 = (int)A.halo + 1;
    const int nthreads = blockDim.x * blockDim.y;
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    for (int q = tid; q < 8 * h1; q += nthreads) {
        const int s = q / (4 * h1);
        const int r = q % (4 * h1);
        const int w = r / h1, m = r % h1;
        const CpmlSide<T>& S = A.side[axis][s];
        T v = T(0);
        if (w == 0) v = S.ah[m];
        else if (w == 1) v = S.bKh[m];
        else if (m < h1 - 1) v = (w == 2) ? S.a[m] : S.bK[m];
        lds[q] = v;
    }
}

template <typename T>
__device__ __forceinline__ LayerCoef<T> lds_coef(const T* lds, int h1, int side, int m) {
    const T* b = lds + side * 4 * h1;
    LayerCoef<T> c;
    c.ah = b[m];
    c.ah1 = b[m - 1];
    c.bKh = b[h1 + m];
    c.bKh1 = b[h1 + m - 1];
    c.a1 = b[2 * h1 + m - 1];
    c.bK1 = b[3 * h1 + m - 1];
    return c;
}

// Per-node loads of the march (one ring slot).
template <typename T>
struct NodeLoads {
    T pm, pc, pp;         // pcur one node behind, at, and ahead along the march
    T xm, xp, ym, yp;     // in-plane neighbours (ym/yp unused in 2D)
    T po, f;              // pold, fact
    T psa0, psa1, xia;    // first in-plane slab (x): psi[m], psi[m-1], xi[m-1]
    T psb0, psb1, xib;    // second in-plane slab (y, 3D only)
    T psm0, psm1, xim;    // march-axis slab
};

// Everything a column's march needs; built once per thread.
template <typename T, typename Off, int NDIM>
struct March {
    const T* pcur;
    const T* pold;
    T* pnew;
    const T* fact;
    Off col;      // byte offset of the column at march position 0
    Off sm;       // byte stride along the march axis
    Off sy;       // byte stride along y (3D: in-plane y; 2D: == sm)
    T i2a, i2b, i2m;  // 1/d^2 of x, y (3D), march axis
    T ia, ib, imm;    // 1/d   of x, y (3D), march axis
    ColSlab<T, Off> sa, sb;   // x slab, y slab (3D)
    SlabPtrs<T> pm_side[2];   // march-axis slab pointers per side
    int mshift[2];            // march layer = node - mshift[side]
    int mfirst[2];            // march layer of each side's first slab node
    const T* lds;             // march-axis coefficients
    int h1;
    T* qout;                  // optional d(pnew)/d(fact) output (this shot)
    bool store_p;             // false on the x = 0 lane kept for the x-slab psi[0]
};

template <typename T, typename Off, int NDIM, int SA, bool SB, bool SM>
__device__ __forceinline__ void load_node(const March<T, Off, NDIM>& M, int k, int mside,
                                          NodeLoads<T>& q) {
    constexpr Off E = (Off)sizeof(T);
    const Off c = M.col + (Off)k * M.sm;
    q.pm = ldg(M.pcur, c - M.sm);
    q.pc = ldg(M.pcur, c);
    q.pp = ldg(M.pcur, c + M.sm);
    q.xm = ldg(M.pcur, c - E);
    q.xp = ldg(M.pcur, c + E);
    if (NDIM == 3) {
        q.ym = ldg(M.pcur, c - M.sy);
        q.yp = ldg(M.pcur, c + M.sy);
    }
    q.po = ldg_stream(M.pold, c);
    q.f = ldg(M.fact, c);
    if (SA == 1) {  // lane-shuffle path: psi[m] of the lane's layer, xi[m-1]
        q.psa0 = ldg(M.sa.p.psi_rd, M.sa.o_psi + (Off)k * M.sa.st_psi);
        q.xia = ldg((const T*)M.sa.p.xi, M.sa.o_xi + (Off)k * M.sa.st_xi);
    } else if (SA == 2) {
        q.psa0 = ldg(M.sa.p.psi_rd, M.sa.o_psi + (Off)k * M.sa.st_psi);
        q.psa1 = ldg(M.sa.p.psi_rd, M.sa.o_psi1 + (Off)k * M.sa.st_psi);
        q.xia = ldg((const T*)M.sa.p.xi, M.sa.o_xi + (Off)k * M.sa.st_xi);
    }
    if (SB) {
        q.psb0 = ldg(M.sb.p.psi_rd, M.sb.o_psi + (Off)k * M.sb.st_psi);
        q.psb1 = ldg(M.sb.p.psi_rd, M.sb.o_psi1 + (Off)k * M.sb.st_psi);
        q.xib = ldg((const T*)M.sb.p.xi, M.sb.o_xi + (Off)k * M.sb.st_xi);
    }
    if (SM) {
        // march-axis slab arrays share the field layout with the march axis
        // replaced by the layer index: offset col + layer*sm
        const int mz = k - M.mshift[mside];
        const SlabPtrs<T>& P = M.pm_side[mside];
        q.psm0 = ldg(P.psi_rd, M.col + (Off)mz * M.sm);
        q.psm1 = ldg(P.psi_rd, M.col + (Off)(mz - 1) * M.sm);
        q.xim = ldg((const T*)P.xi, M.col + (Off)(mz - 1) * M.sm);
    }
}

template <typename T, typename Off, int NDIM, int SA, bool SB, bool SM>
__device__ __forceinline__ void compute_node(const March<T, Off, NDIM>& M, int k, int mside,
                                             const NodeLoads<T>& q) {
    const T two = T(2);
    const T pm = q.pm, pc = q.pc, pp = q.pp;
    const T f = q.f;
    const T d2a = (q.xp - two * pc + q.xm) * M.i2a;
    T d2b = T(0);
    if (NDIM == 3) d2b = (q.yp - two * pc + q.ym) * M.i2b;
    const T d2m = (pp - two * pc + pm) * M.i2m;
    T damp = T(0);
    T slab = T(0);  // sum of (dpsi + xi) over the node's slabs, for qout
    // Only lane-shuffle x waves hold the x = 0 lane, which computes psi[0]
    // of the low x side and stores nothing else.
    const bool st = SA != 1 || M.store_p;
    if (SA == 1) {
        // psi[m] of this lane's layer; psi[m-1] is the same expression
        // evaluated by lane i-1 (layer m-1), taken with a lane shift
        const T ph = M.sa.c.bKh * q.psa0 + M.sa.c.ah * (q.xp - pc) * M.ia;
        const T pprev = __shfl_up(ph, 1u, 64);
        if (M.sa.own) stg(M.sa.p.psi_wr, M.sa.o_psi + (Off)k * M.sa.st_psi, ph);
        if (M.sa.in) {
            const T dpsi = (ph - pprev) * M.ia;
            const T xin = M.sa.c.bK1 * q.xia + M.sa.c.a1 * (d2a + dpsi);
            stg(M.sa.p.xi, M.sa.o_xi + (Off)k * M.sa.st_xi, xin);
            const T t = dpsi + xin;
            damp += f * t;
            slab += t;
        }
    } else if (SA == 2) {
        if (M.sa.in) {
            const T t = cpml_apply(M.sa.c, M.sa.p.psi_wr, M.sa.p.xi, M.sa.o_psi + (Off)k * M.sa.st_psi,
                                   M.sa.o_psi1 + (Off)k * M.sa.st_psi, M.sa.o_xi + (Off)k * M.sa.st_xi,
                                   M.sa.first, q.psa0, q.psa1, q.xia, q.xm, pc, q.xp, d2a, M.ia, f);
            damp += f * t;
            slab += t;
        }
    }
    if (NDIM == 3 && SB) {
        const T t = cpml_apply(M.sb.c, M.sb.p.psi_wr, M.sb.p.xi, M.sb.o_psi + (Off)k * M.sb.st_psi,
                               M.sb.o_psi1 + (Off)k * M.sb.st_psi, M.sb.o_xi + (Off)k * M.sb.st_xi,
                               M.sb.first, q.psb0, q.psb1, q.xib, q.ym, pc, q.yp, d2b, M.ib, f, st);
        damp += f * t;
        slab += t;
    }
    if (SM) {
        const int mz = k - M.mshift[mside];
        const LayerCoef<T> cz = lds_coef(M.lds, M.h1, mside, mz);
        const SlabPtrs<T>& P = M.pm_side[mside];
        const T t = cpml_apply(cz, P.psi_wr, P.xi, M.col + (Off)mz * M.sm,
                               M.col + (Off)(mz - 1) * M.sm, M.col + (Off)(mz - 1) * M.sm,
                               mz == M.mfirst[mside], q.psm0, q.psm1, q.xim, pm, pc, pp, d2m, M.imm,
                               f, st);
        damp += f * t;
        slab += t;
    }
    // update_p!: pnew = 2 pcur - pold + fact*(d2x + d2y (+ d2z)) + damp
    T lap;
    if (NDIM == 3) lap = d2a + d2b + d2m;   // (d2x + d2y) + d2z
    else lap = d2a + d2m;                   // d2x + d2y
    if (st) stg_stream(M.pnew, M.col + (Off)k * M.sm, two * pc - q.po + f * lap + damp);
    if (st && M.qout) stg(M.qout, M.col + (Off)k * M.sm, lap + slab);
}

// March over nodes [k_from, k_to) with a fixed load set.  Each node's slot
// holds all its operands (pcur below/at/above included), so no value is
// carried from one slot to the next and refills never need register copies.
template <typename T, typename Off, int NDIM, int D, int SA, bool SB, bool SM>
__device__ __forceinline__ void march_segment(const March<T, Off, NDIM>& M, int k_from, int k_to,
                                              int mside) {
    if (k_to - k_from < 2 * D) {
        // short segment: no pipeline
        for (int k = k_from; k < k_to; ++k) {
            NodeLoads<T> q;
            load_node<T, Off, NDIM, SA, SB, SM>(M, k, mside, q);
            compute_node<T, Off, NDIM, SA, SB, SM>(M, k, mside, q);
        }
        return;
    }
    // unconditional prologue in steady-state order, so the waits the
    // compiler derives at the loop head (a minimum over the entry and the
    // back edge) stay close to the steady-state ones
    NodeLoads<T> ring[D];
#pragma unroll
    for (int u = 0; u < D; ++u) {
        load_node<T, Off, NDIM, SA, SB, SM>(M, k_from + u, mside, ring[u]);
        __builtin_amdgcn_sched_barrier(0);
    }
    int k0 = k_from;
    // steady state: every node refills its slot with node k+D.  The
    // scheduling barrier keeps each refill right after its node: left free,
    // the scheduler sinks all refills to the end of the unrolled body and
    // the next iteration waits on them at once.
    for (; k0 + 2 * D <= k_to; k0 += D) {
#pragma unroll
        for (int u = 0; u < D; ++u) {
            const int k = k0 + u;
            compute_node<T, Off, NDIM, SA, SB, SM>(M, k, mside, ring[u]);
            load_node<T, Off, NDIM, SA, SB, SM>(M, k + D, mside, ring[u]);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    // tail: at most 2D-1 nodes, refills only while inside the segment
#pragma unroll
    for (int u = 0; u < 2 * D - 1; ++u) {
        const int k = k0 + u;
        if (k < k_to) {
            compute_node<T, Off, NDIM, SA, SB, SM>(M, k, mside, ring[u % D]);
            if (k + D < k_to) load_node<T, Off, NDIM, SA, SB, SM>(M, k + D, mside, ring[u % D]);
        }
    }
}

// Two-distance pipeline.  The loads of a node fall in two groups:
//  - far: the new pcur plane ahead of the march, pold and fact, the only
//    loads that normally come from HBM;
//  - near: the in-plane neighbours (xm, xp, ym, yp: the same lines as pcur
//    of this and the neighbouring rows, cached) and the slab values.
// Far loads are issued F nodes ahead and near loads DN nodes ahead (DN <=
// F), near before far within each node's issue group: a wave then waits
// only for the near loads of the node it computes while the far loads of
// the next F nodes stay in flight, which keeps more HBM bytes outstanding
// per wave than one ring of whole-node slots at the same register cost.
// pcur values along the march are kept in a register ring of CR >= F + 2
// entries (each plane is loaded once and used as pp, pc and pm); pold/fact
// use a ring of CR entries as well so every slot index is a compile-time
// constant in the loop unrolled CR times (CR is a multiple of DN).
template <int P>
struct Pipe;
template <>
struct Pipe<4> {
    static constexpr int CR = 5, F = 3, DN = 1;
};
template <>
struct Pipe<5> {
    static constexpr int CR = 6, F = 4, DN = 2;
};

template <typename T, typename Off, int NDIM, int SA, bool SB, bool SM>
__device__ __forceinline__ void load_near(const March<T, Off, NDIM>& M, int k, int mside,
                                          NodeLoads<T>& q) {
    constexpr Off E = (Off)sizeof(T);
    const Off c = M.col + (Off)k * M.sm;
    q.xm = ldg(M.pcur, c - E);
    q.xp = ldg(M.pcur, c + E);
    if (NDIM == 3) {
        q.ym = ldg(M.pcur, c - M.sy);
        q.yp = ldg(M.pcur, c + M.sy);
    }
    if (SA == 1) {  // lane-shuffle path: psi[m] of the lane's layer, xi[m-1]
        q.psa0 = ldg(M.sa.p.psi_rd, M.sa.o_psi + (Off)k * M.sa.st_psi);
        q.xia = ldg((const T*)M.sa.p.xi, M.sa.o_xi + (Off)k * M.sa.st_xi);
    } else if (SA == 2) {
        q.psa0 = ldg(M.sa.p.psi_rd, M.sa.o_psi + (Off)k * M.sa.st_psi);
        q.psa1 = ldg(M.sa.p.psi_rd, M.sa.o_psi1 + (Off)k * M.sa.st_psi);
        q.xia = ldg((const T*)M.sa.p.xi, M.sa.o_xi + (Off)k * M.sa.st_xi);
    }
    if (SB) {
        q.psb0 = ldg(M.sb.p.psi_rd, M.sb.o_psi + (Off)k * M.sb.st_psi);
        q.psb1 = ldg(M.sb.p.psi_rd, M.sb.o_psi1 + (Off)k * M.sb.st_psi);
        q.xib = ldg((const T*)M.sb.p.xi, M.sb.o_xi + (Off)k * M.sb.st_xi);
    }
    if (SM) {
        const int mz = k - M.mshift[mside];
        const SlabPtrs<T>& P = M.pm_side[mside];
        q.psm0 = ldg(P.psi_rd, M.col + (Off)mz * M.sm);
        q.psm1 = ldg(P.psi_rd, M.col + (Off)(mz - 1) * M.sm);
        q.xim = ldg((const T*)P.xi, M.col + (Off)(mz - 1) * M.sm);
    }
}

template <typename T, typename Off, int NDIM, int PK, int SA, bool SB, bool SM>
__device__ __forceinline__ void march_segment_far(const March<T, Off, NDIM>& M, int k_from, int k_to,
                                                  int mside) {
    constexpr int CR = Pipe<PK>::CR, F = Pipe<PK>::F, DN = Pipe<PK>::DN;
    static_assert(CR >= F + 2 && CR % DN == 0 && DN <= F, "pipeline shape");
    if (k_to - k_from < F + 2) {
        for (int k = k_from; k < k_to; ++k) {
            NodeLoads<T> q;
            load_node<T, Off, NDIM, SA, SB, SM>(M, k, mside, q);
            compute_node<T, Off, NDIM, SA, SB, SM>(M, k, mside, q);
        }
        return;
    }
    T col[CR];   // pcur of node m in slot (m - k_from + 1) % CR
    T po[CR];    // pold of node m in slot (m - k_from) % CR
    T fa[CR];    // fact of node m in slot (m - k_from) % CR
    NodeLoads<T> nr[DN];  // near loads of node m in slot (m - k_from) % DN
    // prologue, in steady-state order: near loads first, then far loads
#pragma unroll
    for (int u = 0; u < DN; ++u) load_near<T, Off, NDIM, SA, SB, SM>(M, k_from + u, mside, nr[u]);
#pragma unroll
    for (int u = 0; u < F + 2; ++u) col[u] = ldg(M.pcur, M.col + (Off)(k_from - 1 + u) * M.sm);
#pragma unroll
    for (int u = 0; u < F; ++u) {
        po[u] = ldg_stream(M.pold, M.col + (Off)(k_from + u) * M.sm);
        fa[u] = ldg(M.fact, M.col + (Off)(k_from + u) * M.sm);
    }
    __builtin_amdgcn_sched_barrier(0);
    int k0 = k_from;
    for (; k0 + CR + F <= k_to; k0 += CR) {
#pragma unroll
        for (int u = 0; u < CR; ++u) {
            const int k = k0 + u;
            NodeLoads<T>& q = nr[u % DN];
            q.pm = col[u % CR];
            q.pc = col[(u + 1) % CR];
            q.pp = col[(u + 2) % CR];
            q.po = po[u % CR];
            q.f = fa[u % CR];
            compute_node<T, Off, NDIM, SA, SB, SM>(M, k, mside, q);
            load_near<T, Off, NDIM, SA, SB, SM>(M, k + DN, mside, nr[u % DN]);
            col[(u + F + 2) % CR] = ldg(M.pcur, M.col + (Off)(k + F + 1) * M.sm);
            po[(u + F) % CR] = ldg_stream(M.pold, M.col + (Off)(k + F) * M.sm);
            fa[(u + F) % CR] = ldg(M.fact, M.col + (Off)(k + F) * M.sm);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    // tail: fewer than CR + F nodes, loads only inside the segment
#pragma unroll
    for (int u = 0; u < CR + F - 1; ++u) {
        const int k = k0 + u;
        if (k < k_to) {
            NodeLoads<T>& q = nr[u % DN];
            q.pm = col[u % CR];
            q.pc = col[(u + 1) % CR];
            q.pp = col[(u + 2) % CR];
            q.po = po[u % CR];
            q.f = fa[u % CR];
            compute_node<T, Off, NDIM, SA, SB, SM>(M, k, mside, q);
            if (k + DN < k_to) load_near<T, Off, NDIM, SA, SB, SM>(M, k + DN, mside, nr[u % DN]);
            if (k + F + 1 <= k_to) col[(u + F + 2) % CR] = ldg(M.pcur, M.col + (Off)(k + F + 1) * M.sm);
            if (k + F < k_to) {
                po[(u + F) % CR] = ldg_stream(M.pold, M.col + (Off)(k + F) * M.sm);
                fa[(u + F) % CR] = ldg(M.fact, M.col + (Off)(k + F) * M.sm);
            }
        }
    }
}

// Split [kbeg, kend) into low-slab / interior / high-slab segments of the
// march axis.
template <typename T, typename Off, int NDIM, int D, int SA, bool SB>
__device__ __forceinline__ void march_column(const March<T, Off, NDIM>& M, const int* active_m,
                                             int64_t n_m, int64_t halo, int kbeg, int kend) {
    // the high side starts after the low side's nodes even when the low
    // side is idle (overlapping slabs: the low branch wins, classify())
    const int lo_end = active_m[0] ? (int)min((int64_t)kend, halo + 1) : kbeg;
    const int hi_beg =
        active_m[1] ? (int)max((int64_t)kbeg, max(n_m - halo - 1, halo + 1)) : kend;
    const int a = max(kbeg, lo_end), b = min(kend, max(a, hi_beg));
    if constexpr (D >= 4) {
        march_segment_far<T, Off, NDIM, D, SA, SB, true>(M, kbeg, min(lo_end, kend), 0);
        march_segment_far<T, Off, NDIM, D, SA, SB, false>(M, a, b, 0);
        march_segment_far<T, Off, NDIM, D, SA, SB, true>(M, max(b, hi_beg), kend, 1);
    } else {
        march_segment<T, Off, NDIM, D, SA, SB, true>(M, kbeg, min(lo_end, kend), 0);
        march_segment<T, Off, NDIM, D, SA, SB, false>(M, a, b, 0);
        march_segment<T, Off, NDIM, D, SA, SB, true>(M, max(b, hi_beg), kend, 1);
    }
}

// Fill a ColSlab for an in-plane axis.  q: node index along the axis;
// psi_base(m)/xi_base(m): byte offsets of layer m at march position 0;
// steps: byte increments per march step.  Lanes outside the slab get a
// valid dummy location (layer 1 of the low side) and in = false.
template <typename T, typename Off>
__device__ __forceinline__ void col_slab(const StencilArgs<T>& A, int axis, int64_t q, int64_t n,
                                         int shot, Off psi_l0, Off psi_dl, Off xi_l0, Off xi_dl,
                                         Off st_psi, Off st_xi, ColSlab<T, Off>& S) {
    int m = 1;
    int side = classify(q, n, A.halo, A.active[axis], &m);
    S.in = side >= 0;
    if (!S.in) {
        side = 0;
        m = 1;
    }
    S.own = false;
    S.first = S.in && m == (side ? right_first(n, A.halo) : 1);
    S.m = m;
    S.p = slab_ptrs(A, axis, side, shot);
    S.c = LayerCoef<T>{};
    if (S.in) S.c = layer_coef(A.side[axis][side], m);
    S.o_psi = psi_l0 + (Off)m * psi_dl;
    S.o_psi1 = psi_l0 + (Off)(m - 1) * psi_dl;
    S.o_xi = xi_l0 + (Off)(m - 1) * xi_dl;
    S.st_psi = st_psi;
    S.st_xi = st_xi;
}

// x slab for the lane-shuffle path: lanes q = 0..halo (low side) and
// q = n-halo-2..n-2 (high side) own psi layer m = q - offset, 0..halo; the
// lanes with m >= 1 are the slab nodes.  Same if/elseif order as
// classify(); the sides cannot overlap because n >= 2*halo+3.
template <typename T, typename Off>
__device__ __forceinline__ void col_slab_xfast(const StencilArgs<T>& A, int64_t q, int64_t n,
                                               int shot, Off psi_l0, Off psi_dl, Off xi_l0,
                                               Off xi_dl, Off st_psi, Off st_xi, ColSlab<T, Off>& S) {
    const int64_t h = A.halo;
    int side = -1, m = 0;
    if (q <= h) {
        if (A.active[0][0]) {
            side = 0;
            m = (int)q;
        }
    } else if (q >= n - h - 2 && q <= n - 2 && A.active[0][1]) {
        side = 1;
        m = (int)(q - (n - h - 2));
    }
    S.own = side >= 0;
    S.in = S.own && m >= 1;
    S.first = false;  // the owning lanes write every psi layer themselves
    if (!S.own) {
        side = 0;
        m = 1;  // a valid dummy location
    }
    S.m = m;
    S.p = slab_ptrs(A, 0, side, shot);
    S.c = LayerCoef<T>{};
    if (S.own) {
        const CpmlSide<T>& cs = A.side[0][side];
        S.c.bKh = cs.bKh[m];
        S.c.ah = cs.ah[m];
        if (m >= 1) {
            S.c.bK1 = cs.bK[m - 1];
            S.c.a1 = cs.a[m - 1];
        }
    }
    S.o_psi = psi_l0 + (Off)m * psi_dl;
    S.o_psi1 = S.o_psi;
    S.o_xi = xi_l0 + (Off)(m >= 1 ? m - 1 : 0) * xi_dl;
    S.st_psi = st_psi;
    S.st_xi = st_xi;
}

// ---------------------------------------------------------------------
// 3D: z-march.  Block (64, BY): lane -> x, threadIdx.y -> y (one row per
// wave, so y is wave-uniform).  D = ring depth.
// ---------------------------------------------------------------------
// One block of a 3D launch: logical tile `lb` of the plan described by A.
template <typename T, typename Off, int D>
__device__ __forceinline__ void stencil3d_block(const StencilArgs<T>& A, int lb, T* lds) {
    if (A.active[2][0] || A.active[2][1]) stage_march_coef(A, 2, lds);
    __syncthreads();

    // logical order: x tile fastest, then y tile, chunk, shot group, so the
    // blocks resident on one XCD at a time cover whole planes of one shot
    // (their halo rows and lines are shared through that XCD's L2); with
    // the shot fastest instead, an 8-shot launch moved ~10 % more bytes
    int rest = lb;
    const int tx = rest % A.tiles_x; rest /= A.tiles_x;
    const int ty = rest % A.tiles_y; rest /= A.tiles_y;
    const int chunk = rest % A.nchunks;
    const int sg = rest / A.nchunks;
    // blockDim.x == 64 == wavefront: threadIdx.y is the same on all lanes
    const int wy = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int G = A.shot_group;
    const int shot = sg * G + wy % G;
    const int rows = (int)blockDim.y / G;
    const int64_t i = (int64_t)tx * kTileX + threadIdx.x;
    const int64_t j = 1 + (int64_t)ty * rows + wy / G;
    const int64_t nx = A.nx, ny = A.ny, nz = A.nz, halo = A.halo;
    if (shot >= A.nshots) return;
    if (j > ny - 2) return;
    // the x = 0 lane stays when it computes psi[0] of the low x side
    if (i < ((A.x_fast && A.active[0][0]) ? 0 : 1) || i > nx - 2) return;
    const int kbeg = 1 + chunk * A.march_len;
    const int kend = (int)min((int64_t)kbeg + A.march_len, nz - 1);
    if (kbeg >= kend) return;

    constexpr Off E = (Off)sizeof(T);
    const int64_t fo = (int64_t)shot * A.field_stride;
    March<T, Off, 3> M;
    M.pcur = A.pcur + fo;
    M.pold = A.pold + fo;
    M.pnew = A.pnew + fo;
    M.fact = A.fact;
    M.qout = A.qout ? A.qout + fo : nullptr;
    M.sy = (Off)A.px * E;
    M.sm = (Off)(A.px * ny) * E;
    M.col = ((Off)i + (Off)A.px * (Off)j) * E;
    M.i2a = A.inv_d2[0];
    M.i2b = A.inv_d2[1];
    M.i2m = A.inv_d2[2];
    M.ia = A.inv_d[0];
    M.ib = A.inv_d[1];
    M.imm = A.inv_d[2];
    M.lds = lds;
    M.h1 = (int)halo + 1;
    const Off hp1 = (Off)halo + 1, hl = (Off)halo;
    M.store_p = i >= 1;
    // psi_x: (halo+1) x ny x nz, xi_x: halo x ny x nz
    if (A.x_fast)
        col_slab_xfast<T, Off>(A, i, nx, shot, hp1 * (Off)j * E, E, hl * (Off)j * E, E,
                               hp1 * (Off)ny * E, hl * (Off)ny * E, M.sa);
    else
        col_slab<T, Off>(A, 0, i, nx, shot, hp1 * (Off)j * E, E, hl * (Off)j * E, E,
                         hp1 * (Off)ny * E, hl * (Off)ny * E, M.sa);
    // psi_y: px x (halo+1) x nz, xi_y: px x halo x nz
    col_slab<T, Off>(A, 1, j, ny, shot, (Off)i * E, M.sy, (Off)i * E, M.sy, hp1 * M.sy,
                     hl * M.sy, M.sb);
    // psi_z: px x ny x (halo+1), xi_z: px x ny x halo
    for (int s = 0; s < 2; ++s) M.pm_side[s] = slab_ptrs(A, 2, s, shot);
    M.mshift[0] = 0;
    M.mshift[1] = (int)(nz - halo - 2);
    M.mfirst[0] = 1;
    M.mfirst[1] = right_first(nz, halo);

    const bool xs = A.x_fast ? __any(M.sa.own) : __any(M.sa.in);
    const bool ys = M.sb.in;  // wave-uniform
    if (xs && A.x_fast) {
        if (ys) march_column<T, Off, 3, D, 1, true>(M, A.active[2], nz, halo, kbeg, kend);
        else march_column<T, Off, 3, D, 1, false>(M, A.active[2], nz, halo, kbeg, kend);
    } else if (xs) {
        if (ys) march_column<T, Off, 3, D, 2, true>(M, A.active[2], nz, halo, kbeg, kend);
        else march_column<T, Off, 3, D, 2, false>(M, A.active[2], nz, halo, kbeg, kend);
    } else {
        if (ys) march_column<T, Off, 3, D, 0, true>(M, A.active[2], nz, halo, kbeg, kend);
        else march_column<T, Off, 3, D, 0, false>(M, A.active[2], nz, halo, kbeg, kend);
    }
}

// ---------------------------------------------------------------------
// 2D: y-march.  Block (64, BY): lane -> x, threadIdx.y -> march chunk.
// ---------------------------------------------------------------------
template <typename T, typename Off, int D>
__device__ __forceinline__ void stencil2d_block(const StencilArgs<T>& A, int lb, T* lds) {
    if (A.active[1][0] || A.active[1][1]) stage_march_coef(A, 1, lds);
    __syncthreads();

    int rest = lb;  // x tile fastest, then chunk group, shot group
    const int tx = rest % A.tiles_x; rest /= A.tiles_x;
    const int ty = rest % A.tiles_y;  // group of blockDim.y / shot_group chunks
    const int sg = rest / A.tiles_y;
    const int wy = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int G = A.shot_group;
    const int shot = sg * G + wy % G;
    const int chunk = ty * ((int)blockDim.y / G) + wy / G;
    const int64_t i = (int64_t)tx * kTileX + threadIdx.x;
    const int64_t nx = A.nx, ny = A.ny, halo = A.halo;
    if (shot >= A.nshots) return;
    if (chunk >= A.nchunks) return;
    if (i < ((A.x_fast && A.active[0][0]) ? 0 : 1) || i > nx - 2) return;
    const int jbeg = 1 + chunk * A.march_len;
    const int jend = (int)min((int64_t)jbeg + A.march_len, ny - 1);
    if (jbeg >= jend) return;

    constexpr Off E = (Off)sizeof(T);
    const int64_t fo = (int64_t)shot * A.field_stride;
    March<T, Off, 2> M;
    M.pcur = A.pcur + fo;
    M.pold = A.pold + fo;
    M.pnew = A.pnew + fo;
    M.fact = A.fact;
    M.qout = A.qout ? A.qout + fo : nullptr;
    M.sy = (Off)A.px * E;
    M.sm = M.sy;
    M.col = (Off)i * E;
    M.i2a = A.inv_d2[0];
    M.i2b = T(0);
    M.i2m = A.inv_d2[1];
    M.ia = A.inv_d[0];
    M.ib = T(0);
    M.imm = A.inv_d[1];
    M.lds = lds;
    M.h1 = (int)halo + 1;
    const Off hp1 = (Off)halo + 1, hl = (Off)halo;
    M.store_p = i >= 1;
    // psi_x: (halo+1) x ny, xi_x: halo x ny
    if (A.x_fast) col_slab_xfast<T, Off>(A, i, nx, shot, 0, E, 0, E, hp1 * E, hl * E, M.sa);
    else col_slab<T, Off>(A, 0, i, nx, shot, 0, E, 0, E, hp1 * E, hl * E, M.sa);
    M.sb = M.sa;
    M.sb.in = false;
    // psi_y: px x (halo+1), xi_y: px x halo
    for (int s = 0; s < 2; ++s) M.pm_side[s] = slab_ptrs(A, 1, s, shot);
    M.mshift[0] = 0;
    M.mshift[1] = (int)(ny - halo - 2);
    M.mfirst[0] = 1;
    M.mfirst[1] = right_first(ny, halo);

    if (A.x_fast && __any(M.sa.own))
        march_column<T, Off, 2, D, 1, false>(M, A.active[1], ny, halo, jbeg, jend);
    else if (!A.x_fast && __any(M.sa.in))
        march_column<T, Off, 2, D, 2, false>(M, A.active[1], ny, halo, jbeg, jend);
    else
        march_column<T, Off, 2, D, 0, false>(M, A.active[1], ny, halo, jbeg, jend);
}

// ---------------------------------------------------------------------
// Entry kernels.  A plain launch covers one plan; a batch launch covers up
// to kMaxBatch plans (different grids, same precision) in one grid: block b
// (after the XCD renumbering) belongs to the plan p with
// block_start[p] <= b < block_start[p+1].  A plan whose run is shorter than
// the batch's is skipped once it0 reaches its nt.
// ---------------------------------------------------------------------
constexpr int kMaxBatch = 32;

template <typename T>
struct BatchArgs {
    const StencilArgs<T>* table;      // per-plan arguments for this step's rotation
    int64_t it0;                      // 0-based step index
    int64_t nt[kMaxBatch];
    int block_start[kMaxBatch + 1];
    int nplans;
    int grid_blocks;
};

__device__ __forceinline__ int batch_plan(const int* start, int nplans, int lb) {
    int p = 0;
    for (int q = 1; q < nplans; ++q)
        if (lb >= start[q]) p = q;
    return p;
}

template <typename T, typename Off, int D>
__global__ __launch_bounds__(512) void stencil3d(const StencilArgs<T> A) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int lb = logical_block(blockIdx.x, A.grid_blocks);
    if (lb >= A.logical_blocks) return;  // whole block
    stencil3d_block<T, Off, D>(A, lb, reinterpret_cast<T*>(smem_raw));
}

template <typename T, typename Off, int D>
__global__ __launch_bounds__(512) void stencil2d(const StencilArgs<T> A) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int lb = logical_block(blockIdx.x, A.grid_blocks);
    if (lb >= A.logical_blocks) return;  // whole block
    stencil2d_block<T, Off, D>(A, lb, reinterpret_cast<T*>(smem_raw));
}

template <typename T, typename Off, int D, int NDIM>
__global__ __launch_bounds__(512) void stencil_batch(const BatchArgs<T> B) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int lb = logical_block(blockIdx.x, B.grid_blocks);
    if (lb >= B.block_start[B.nplans]) return;  // whole block
    const int p = batch_plan(B.block_start, B.nplans, lb);
    if (B.it0 >= B.nt[p]) return;
    const StencilArgs<T>& A = B.table[p];
    if (NDIM == 3)
        stencil3d_block<T, Off, D>(A, lb - B.block_start[p], reinterpret_cast<T*>(smem_raw));
    else
        stencil2d_block<T, Off, D>(A, lb - B.block_start[p], reinterpret_cast<T*>(smem_raw));
}

// ---------------------------------------------------------------------
// Sources and receivers of step `it0` (0-based), one block per shot:
// inject_sources! then record_receivers! (acoustic_2D_xPU.jl:77-91,126-128).
// Sources of a shot that share a node are added by one thread in source
// order (seq[shot] = 1), otherwise one thread per source.
// ---------------------------------------------------------------------
template <typename T>
struct IoArgs {
    T* pnew;                 // this step's pnew (shot 0)
    int64_t field_stride;
    const int64_t* src_lin;  // element offsets of the sources, all shots
    const T* src_wav;        // nt values per source
    const int* src_start;    // nshots+1 CSR offsets
    const int* src_seq;      // 1: the shot has sources sharing a node
    const int64_t* rec_lin;
    T* traces;               // nt values per receiver
    const int* rec_start;
    int64_t nt;
    int inject;              // 0: receivers only (the sources are already in pnew)
};

template <typename T>
__device__ __forceinline__ void io_block(const IoArgs<T>& A, int shot, int64_t it0) {
    T* p = A.pnew + (int64_t)shot * A.field_stride;
    const int s0 = A.src_start[shot], s1 = A.inject ? A.src_start[shot + 1] : s0;
    if (A.src_seq[shot]) {
        if (threadIdx.x == 0)
            for (int e = s0; e < s1; ++e) p[A.src_lin[e]] += A.src_wav[(int64_t)e * A.nt + it0];
    } else {
        for (int e = s0 + threadIdx.x; e < s1; e += blockDim.x)
            p[A.src_lin[e]] += A.src_wav[(int64_t)e * A.nt + it0];
    }
    __syncthreads();
    const int r0 = A.rec_start[shot], r1 = A.rec_start[shot + 1];
    for (int r = r0 + threadIdx.x; r < r1; r += blockDim.x)
        A.traces[(int64_t)r * A.nt + it0] = p[A.rec_lin[r]];
}

template <typename T>
__global__ __launch_bounds__(256) void sources_receivers(const IoArgs<T> A, int64_t it0) {
    io_block(A, blockIdx.x, it0);
}

// Batch: block b serves shot b - shot_start[p] of plan p.
template <typename T>
struct IoBatchArgs {
    const IoArgs<T>* table;   // per-plan arguments for this step's rotation
    int64_t it0;
    int shot_start[kMaxBatch + 1];
    int nplans;
};

template <typename T>
__global__ __launch_bounds__(256) void sources_receivers_batch(const IoBatchArgs<T> B) {
    const int b = blockIdx.x;
    const int p = batch_plan(B.shot_start, B.nplans, b);
    const IoArgs<T>& A = B.table[p];
    if (B.it0 >= A.nt) return;
    io_block(A, b - B.shot_start[p], B.it0);
}

}  // namespace awcpml
