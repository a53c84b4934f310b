// stencil2_kernels.h -- two time steps per launch (temporal blocking) of the
// 3D CPML pressure stencil.
//
// A launch takes (p^{t-1}, p^t) to (p^{t+1}, p^{t+2}).  Block (64, W): lane
// -> x, wave -> y row, every wave marching z through one chunk of planes.
// The block owns the nodes of its inner 62 lanes and W-2 rows; the outer
// lanes and rows are a one-node halo on which only step t+1 is evaluated
// (the neighbouring blocks evaluate it too), so that step t+2 finds all its
// operands inside the block.  At march position k a wave
//   1. evaluates step t+1 at plane k on all its lanes,
//   2. puts that p^{t+1} plane row into a three-plane LDS ring (and stores
//      it to HBM on the owned nodes: the next launch's pold, and the
//      receivers sample it),
//   3. after the block barrier evaluates step t+2 at plane k-1 on the
//      owned nodes: x neighbours of p^{t+1} by DPP lane shifts, y neighbours
//      from the ring, z neighbours from the wave's own registers.
// The chunk's march starts one plane early and ends one plane late.
//
// CPML: every slab node evaluates both psi edges of each of its axes in
// step t+1 (as the single-step kernel's per-node form does) and keeps them
// and its xi in registers for step t+2; only step t+2's psi and xi are
// stored, into buffers distinct from the inputs, since other blocks read
// the inputs on their halo nodes during the launch.  The arithmetic of each
// step is the single-step kernel's, operation for operation, so two steps
// of this kernel equal two single-step launches bit for bit.
//
// Aligned tiles.  Field rows start on 128-byte boundaries, but the 62 owned
// lanes of a wave do not: every row a wave stores then begins and ends in
// cache lines that a neighbouring tile's block completes at some other
// time, and such partial-line writes cost HBM more than twice what whole
// lines cost.  Where no CPML slab reaches a block's rows, its chunk's
// planes or the x nodes next to its tile, the block instead owns all 64
// lanes of a tile starting at x = 64*tx (lanes outside the interior store
// 0, the Dirichlet value, so every stored row is whole lines).  Step t+2 on
// lanes 0 and 63 then needs p^{t+1} at x = 64*tx-1 and 64*tx+64: the two
// halo-row waves, which have no step t+2 of their own, evaluate step t+1 at
// those nodes for every row of the block (one row per lane) and pass the
// values through a small LDS ring.  Ownership is decided per row group and
// chunk, so the two tilings partition the grid between them.
//
// The host uses this kernel only when every field buffer is zero on the
// Dirichlet nodes (no source on such a node since the last reset; the halo
// lanes and rows read those nodes as 0) and the fields take 32-bit byte
// offsets.  The optional d(pnew)/d(fact) outputs of both steps (the
// adjoint's replay) are the single-step kernel's q, operation for
// operation.
#pragma once

#include <utility>

#include "stencil_kernels.h"

namespace awcpml {

constexpr int kOwnX = kTileX - 2;  // owned x nodes per wave
constexpr int kMaxSrc2 = 32;       // sources per shot a two-step launch takes
constexpr int kMaxChunks2 = 32;    // march chunks a two-step launch lists

template <typename T>
struct Stencil2Args {
    // geometry, coefficients, pold = p^{t-1}, pcur = p^t; side[][].psi_rd
    // and .xi are step t+1's inputs, side[][].psi_wr receives step t+2's psi
    StencilArgs<T> s;
    T* p1;                   // p^{t+1}
    T* p2;                   // p^{t+2}
    // optional: d(pnew)/d(fact) of steps t+1 and t+2 (as StencilArgs::qout,
    // for the adjoint's replay), null = off
    T* q1;
    T* q2;
    T* xi_out[3][2];         // xi after step t+2
    const int64_t* src_lin;  // element offsets of the sources (all shots)
    const T* src_wav;        // nt values per source
    const int* src_start;    // nshots+1 CSR offsets
    int64_t nt;
    int64_t it0;             // 0-based index of step t+1
    // 3D: 1 = blocks without CPML in their rows and planes take aligned
    // 64-lane tiles (the host has checked that no tile's outside x
    // neighbours lie in an x slab); tiles_x then counts the larger of the
    // two tilings' tiles
    int align;
    int tiles62, tiles64;    // x tiles of the 62-lane and the aligned tiling
    // fact at the aligned tiles' edge nodes, rows of an edge adjacent: 3D
    // [z][tile][side][y], 2D [y][tile][side] (the edge lanes' fact loads then
    // share cache lines across the rows of a block)
    const T* fact_edge;
    // march chunks: chunk c owns planes (2D: rows) [chunk_start[c],
    // chunk_start[c+1]); 0 chunks listed: chunk c starts at 1 + c*march_len.
    // The host cuts the CPML layers of the march axis into chunks of their
    // own, so the interior chunks run the march without the slab terms.
    int nchunk_list;
    int chunk_start[kMaxChunks2 + 1];
};

// The planes (2D: rows) chunk c of a launch owns: [*kbeg, *kend).
template <typename T>
__device__ __forceinline__ void two_chunk(const Stencil2Args<T>& B, int chunk, int64_t n_march,
                                          int* kbeg, int* kend) {
    if (B.nchunk_list > 0) {
        *kbeg = B.chunk_start[chunk];
        *kend = B.chunk_start[chunk + 1];
    } else {
        *kbeg = 1 + chunk * B.s.march_len;
        *kend = (int)min((int64_t)*kbeg + B.s.march_len, n_march - 1);
    }
}

// Lane shifts across the 64-lane wavefront with DPP (VALU moves, no LDS
// traffic): from_below gives lane i the value of lane i-1, from_above the
// value of lane i+1; lanes without a neighbour keep their own value.
__device__ __forceinline__ int dpp_shift(int v, bool below) {
    return below ? __builtin_amdgcn_update_dpp(v, v, 0x138 /* wave_shr:1 */, 0xf, 0xf, false)
                 : __builtin_amdgcn_update_dpp(v, v, 0x130 /* wave_shl:1 */, 0xf, 0xf, false);
}
__device__ __forceinline__ float lane_shift(float v, bool below) {
    return __builtin_bit_cast(float, dpp_shift(__builtin_bit_cast(int, v), below));
}
__device__ __forceinline__ double lane_shift(double v, bool below) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const int lo = dpp_shift((int)(unsigned)(u & 0xffffffffull), below);
    const int hi = dpp_shift((int)(unsigned)(u >> 32), below);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// A copy of v in a register of its own (early-clobber output).  A value
// read from a ring slot that must outlive the slot's refill (pp, which
// becomes the next plane's pc, and fact, which step t+2 reads) is copied
// out before the refill.  The refill can then load straight into the
// slot's registers.  Otherwise the register allocator places the loads
// elsewhere and copies them into the slot at the loop latch, and those
// copies wait for every load in flight.
__device__ __forceinline__ float own_reg(float v) {
    float r;
    asm("v_mov_b32 %0, %1" : "=&v"(r) : "v"(v));
    return r;
}
__device__ __forceinline__ double own_reg(double v) {
    double r;
    asm("v_mov_b64 %0, %1" : "=&v"(r) : "v"(v));
    return r;
}

// Step t+1 results of one node that step t+2 at the same node reads.
template <typename T>
struct Mid {
    T p1;           // p^{t+1}
    T q;            // d(p^{t+1})/d(fact): laplacian + sum of (dpsi + xi)
    T pc, f;        // p^t (step t+2's pold) and fact
    T xl, xh, xx;   // x slab: psi^{t+1} of edges m-1 and m, xi^{t+1}[m-1]
    T yl, yh, yx;   // y slab
    T zl, zh, zx;   // z slab
};

// One axis's CPML terms at a slab node: the node's two psi edges from the
// previous psi values (psi_old = psi[m], psi_old1 = psi[m-1]), xi[m-1]
// from xi_old, and t = dpsi + xi (acoustic_3D_xPU.jl update_psi_*! and the
// damping term of update_p!).  Same operations as cpml_apply.  Returned by
// value (out-parameters of the side-dispatched callers would be merged
// into stores through selected addresses and kept in scratch memory).
template <typename T>
struct Pair {
    T t, lo, hi, xn;
};

template <typename T>
__device__ __forceinline__ Pair<T> cpml_pair(const LayerCoef<T>& c, T psi_old, T psi_old1,
                                             T xi_old, T pm, T pc, T pp, T d2, T inv_d) {
    Pair<T> r;
    r.hi = c.bKh * psi_old + c.ah * (pp - pc) * inv_d;
    r.lo = c.bKh1 * psi_old1 + c.ah1 * (pc - pm) * inv_d;
    const T dpsi = (r.hi - r.lo) * inv_d;
    r.xn = c.bK1 * xi_old + c.a1 * (d2 + dpsi);
    r.t = dpsi + r.xn;
    return r;
}

// Step t+1 at one node from its loads.  ZS: z slab side of the plane (-1:
// none), a template constant so that no per-side array is indexed at run
// time (that would place the array in scratch memory).
template <typename T, typename Off, int SA, bool SB, int ZS>
__device__ __forceinline__ Mid<T> two_first(const March<T, Off, 3>& M, int mz,
                                            const NodeLoads<T>& q, T pm, T pc, T xm, T xp, T ym,
                                            T yp) {
    Mid<T> r = {};
    const T two = T(2);
    const T f = q.f;
    const T d2a = (xp - two * pc + xm) * M.i2a;
    const T d2b = (yp - two * pc + ym) * M.i2b;
    const T d2m = (q.pp - two * pc + pm) * M.i2m;
    T damp = T(0);
    T slab = T(0);  // sum of (dpsi + xi), for the q outputs
    if (SA == 2 && M.sa.in) {
        const Pair<T> e = cpml_pair(M.sa.c, q.psa0, q.psa1, q.xia, xm, pc, xp, d2a, M.ia);
        damp += f * e.t;
        slab += e.t;
        r.xl = e.lo;
        r.xh = e.hi;
        r.xx = e.xn;
    }
    if (SB) {
        const Pair<T> e = cpml_pair(M.sb.c, q.psb0, q.psb1, q.xib, ym, pc, yp, d2b, M.ib);
        damp += f * e.t;
        slab += e.t;
        r.yl = e.lo;
        r.yh = e.hi;
        r.yx = e.xn;
    }
    if (ZS >= 0) {
        const LayerCoef<T> cz = lds_coef(M.lds, M.h1, ZS, mz);
        const Pair<T> e = cpml_pair(cz, q.psm0, q.psm1, q.xim, pm, pc, q.pp, d2m, M.imm);
        damp += f * e.t;
        slab += e.t;
        r.zl = e.lo;
        r.zh = e.hi;
        r.zx = e.xn;
    }
    const T lap = d2a + d2b + d2m;  // (d2x + d2y) + d2z
    r.p1 = two * pc - q.po + f * lap + damp;
    r.q = lap + slab;
    r.pc = pc;
    r.f = own_reg(f);
    return r;
}

// Output pointers of one launch for one shot.
template <typename T, typename Off>
struct TwoOut {
    T* p1;
    T* p2;
    T* q1;        // optional d(pnew)/d(fact) outputs of the two steps
    T* q2;
    T* xi_x;      // this column's x slab side (valid when M.sa.in)
    T* xi_y;      // this row's y slab side (valid when M.sb.in)
    T* xi_z[2];   // z slab sides
    // aligned tiles: this lane stores p^{t+1} and p^{t+2} (an owned row, x
    // inside the row pitch) at byte offset scol of plane 0 (its own x, not
    // the clamped one of its loads), and its node is interior (else the
    // stored value is 0)
    bool st, in1;
    Off scol;
};

// Step t+2 at one node (plane k) from its step t+1 results d and the
// neighbouring p^{t+1} values; stores on owned nodes only.
template <typename T, typename Off, int SA, bool SB, int ZS, bool AL>
__device__ __forceinline__ void two_second(const March<T, Off, 3>& M, const TwoOut<T, Off>& O,
                                           const Mid<T>& d, T pm, T pp, T xm, T xp, T ym, T yp,
                                           bool own, int mz, int k) {
    const T two = T(2);
    const T pc = d.p1, f = d.f;
    const T d2a = (xp - two * pc + xm) * M.i2a;
    const T d2b = (yp - two * pc + ym) * M.i2b;
    const T d2m = (pp - two * pc + pm) * M.i2m;
    T damp = T(0);
    T slab = T(0);  // sum of (dpsi + xi), for the q outputs
    if (SA == 2 && M.sa.in) {
        const Pair<T> e = cpml_pair(M.sa.c, d.xh, d.xl, d.xx, xm, pc, xp, d2a, M.ia);
        damp += f * e.t;
        slab += e.t;
        if (own) {
            stg(M.sa.p.psi_wr, M.sa.o_psi + (Off)k * M.sa.st_psi, e.hi);
            if (M.sa.first) stg(M.sa.p.psi_wr, M.sa.o_psi1 + (Off)k * M.sa.st_psi, e.lo);
            stg(O.xi_x, M.sa.o_xi + (Off)k * M.sa.st_xi, e.xn);
        }
    }
    if (SB) {
        const Pair<T> e = cpml_pair(M.sb.c, d.yh, d.yl, d.yx, ym, pc, yp, d2b, M.ib);
        damp += f * e.t;
        slab += e.t;
        if (own) {
            stg(M.sb.p.psi_wr, M.sb.o_psi + (Off)k * M.sb.st_psi, e.hi);
            if (M.sb.first) stg(M.sb.p.psi_wr, M.sb.o_psi1 + (Off)k * M.sb.st_psi, e.lo);
            stg(O.xi_y, M.sb.o_xi + (Off)k * M.sb.st_xi, e.xn);
        }
    }
    if (ZS >= 0) {
        const LayerCoef<T> cz = lds_coef(M.lds, M.h1, ZS, mz);
        const Pair<T> e = cpml_pair(cz, d.zh, d.zl, d.zx, pm, pc, pp, d2m, M.imm);
        damp += f * e.t;
        slab += e.t;
        if (own) {
            T* pw = M.pm_side[ZS < 0 ? 0 : ZS].psi_wr;
            stg(pw, M.col + (Off)mz * M.sm, e.hi);
            if (mz == M.mfirst[ZS < 0 ? 0 : ZS]) stg(pw, M.col + (Off)(mz - 1) * M.sm, e.lo);
            stg(O.xi_z[ZS < 0 ? 0 : ZS], M.col + (Off)(mz - 1) * M.sm, e.xn);
        }
    }
    const T lap = d2a + d2b + d2m;
    if (AL) {
        // whole rows: lanes outside the interior store the Dirichlet 0
        const T v = two * pc - d.pc + f * lap + damp;
        if (O.st) stg_stream(O.p2, O.scol + (Off)k * M.sm, O.in1 ? v : T(0));
        if (O.st && O.q2) stg(O.q2, O.scol + (Off)k * M.sm, O.in1 ? lap + slab : T(0));
    } else {
        // 62-node tiles: plain stores (see the p^{t+1} store)
        if (own) stg(O.p2, M.col + (Off)k * M.sm, two * pc - d.pc + f * lap + damp);
        if (own && O.q2) stg(O.q2, M.col + (Off)k * M.sm, lap + slab);
    }
}

// Step t+1 loads of march position k into a ring slot.  The wave's own
// pcur column is loaded once per plane (pp = pcur at k+1; pm and pc are the
// two planes before, carried in registers); the x neighbours are lane
// shifts of pc, so only the two edge lanes load one (q.xm: pcur at x-1 on
// lane 0, at x+1 on lane 63); the y neighbours come from the other waves'
// columns through LDS, so only the block's outer rows need one (q.ym: pcur
// at y-1 on row 0, at y+1 on row W-1).  pp is loaded for every position
// below the last plane (the next position needs it as pc); the other loads
// only for computed planes (1 <= k <= nz-2).  The slot is returned by
// value: branches that fill different fields of a slot held in memory get
// merged into stores through a selected address, which would keep the
// whole ring out of registers.
template <typename T, typename Off, int W, int SA, bool SB, bool GEN>
__device__ __forceinline__ NodeLoads<T> two_load_node(const March<T, Off, 3>& M, Off dxm, Off dym,
                                                      int nz, int k, int mside) {
    NodeLoads<T> q = {};
    const Off c = M.col + (Off)k * M.sm;
    if (!GEN || k + 1 <= nz - 1) q.pp = ldg(M.pcur, c + M.sm);
    if (GEN && (k < 1 || k > nz - 2)) return q;
    // issued by every lane and wave (the others reload their own node, a
    // cache hit), so that each march position issues the same loads and the
    // waits the compiler places for the ring slots stay exact; dxm / dym:
    // the lane's (wave's) offset to its x (y) neighbour outside the block,
    // 0 inside
    q.xm = ldg(M.pcur, c + dxm);
    q.ym = ldg(M.pcur, c + dym);
    q.po = ldg_stream(M.pold, c);
    q.f = ldg(M.fact, c);
    if (SA == 2) {
        q.psa0 = ldg(M.sa.p.psi_rd, M.sa.o_psi + (Off)k * M.sa.st_psi);
        q.psa1 = ldg(M.sa.p.psi_rd, M.sa.o_psi1 + (Off)k * M.sa.st_psi);
        q.xia = ldg((const T*)M.sa.p.xi, M.sa.o_xi + (Off)k * M.sa.st_xi);
    }
    if (SB) {
        q.psb0 = ldg(M.sb.p.psi_rd, M.sb.o_psi + (Off)k * M.sb.st_psi);
        q.psb1 = ldg(M.sb.p.psi_rd, M.sb.o_psi1 + (Off)k * M.sb.st_psi);
        q.xib = ldg((const T*)M.sb.p.xi, M.sb.o_xi + (Off)k * M.sb.st_xi);
    }
    if (mside >= 0) {  // z slab layer (mside: its side)
        const int mz = k - (mside ? M.mshift[1] : M.mshift[0]);
        const T* pr = mside ? M.pm_side[1].psi_rd : M.pm_side[0].psi_rd;
        const T* xr = mside ? M.pm_side[1].xi : M.pm_side[0].xi;
        q.psm0 = ldg(pr, M.col + (Off)mz * M.sm);
        q.psm1 = ldg(pr, M.col + (Off)(mz - 1) * M.sm);
        q.xim = ldg(xr, M.col + (Off)(mz - 1) * M.sm);
    }
    return q;
}

// Uniform per-launch values of a wave's march.
struct TwoGeom {
    int wy, kbeg, kend, nz, ksrc_lo, ksrc_hi, ns;
    int halo, act[2];  // march axis: CPML layers, active sides (no argument reads in the march)
    int64_t col_el, plane;
    bool in1, own;
    bool inx;     // the lane's x is its clamped x (pcur there is the lane's own column)
    bool iny;     // the wave's y is its clamped y
    uint32_t dxm; // 3D: byte offset from the lane's node to the x neighbour it loads
    uint32_t dym; // 3D: byte offset from the wave's row to the y neighbour it loads
    // aligned tiles: the nodes just outside the tile, x = 64*tx-1 (side 0)
    // and 64*tx+64 (side 1).  Lanes 0 and 63 of every row load what step
    // t+1 at their side's node needs beyond their own loads (its outer x
    // neighbour, pold, fact) and pass it through LDS; lane r < W of the
    // halo-row waves (wave 0: side 0, wave W-1: side 1) evaluates that node
    // in the block's row r.
    bool eln;         // lane 0 or 63 of an aligned tile
    uint32_t dxe;     // byte offset from the lane's column to its side's node (clamped)
    uint32_t dxo;     // ... and to that node's outer x neighbour
    int64_t hcol_el;  // halo-row waves: element offset of lane r's node column (sources)
    bool hin;         // ... that node is interior
    const void* fedge;  // the plan's edge copy of fact
    uint32_t fe_off;    // byte offset of this edge lane's entry at march position 0
    uint32_t fe_step;   // bytes per march position
};

// Step t+1 inputs of the node just outside an aligned tile that its row's
// edge lane loads (see TwoGeom): the outer x neighbour of pcur, pold and
// fact.  The node's pcur is the lane's x-neighbour load.
template <typename T>
struct EdgeSlot {
    T xo, pe, fe;
};

template <typename T, typename Off, bool GEN>
__device__ __forceinline__ EdgeSlot<T> edge_slot_load(const TwoGeom& G, const March<T, Off, 3>& M,
                                                      int k) {
    EdgeSlot<T> q = {};
    if (GEN && (k < 1 || k > G.nz - 2)) return q;
    if (G.eln) {
        const Off c = M.col + (Off)k * M.sm;
        q.xo = ldg(M.pcur, c + (Off)G.dxo);
        q.pe = ldg(M.pold, c + (Off)G.dxe);
        q.fe = ldg((const T*)G.fedge, (Off)G.fe_off + (Off)k * (Off)G.fe_step);
    }
    return q;
}

// 2D: the same loads for row k.
template <typename T, typename Off, bool GEN>
__device__ __forceinline__ EdgeSlot<T> edge_slot_load2(const TwoGeom& G, const March<T, Off, 2>& M,
                                                       int k) {
    EdgeSlot<T> q = {};
    if (GEN && (k < 1 || k > G.nz - 2)) return q;
    if (G.eln) {
        const Off c = M.col + (Off)k * M.sm;
        q.xo = ldg(M.pcur, c + (Off)G.dxo);
        q.pe = ldg(M.pold, c + (Off)G.dxe);
        q.fe = ldg((const T*)G.fedge, (Off)G.fe_off + (Off)k * (Off)G.fe_step);
    }
    return q;
}

// LDS of the aligned tiles' outside nodes: their pcur per row over four
// planes (slot: plane & 3), and the other step t+1 inputs of plane k
// (slot: k & 1) -- outer x neighbour, pold, fact, inner x neighbour.
template <typename T, int W>
struct EdgeLds {
    T pc[4][2][W];
    T in[2][4][2][W];
};

// ZM = false: no z slab layer in the chunk.  One code path for both sides
// (a branch per side would be merged into stores through selected
// addresses, which keeps the slot in scratch memory).
template <typename T, typename Off, int W, int SA, bool SB, bool ZM, bool GEN>
__device__ __forceinline__ NodeLoads<T> two_load(const TwoGeom& G, const March<T, Off, 3>& M,
                                                 int k, int wy) {
    int mz;
    const int zs = ZM ? classify(k, G.nz, G.halo, G.act, &mz) : -1;
    return two_load_node<T, Off, W, SA, SB, GEN>(M, (Off)G.dxm, (Off)G.dym, G.nz, k, zs);
}

template <typename T, typename Off, int SA, bool SB, bool ZM>
__device__ __forceinline__ Mid<T> two_first_at(const TwoGeom& G, const March<T, Off, 3>& M,
                                               int k, const NodeLoads<T>& q, T pm, T pc, T xm,
                                               T xp, T ym, T yp) {
    if (!ZM) return two_first<T, Off, SA, SB, -1>(M, 0, q, pm, pc, xm, xp, ym, yp);
    int mz;
    const int zs = classify(k, G.nz, G.halo, G.act, &mz);
    if (zs == 0) return two_first<T, Off, SA, SB, 0>(M, mz, q, pm, pc, xm, xp, ym, yp);
    if (zs == 1) return two_first<T, Off, SA, SB, 1>(M, mz, q, pm, pc, xm, xp, ym, yp);
    return two_first<T, Off, SA, SB, -1>(M, mz, q, pm, pc, xm, xp, ym, yp);
}

template <typename T, typename Off, int SA, bool SB, bool ZM, bool AL>
__device__ __forceinline__ void two_second_at(const TwoGeom& G, const March<T, Off, 3>& M,
                                              const TwoOut<T, Off>& O, const Mid<T>& d, T pm,
                                              T pp, T xm, T xp, T ym, T yp, bool own, int k) {
    if (!ZM) {
        two_second<T, Off, SA, SB, -1, AL>(M, O, d, pm, pp, xm, xp, ym, yp, own, 0, k);
        return;
    }
    int mz;
    const int zs = classify(k, G.nz, G.halo, G.act, &mz);
    if (zs == 0) two_second<T, Off, SA, SB, 0, AL>(M, O, d, pm, pp, xm, xp, ym, yp, own, mz, k);
    else if (zs == 1) two_second<T, Off, SA, SB, 1, AL>(M, O, d, pm, pp, xm, xp, ym, yp, own, mz, k);
    else two_second<T, Off, SA, SB, -1, AL>(M, O, d, pm, pp, xm, xp, ym, yp, own, mz, k);
}

// Per-wave march state: the ring of step t+1 loads (PF planes in flight),
// step t+1's results of the previous plane, p^{t+1} two planes back, and
// the LDS ring slot of the current plane.
template <typename T, int PF>
struct TwoCtx {
    NodeLoads<T> qr[PF];
    Mid<T> d;
    T p1m;
    T pz0, pz1;   // pcur of the wave's column at planes k-1 and k
    int slot;
    // 3D: element offsets of the LDS ring slots of planes k-1, k, k+1
    int so_p, so_c, so_n;
    // aligned tiles, edge lanes: their side node's loads, PF planes in flight
    EdgeSlot<T> es[PF];
    // 2D aligned tiles, edge lanes: their side node's pcur at row k-1 and its
    // step t+1 result at row k-1
    T ez0, ep1;
    // element offsets of the edge ring's slots of planes k-1, k, k+1
    int eo_p, eo_c, eo_n;
};

// The shot's sources for step t+1, staged in LDS at the start of a launch
// (their element offsets and wavelet values), so that the march reads them
// with LDS loads.
template <typename T>
struct TwoSrc {
    int64_t lin[kMaxSrc2];
    T w[kMaxSrc2];
};

// Copy the shot's first kMaxSrc2 sources into S (the host launches two-step
// kernels only for shots with at most that many); returns their count.
template <typename T>
__device__ __forceinline__ int two_stage_sources(const Stencil2Args<T>& B, int64_t it0, int shot,
                                                 TwoSrc<T>& S) {
    const int s0 = B.src_start[shot];
    int ns = B.src_start[shot + 1] - s0;
    ns = ns < kMaxSrc2 ? ns : kMaxSrc2;
    const int tid = (int)(threadIdx.y * blockDim.x + threadIdx.x);
    if (tid < ns) {
        S.lin[tid] = B.src_lin[s0 + tid];
        S.w[tid] = B.src_wav[(int64_t)(s0 + tid) * B.nt + it0];
    }
    return ns;
}

// One march position k: step t+1 at plane k, refill of ring slot U with
// plane k+PF, the LDS exchange, step t+2 at plane k-1.  GEN = false: k is
// in the steady state, where every plane index condition holds (computed,
// refilled, owned plane, step t+2 due), so none is tested.
template <typename T, typename Off, int W, int PF, int SA, bool SB, bool ZM, bool AL, bool GEN,
          int U>
__device__ __forceinline__ void two_iter(const Stencil2Args<T>& B, const March<T, Off, 3>& M,
                                         const TwoOut<T, Off>& O, T (*ring)[W + 2][kTileX],
                                         T (*pring)[W + 2][kTileX], T (*eg)[2][W],
                                         EdgeLds<T, W>* EL, const TwoGeom& G, const TwoSrc<T>& S,
                                         TwoCtx<T, PF>& C, int k) {
    const int lane = (int)threadIdx.x;
    // this lane's column in the padded LDS rings, one row above its own:
    // slot c's rows wy-1, wy, wy+1 are at [so], [so + 64], [so + 128]
    T* const pb = &pring[0][0][0] + G.wy * kTileX + lane;
    T* const rb = &ring[0][0][0] + G.wy * kTileX + lane;
    Mid<T> dn = {};
    const T pp = own_reg(C.qr[U].pp);
    if (!GEN || (k >= 1 && k <= G.nz - 2)) {
        // x neighbours: lane shifts of pc; a lane outside the grid's x range
        // (clamped) stands for the Dirichlet node there, 0
        const T pc = C.pz1;
        const T pcx = G.inx ? pc : T(0);
        const T sl = lane_shift(pcx, true), sr = lane_shift(pcx, false);
        const T edge = C.qr[U].xm;
        const T xm = lane == 0 ? edge : sl;
        const T xp = lane == kTileX - 1 ? edge : sr;
        // rows wy-1 and wy+1 of the padded ring (rows 0 and W+1 are padding,
        // read and discarded by the block's outer rows)
        const T ymr = pb[C.so_c];
        const T ypr = pb[C.so_c + 2 * kTileX];
        const T ym = G.wy == 0 ? C.qr[U].ym : ymr;
        const T yp = G.wy == W - 1 ? C.qr[U].ym : ypr;
        dn = two_first_at<T, Off, SA, SB, ZM>(G, M, k, C.qr[U], C.pz0, pc, xm, xp, ym, yp);
        // inject_sources! of step t+1, in source order
        if (k >= G.ksrc_lo && k <= G.ksrc_hi) {
            const int64_t lin = G.col_el + (int64_t)k * G.plane;
            for (int e = 0; e < G.ns; ++e)
                if (S.lin[e] == lin) dn.p1 += S.w[e];
        }
        dn.p1 = G.in1 ? dn.p1 : T(0);
        if (AL && G.eln) {
            // the side node's step t+1 inputs of plane k for the halo-row
            // waves: outer x neighbour, pold, fact, and this lane's pcur
            // (the node's inner x neighbour)
            const int sd = lane == 0 ? 0 : 1;
            EL->in[k & 1][0][sd][G.wy] = C.es[U].xo;
            EL->in[k & 1][1][sd][G.wy] = C.es[U].pe;
            EL->in[k & 1][2][sd][G.wy] = C.es[U].fe;
            EL->in[k & 1][3][sd][G.wy] = pc;
        }
    }
    if (AL && G.eln) {
        // the side node's pcur at plane k+1 (this lane's x-neighbour load of
        // the next ring slot; 0 on a Dirichlet row)
        const T pn = C.qr[(U + 1) % PF].xm;
        EL->pc[(k + 1) & 3][lane == 0 ? 0 : 1][G.wy] = G.iny ? pn : T(0);
    }
    C.pz0 = C.pz1;
    C.pz1 = pp;
    const int kr = k + PF;
    if (!GEN || kr <= G.kend) {
        C.qr[U] = two_load<T, Off, W, SA, SB, ZM, GEN>(G, M, kr, G.wy);
        if (AL) C.es[U] = edge_slot_load<T, Off, GEN>(G, M, kr);
    }
    __builtin_amdgcn_sched_barrier(0);
    // this column's pcur at plane k+1 for the neighbouring rows (0 for a
    // Dirichlet row), and p^{t+1} at plane k
    pb[C.so_n + kTileX] = G.iny ? pp : T(0);
    rb[C.so_c + kTileX] = dn.p1;
    if (AL) {
        // whole rows (dn.p1 is 0 outside the interior)
        if (O.st && (!GEN || (k >= G.kbeg && k < G.kend))) {
            stg_stream(O.p1, O.scol + (Off)k * M.sm, dn.p1);
            if (O.q1) stg(O.q1, O.scol + (Off)k * M.sm, G.in1 ? dn.q : T(0));
        }
    } else if (G.own && (!GEN || (k >= G.kbeg && k < G.kend))) {
        // 62-node tiles: rows end in cache lines the neighbouring tile's
        // block completes; plain stores let L2 merge the two halves
        stg(O.p1, M.col + (Off)k * M.sm, dn.p1);
        if (O.q1) stg(O.q1, M.col + (Off)k * M.sm, dn.q);
    }
    __syncthreads();
    // step t+2 on the owned rows only: the first and last wave are the
    // block's halo rows, which store nothing (a wave-uniform branch that
    // leaves their issue slots to the other waves)
    if ((!GEN || k - 1 >= G.kbeg) && G.wy >= 1 && G.wy <= W - 2) {
        T xm = lane_shift(C.d.p1, true);
        T xp = lane_shift(C.d.p1, false);
        if (AL) {
            // aligned tiles: lanes 0 and 63 take the nodes outside the tile
            // from the edge ring (plane k-1)
            const T* ep = &eg[0][0][0] + C.eo_p;
            const T el = ep[G.wy], er = ep[W + G.wy];
            xm = lane == 0 ? el : xm;
            xp = lane == kTileX - 1 ? er : xp;
        }
        const T ym = rb[C.so_p];                 // plane k-1, row wy-1
        const T yp = rb[C.so_p + 2 * kTileX];    // plane k-1, row wy+1
        two_second_at<T, Off, SA, SB, ZM, AL>(G, M, O, C.d, C.p1m, dn.p1, xm, xp, ym, yp, G.own,
                                              k - 1);
    }
    if (AL && (G.wy == 0 || G.wy == W - 1) && (!GEN || (k >= 1 && k <= G.nz - 2))) {
        // aligned tiles: the halo-row waves evaluate step t+1 at plane k on
        // their side's node outside the tile in every row of the block
        // (lane r: row r), the plain stencil in the owning block's order of
        // operations (no CPML slab reaches these nodes), from the inputs the
        // rows' edge lanes put in LDS
        const int sd = G.wy == 0 ? 0 : 1;
        const int r = lane < W ? lane : W - 1;
        const int rm = r > 0 ? r - 1 : 0, rp = r < W - 1 ? r + 1 : W - 1;
        const T* pk = EL->pc[k & 3][sd];
        const T hpc = pk[r], hym = pk[rm], hyp = pk[rp];
        const T hpm = EL->pc[(k - 1) & 3][sd][r];
        const T hpp = EL->pc[(k + 1) & 3][sd][r];
        const T xo = EL->in[k & 1][0][sd][r];
        const T inner = EL->in[k & 1][3][sd][r];
        NodeLoads<T> q = {};
        q.pp = hpp;
        q.po = EL->in[k & 1][1][sd][r];
        q.f = EL->in[k & 1][2][sd][r];
        Mid<T> e = two_first<T, Off, 0, false, -1>(M, 0, q, hpm, hpc, sd ? inner : xo,
                                                   sd ? xo : inner, hym, hyp);
        if (k >= G.ksrc_lo && k <= G.ksrc_hi) {
            const int64_t lin = G.hcol_el + (int64_t)k * G.plane;
            for (int s = 0; s < G.ns; ++s)
                if (S.lin[s] == lin) e.p1 += S.w[s];
        }
        if (lane < W) (&eg[0][0][0])[C.eo_c + sd * W + lane] = G.hin ? e.p1 : T(0);
    }
    C.p1m = C.d.p1;
    C.d = dn;
    const int so = C.so_p;
    C.so_p = C.so_c;
    C.so_c = C.so_n;
    C.so_n = so;
    if (AL) {
        const int eo = C.eo_p;
        C.eo_p = C.eo_c;
        C.eo_c = C.eo_n;
        C.eo_n = eo;
    }
}

// Positions k+u, u in Us..., in order, ring slot u % PF (a constant), each
// only while k+u <= kend when CHECK.
template <typename T, typename Off, int W, int PF, int SA, bool SB, bool ZM, bool AL, bool GEN,
          bool CHECK, int... Us>
__device__ __forceinline__ void two_group(const Stencil2Args<T>& B, const March<T, Off, 3>& M,
                                          const TwoOut<T, Off>& O, T (*ring)[W + 2][kTileX],
                                          T (*pring)[W + 2][kTileX], T (*eg)[2][W],
                                          EdgeLds<T, W>* EL, const TwoGeom& G, const TwoSrc<T>& S,
                                          TwoCtx<T, PF>& C, int k,
                                          std::integer_sequence<int, Us...>) {
    ((!CHECK || k + Us <= G.kend
          ? (two_iter<T, Off, W, PF, SA, SB, ZM, AL, GEN, Us % PF>(B, M, O, ring, pring, eg, EL, G,
                                                                   S, C, k + Us),
             0)
          : 0),
     ...);
}

// The march of one wave over planes kbeg-1 .. kend (step t+1) and
// kbeg .. kend-1 (step t+2): a head of H generic positions (H: 2 rounded
// up to a multiple of PF, so that the steady state starts at ring slot 0),
// the steady state unrolled PF times (constant slot indices), and a
// generic tail.  Every wave of the block runs the same positions (kbeg,
// kend are block-uniform) and so meets every barrier.
template <typename T, typename Off, int W, int PF, int SA, bool SB, bool ZM, bool AL>
__device__ __forceinline__ void two_march(const Stencil2Args<T>& B, const March<T, Off, 3>& M,
                                          const TwoOut<T, Off>& O, T (*ring)[W + 2][kTileX],
                                          T (*pring)[W + 2][kTileX], T (*eg)[2][W],
                                          EdgeLds<T, W>* EL, const TwoGeom& G,
                                          const TwoSrc<T>& S) {
    static_assert(!AL || PF >= 2, "aligned tiles read the next ring slot");
    TwoCtx<T, PF> C;
    C.pz0 = G.kbeg >= 2 ? ldg(M.pcur, M.col + (Off)(G.kbeg - 2) * M.sm) : T(0);
    C.pz1 = ldg(M.pcur, M.col + (Off)(G.kbeg - 1) * M.sm);
#pragma unroll
    for (int u = 0; u < PF; ++u) {
        const int kk = G.kbeg - 1 + u;
        C.qr[u] = kk <= G.kend ? two_load<T, Off, W, SA, SB, ZM, true>(G, M, kk, G.wy) : NodeLoads<T>{};
        __builtin_amdgcn_sched_barrier(0);
    }
    if (AL) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int kk = G.kbeg - 1 + u;
            C.es[u] = kk <= G.kend ? edge_slot_load<T, Off, true>(G, M, kk) : EdgeSlot<T>{};
        }
        // the side nodes' pcur at the first march position's plane
        if (G.eln) EL->pc[(G.kbeg - 1) & 3][threadIdx.x == 0 ? 0 : 1][G.wy] = G.iny ? C.qr[0].xm : T(0);
    }
    C.d = Mid<T>{};
    C.p1m = T(0);
    {
        constexpr int SS = (W + 2) * kTileX;  // elements per ring slot
        const int s0 = (G.kbeg - 1) % 3;
        C.so_c = s0 * SS;
        C.so_n = (s0 == 2 ? 0 : s0 + 1) * SS;
        C.so_p = (s0 == 0 ? 2 : s0 - 1) * SS;
        constexpr int ES = 2 * W;  // elements per edge ring slot
        C.eo_c = s0 * ES;
        C.eo_n = (s0 == 2 ? 0 : s0 + 1) * ES;
        C.eo_p = (s0 == 0 ? 2 : s0 - 1) * ES;
    }
    (&pring[0][0][0])[C.so_c + (G.wy + 1) * kTileX + threadIdx.x] = G.iny ? C.pz1 : T(0);
    __syncthreads();
    constexpr int H = (2 + PF - 1) / PF * PF;
    int k = G.kbeg - 1;
    two_group<T, Off, W, PF, SA, SB, ZM, AL, true, true>(B, M, O, ring, pring, eg, EL, G, S, C, k,
                                                         std::make_integer_sequence<int, H>{});
    k += H;
    const int last = (G.kend < G.nz - 2 ? G.kend : G.nz - 2) - PF;  // last steady position
    // drain the head's memory operations once: the steady loop is then
    // entered with nothing in flight, and the counter waits the compiler
    // derives at its head follow the loop's own issue order (merged with the
    // head's different order they fall back to waiting for everything)
    __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0)
    for (; k + PF - 1 <= last; k += PF)
        two_group<T, Off, W, PF, SA, SB, ZM, AL, false, false>(B, M, O, ring, pring, eg, EL, G, S, C,
                                                               k, std::make_integer_sequence<int, PF>{});
    for (; k <= G.kend; k += PF)
        two_group<T, Off, W, PF, SA, SB, ZM, AL, true, true>(B, M, O, ring, pring, eg, EL, G, S, C,
                                                             k, std::make_integer_sequence<int, PF>{});
}

// Logical block lb of a launch over one plan; it0 is the 0-based index of
// step t+1 (the wavelet sample its sources add).
template <typename T, typename Off, int W, int PF>
__device__ __forceinline__ void stencil3d_two_block(const Stencil2Args<T>& B, int64_t it0, int lb) {
    // three-plane rings of p^{t+1} and pcur rows, padded with one row on
    // each side: wave wy's row is row wy+1
    __shared__ T ring[3][W + 2][kTileX];
    __shared__ T pring[3][W + 2][kTileX];
    __shared__ T eg[3][2][W];  // aligned tiles: p^{t+1} just outside the tile, per row
    __shared__ EdgeLds<T, W> EL;
    __shared__ TwoSrc<T> S;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* lds = reinterpret_cast<T*>(smem_raw);
    const StencilArgs<T>& A = B.s;
    const int64_t nx = A.nx, ny = A.ny, nz = A.nz, halo = A.halo;

    // logical order: x tile fastest, then y tile, chunk, shot
    int rest = lb;
    const int tx = rest % A.tiles_x; rest /= A.tiles_x;
    const int ty = rest % A.tiles_y; rest /= A.tiles_y;
    const int chunk = rest % A.nchunks;
    const int shot = rest / A.nchunks;
    int kbeg, kend;
    two_chunk(B, chunk, nz, &kbeg, &kend);
    // any z slab layer among the planes this chunk reads (block-uniform)
    const bool zm = (A.active[2][0] && kbeg - 1 <= halo) || (A.active[2][1] && kend >= nz - halo - 1);
    // aligned tiles where no y slab reaches the block's rows and no z slab
    // its planes (decided per row group and chunk, so that all tiles of a
    // row group and chunk share one tiling)
    bool al = false;
    if (B.align) {
        const int64_t y0 = (int64_t)ty * (W - 2), y1 = y0 + W - 1;
        const int64_t r0 = y0 < 1 ? 1 : y0, r1 = y1 > ny - 2 ? ny - 2 : y1;
        const bool yslab = (A.active[1][0] && r0 <= halo) || (A.active[1][1] && r1 >= ny - halo - 1);
        al = !yslab && !zm;
    }
    if (PF < 2) al = false;  // (the aligned march reads the next ring slot)
    if (tx >= (al ? B.tiles64 : B.tiles62)) return;  // whole block: a tile of the other tiling
    const int ns = two_stage_sources(B, it0, shot, S);
    if (A.active[2][0] || A.active[2][1]) stage_march_coef(A, 2, lds);
    __syncthreads();
    const int wy = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int lane = (int)threadIdx.x;
    const int64_t i = al ? (int64_t)tx * kTileX + lane : (int64_t)tx * kOwnX + lane;
    const int64_t j = (int64_t)ty * (W - 2) + wy;
    // clamped to the interior so every load stays inside the field
    const int64_t ic = i < 1 ? 1 : (i > nx - 2 ? nx - 2 : i);
    const int64_t jc = j < 1 ? 1 : (j > ny - 2 ? ny - 2 : j);
    const bool in1 = i == ic && j == jc;
    const bool orow = wy >= 1 && wy <= W - 2;
    const bool own = in1 && orow && (al || (lane >= 1 && lane <= kOwnX));

    constexpr Off E = (Off)sizeof(T);
    const int64_t fo = (int64_t)shot * A.field_stride;
    March<T, Off, 3> M;
    M.pcur = A.pcur + fo;
    M.pold = A.pold + fo;
    M.pnew = nullptr;
    M.fact = A.fact;
    M.qout = nullptr;
    M.sy = (Off)A.px * E;
    M.sm = (Off)(A.px * ny) * E;
    M.col = ((Off)ic + (Off)A.px * (Off)jc) * E;
    M.i2a = A.inv_d2[0];
    M.i2b = A.inv_d2[1];
    M.i2m = A.inv_d2[2];
    M.ia = A.inv_d[0];
    M.ib = A.inv_d[1];
    M.imm = A.inv_d[2];
    M.lds = lds;
    M.h1 = (int)halo + 1;
    M.store_p = true;
    const Off hp1 = (Off)halo + 1, hl = (Off)halo;
    // psi_x: (halo+1) x ny x nz, xi_x: halo x ny x nz
    col_slab<T, Off>(A, 0, ic, nx, shot, hp1 * (Off)jc * E, E, hl * (Off)jc * E, E,
                     hp1 * (Off)ny * E, hl * (Off)ny * E, M.sa);
    // psi_y: px x (halo+1) x nz, xi_y: px x halo x nz
    col_slab<T, Off>(A, 1, jc, ny, shot, (Off)ic * E, M.sy, (Off)ic * E, M.sy, hp1 * M.sy,
                     hl * M.sy, M.sb);
    for (int s = 0; s < 2; ++s) M.pm_side[s] = slab_ptrs(A, 2, s, shot);
    M.mshift[0] = 0;
    M.mshift[1] = (int)(nz - halo - 2);
    M.mfirst[0] = 1;
    M.mfirst[1] = right_first(nz, halo);

    TwoOut<T, Off> O;
    O.p1 = B.p1 + fo;
    O.p2 = B.p2 + fo;
    O.q1 = B.q1 ? B.q1 + fo : nullptr;
    O.q2 = B.q2 ? B.q2 + fo : nullptr;
    int mx, my;
    const int sx = classify(ic, nx, halo, A.active[0], &mx);
    const int sy = classify(jc, ny, halo, A.active[1], &my);
    O.xi_x = (sx == 1 ? B.xi_out[0][1] : B.xi_out[0][0]) + shot * A.xi_stride[0];
    O.xi_y = (sy == 1 ? B.xi_out[1][1] : B.xi_out[1][0]) + shot * A.xi_stride[1];
    O.xi_z[0] = B.xi_out[2][0] + shot * A.xi_stride[2];
    O.xi_z[1] = B.xi_out[2][1] + shot * A.xi_stride[2];
    // aligned tiles store whole rows: every lane inside the row pitch
    O.st = al && orow && j == jc && i < A.px;
    O.in1 = in1;
    O.scol = O.st ? ((Off)i + (Off)A.px * (Off)jc) * E : M.col;

    // planes holding this shot's sources
    TwoGeom G;
    G.ksrc_lo = 1 << 30;
    G.ksrc_hi = -1;
    G.plane = A.px * ny;
    G.ns = ns;
    for (int e = 0; e < ns; ++e) {
        const int kk = (int)(S.lin[e] / G.plane);
        G.ksrc_lo = kk < G.ksrc_lo ? kk : G.ksrc_lo;
        G.ksrc_hi = kk > G.ksrc_hi ? kk : G.ksrc_hi;
    }
    G.col_el = ic + A.px * jc;
    G.wy = wy;
    G.kbeg = kbeg;
    G.kend = kend;
    G.nz = (int)nz;
    G.halo = (int)halo;
    G.act[0] = A.active[2][0];
    G.act[1] = A.active[2][1];
    G.in1 = in1;
    G.own = own;
    G.inx = i == ic;
    G.iny = j == jc;
    G.dxm = lane == 0 ? (uint32_t)(0u - (uint32_t)E) : (lane == kTileX - 1 ? (uint32_t)E : 0u);
    G.dym = wy == 0 ? (uint32_t)(0u - (uint32_t)M.sy) : (wy == W - 1 ? (uint32_t)M.sy : 0u);
    {
        // aligned tiles: lanes 0 and 63 load for their side's node outside
        // the tile (x = 64*tx-1, 64*tx+64; clamped into the interior, like
        // every load); lane r < W of a halo-row wave evaluates its side's
        // node in the block's row r (wave 0: left side, wave W-1: right)
        G.eln = al && (lane == 0 || lane == kTileX - 1);
        const int64_t xs = lane == 0 ? i - 1 : i + 1;
        const int64_t xsc = xs < 1 ? 1 : (xs > nx - 2 ? nx - 2 : xs);
        const int64_t xso = lane == 0 ? xsc - 1 : xsc + 1;
        G.dxe = (uint32_t)((Off)(xsc - ic) * E);
        G.dxo = (uint32_t)((Off)(xso - ic) * E);
        const int64_t xe = wy == 0 ? (int64_t)tx * kTileX - 1 : (int64_t)tx * kTileX + kTileX;
        const int64_t jr = (int64_t)ty * (W - 2) + (lane < W ? lane : W - 1);
        const int64_t xec = xe < 1 ? 1 : (xe > nx - 2 ? nx - 2 : xe);
        const int64_t jrc = jr < 1 ? 1 : (jr > ny - 2 ? ny - 2 : jr);
        G.hcol_el = xec + A.px * jrc;
        G.hin = xe == xec && jr == jrc;
        // this edge lane's entry of the edge copy of fact: [z][tile][side][y]
        G.fedge = B.fact_edge;
        G.fe_off = (uint32_t)((((Off)tx * 2 + (lane == 0 ? 0 : 1)) * (Off)ny + (Off)jc) * E);
        G.fe_step = (uint32_t)((Off)B.tiles64 * 2 * (Off)ny * E);
    }

    const bool xs = __any(M.sa.in);
    const bool ys = M.sb.in;  // wave-uniform
#define AWCPML_TWO(SA_, SB_, ZM_, AL_)                                                        \
    two_march<T, Off, W, PF, SA_, SB_, ZM_, AL_>(B, M, O, ring, pring, eg, &EL, G, S)
    if constexpr (PF >= 2) {
        if (al) {
            // no y or z slab in the block (its rows and planes were checked)
            if (xs) AWCPML_TWO(2, false, false, true);
            else AWCPML_TWO(0, false, false, true);
            return;
        }
    }
    if (zm) {
        if (xs) {
            if (ys) AWCPML_TWO(2, true, true, false);
            else AWCPML_TWO(2, false, true, false);
        } else {
            if (ys) AWCPML_TWO(0, true, true, false);
            else AWCPML_TWO(0, false, true, false);
        }
    } else {
        if (xs) {
            if (ys) AWCPML_TWO(2, true, false, false);
            else AWCPML_TWO(2, false, false, false);
        } else {
            if (ys) AWCPML_TWO(0, true, false, false);
            else AWCPML_TWO(0, false, false, false);
        }
    }
#undef AWCPML_TWO
}

// ---------------------------------------------------------------------
// 2D: two time steps per launch.  Block (64, BY): lane -> x (62 owned
// lanes, as in 3D), wave -> one chunk of rows along y, the march axis.
// Step t+2 at row k-1 needs p^{t+1} at rows k-2..k of the wave's own
// column (registers) and at x-1, x+1 (DPP lane shifts), so the waves of a
// block are independent: no LDS exchange and no barrier.
// ---------------------------------------------------------------------

// Step t+1 at one node (2D).  ZS: y slab side of the row (-1: none).
template <typename T, typename Off, int SA, int ZS>
__device__ __forceinline__ Mid<T> two2_first(const March<T, Off, 2>& M, int mz,
                                             const NodeLoads<T>& q, T pm, T pc, T xm, T xp) {
    Mid<T> r = {};
    const T two = T(2);
    const T f = q.f;
    const T d2a = (xp - two * pc + xm) * M.i2a;
    const T d2m = (q.pp - two * pc + pm) * M.i2m;
    T damp = T(0);
    T slab = T(0);  // sum of (dpsi + xi), for the q outputs
    if (SA == 2 && M.sa.in) {
        const Pair<T> e = cpml_pair(M.sa.c, q.psa0, q.psa1, q.xia, xm, pc, xp, d2a, M.ia);
        damp += f * e.t;
        slab += e.t;
        r.xl = e.lo;
        r.xh = e.hi;
        r.xx = e.xn;
    }
    if (ZS >= 0) {
        const LayerCoef<T> cz = lds_coef(M.lds, M.h1, ZS, mz);
        const Pair<T> e = cpml_pair(cz, q.psm0, q.psm1, q.xim, pm, pc, q.pp, d2m, M.imm);
        damp += f * e.t;
        slab += e.t;
        r.zl = e.lo;
        r.zh = e.hi;
        r.zx = e.xn;
    }
    const T lap = d2a + d2m;  // d2x + d2y
    r.p1 = two * pc - q.po + f * lap + damp;
    r.q = lap + slab;
    r.pc = pc;
    r.f = own_reg(f);
    return r;
}

// Step t+2 at one node (2D, row k); stores on owned nodes only.
template <typename T, typename Off, int SA, int ZS, bool AL>
__device__ __forceinline__ void two2_second(const March<T, Off, 2>& M, const TwoOut<T, Off>& O,
                                            const Mid<T>& d, T pm, T pp, T xm, T xp, bool own,
                                            int mz, int k) {
    const T two = T(2);
    const T pc = d.p1, f = d.f;
    const T d2a = (xp - two * pc + xm) * M.i2a;
    const T d2m = (pp - two * pc + pm) * M.i2m;
    T damp = T(0);
    T slab = T(0);  // sum of (dpsi + xi), for the q outputs
    if (SA == 2 && M.sa.in) {
        const Pair<T> e = cpml_pair(M.sa.c, d.xh, d.xl, d.xx, xm, pc, xp, d2a, M.ia);
        damp += f * e.t;
        slab += e.t;
        if (own) {
            stg(M.sa.p.psi_wr, M.sa.o_psi + (Off)k * M.sa.st_psi, e.hi);
            if (M.sa.first) stg(M.sa.p.psi_wr, M.sa.o_psi1 + (Off)k * M.sa.st_psi, e.lo);
            stg(O.xi_x, M.sa.o_xi + (Off)k * M.sa.st_xi, e.xn);
        }
    }
    if (ZS >= 0) {
        const LayerCoef<T> cz = lds_coef(M.lds, M.h1, ZS, mz);
        const Pair<T> e = cpml_pair(cz, d.zh, d.zl, d.zx, pm, pc, pp, d2m, M.imm);
        damp += f * e.t;
        slab += e.t;
        if (own) {
            T* pw = M.pm_side[ZS < 0 ? 0 : ZS].psi_wr;
            stg(pw, M.col + (Off)mz * M.sm, e.hi);
            if (mz == M.mfirst[ZS < 0 ? 0 : ZS]) stg(pw, M.col + (Off)(mz - 1) * M.sm, e.lo);
            stg(O.xi_z[ZS < 0 ? 0 : ZS], M.col + (Off)(mz - 1) * M.sm, e.xn);
        }
    }
    const T lap = d2a + d2m;
    if (AL) {
        // whole rows: lanes outside the interior store the Dirichlet 0
        const T v = two * pc - d.pc + f * lap + damp;
        if (O.st) stg_stream(O.p2, O.scol + (Off)k * M.sm, O.in1 ? v : T(0));
        if (O.st && O.q2) stg(O.q2, O.scol + (Off)k * M.sm, O.in1 ? lap + slab : T(0));
    } else {
        if (own) stg_stream(O.p2, M.col + (Off)k * M.sm, two * pc - d.pc + f * lap + damp);
        if (own && O.q2) stg(O.q2, M.col + (Off)k * M.sm, lap + slab);
    }
}

// Step t+1 loads of row k (2D): pp = pcur at row k+1 (every row below the
// last), the edge lanes' x neighbour, pold, fact and the slab values.
template <typename T, typename Off, int SA, bool GEN>
__device__ __forceinline__ NodeLoads<T> two2_load_node(const March<T, Off, 2>& M, int ny, int k,
                                                       int mside) {
    NodeLoads<T> q = {};
    constexpr Off E = (Off)sizeof(T);
    const Off c = M.col + (Off)k * M.sm;
    const int lane = (int)threadIdx.x;
    if (!GEN || k + 1 <= ny - 1) q.pp = ldg(M.pcur, c + M.sm);
    if (GEN && (k < 1 || k > ny - 2)) return q;
    q.xm = ldg(M.pcur, lane == 0 ? c - E : (lane == kTileX - 1 ? c + E : c));
    q.po = ldg_stream(M.pold, c);
    q.f = ldg(M.fact, c);
    if (SA == 2) {
        q.psa0 = ldg(M.sa.p.psi_rd, M.sa.o_psi + (Off)k * M.sa.st_psi);
        q.psa1 = ldg(M.sa.p.psi_rd, M.sa.o_psi1 + (Off)k * M.sa.st_psi);
        q.xia = ldg((const T*)M.sa.p.xi, M.sa.o_xi + (Off)k * M.sa.st_xi);
    }
    if (mside >= 0) {
        const int mz = k - (mside ? M.mshift[1] : M.mshift[0]);
        const T* pr = mside ? M.pm_side[1].psi_rd : M.pm_side[0].psi_rd;
        const T* xr = mside ? M.pm_side[1].xi : M.pm_side[0].xi;
        q.psm0 = ldg(pr, M.col + (Off)mz * M.sm);
        q.psm1 = ldg(pr, M.col + (Off)(mz - 1) * M.sm);
        q.xim = ldg(xr, M.col + (Off)(mz - 1) * M.sm);
    }
    return q;
}

template <typename T, typename Off, int SA, bool ZM, bool GEN>
__device__ __forceinline__ NodeLoads<T> two2_load(const TwoGeom& G, const March<T, Off, 2>& M,
                                                  int k) {
    int mz;
    const int zs = ZM ? classify(k, G.nz, G.halo, G.act, &mz) : -1;
    return two2_load_node<T, Off, SA, GEN>(M, G.nz, k, zs);
}

template <typename T, typename Off, int SA, bool ZM>
__device__ __forceinline__ Mid<T> two2_first_at(const TwoGeom& G, const March<T, Off, 2>& M,
                                                int k, const NodeLoads<T>& q, T pm, T pc, T xm,
                                                T xp) {
    if (!ZM) return two2_first<T, Off, SA, -1>(M, 0, q, pm, pc, xm, xp);
    int mz;
    const int zs = classify(k, G.nz, G.halo, G.act, &mz);
    if (zs == 0) return two2_first<T, Off, SA, 0>(M, mz, q, pm, pc, xm, xp);
    if (zs == 1) return two2_first<T, Off, SA, 1>(M, mz, q, pm, pc, xm, xp);
    return two2_first<T, Off, SA, -1>(M, mz, q, pm, pc, xm, xp);
}

template <typename T, typename Off, int SA, bool ZM, bool AL>
__device__ __forceinline__ void two2_second_at(const TwoGeom& G, const March<T, Off, 2>& M,
                                               const TwoOut<T, Off>& O, const Mid<T>& d, T pm,
                                               T pp, T xm, T xp, bool own, int k) {
    if (!ZM) {
        two2_second<T, Off, SA, -1, AL>(M, O, d, pm, pp, xm, xp, own, 0, k);
        return;
    }
    int mz;
    const int zs = classify(k, G.nz, G.halo, G.act, &mz);
    if (zs == 0) two2_second<T, Off, SA, 0, AL>(M, O, d, pm, pp, xm, xp, own, mz, k);
    else if (zs == 1) two2_second<T, Off, SA, 1, AL>(M, O, d, pm, pp, xm, xp, own, mz, k);
    else two2_second<T, Off, SA, -1, AL>(M, O, d, pm, pp, xm, xp, own, mz, k);
}

// One march position k (row) of a 2D wave; GEN as in two_iter.
template <typename T, typename Off, int PF, int SA, bool ZM, bool AL, bool GEN, int U>
__device__ __forceinline__ void two2_iter(const Stencil2Args<T>& B, const March<T, Off, 2>& M,
                                          const TwoOut<T, Off>& O, const TwoGeom& G,
                                          const TwoSrc<T>& S, TwoCtx<T, PF>& C, int k) {
    const int lane = (int)threadIdx.x;
    Mid<T> dn = {};
    T en = T(0);  // aligned tiles, edge lanes: their side node's p^{t+1} at row k
    const T pp = own_reg(C.qr[U].pp);
    const T epc = C.qr[U].xm;  // edge lanes: pcur at their side node, row k
    if (!GEN || (k >= 1 && k <= G.nz - 2)) {
        const T pc = C.pz1;
        const T pcx = G.inx ? pc : T(0);
        const T sl = lane_shift(pcx, true), sr = lane_shift(pcx, false);
        const T edge = C.qr[U].xm;
        const T xm = lane == 0 ? edge : sl;
        const T xp = lane == kTileX - 1 ? edge : sr;
        dn = two2_first_at<T, Off, SA, ZM>(G, M, k, C.qr[U], C.pz0, pc, xm, xp);
        // inject_sources! of step t+1, in source order
        if (k >= G.ksrc_lo && k <= G.ksrc_hi) {
            const int64_t lin = G.col_el + (int64_t)k * G.plane;
            for (int e = 0; e < G.ns; ++e)
                if (S.lin[e] == lin) dn.p1 += S.w[e];
        }
        dn.p1 = G.in1 ? dn.p1 : T(0);
        if (AL && G.eln) {
            // aligned tiles: lanes 0 and 63 also evaluate step t+1 at their
            // side's node outside the tile (x = 64*tx-1, 64*tx+64), the plain
            // stencil in the owning wave's order of operations (no CPML slab
            // reaches it); its pcur at row k+1 is the next ring slot's
            // x-neighbour load, its inner x neighbour this lane's pcur
            NodeLoads<T> q = {};
            q.pp = C.qr[(U + 1) % PF].xm;
            q.po = C.es[U].pe;
            q.f = C.es[U].fe;
            const T xo = C.es[U].xo;
            Mid<T> e = two2_first<T, Off, 0, -1>(M, 0, q, C.ez0, epc, lane == 0 ? xo : pc,
                                                 lane == 0 ? pc : xo);
            if (k >= G.ksrc_lo && k <= G.ksrc_hi) {
                const int64_t lin = G.hcol_el + (int64_t)k * G.plane;
                for (int s = 0; s < G.ns; ++s)
                    if (S.lin[s] == lin) e.p1 += S.w[s];
            }
            en = G.hin ? e.p1 : T(0);
        }
    }
    if (AL) C.ez0 = epc;
    C.pz0 = C.pz1;
    C.pz1 = pp;
    const int kr = k + PF;
    if (!GEN || kr <= G.kend) {
        C.qr[U] = two2_load<T, Off, SA, ZM, GEN>(G, M, kr);
        if (AL) C.es[U] = edge_slot_load2<T, Off, GEN>(G, M, kr);
    }
    __builtin_amdgcn_sched_barrier(0);
    if (AL) {
        // whole rows (dn.p1 is 0 outside the interior)
        if (O.st && (!GEN || (k >= G.kbeg && k < G.kend))) {
            stg_stream(O.p1, O.scol + (Off)k * M.sm, dn.p1);
            if (O.q1) stg(O.q1, O.scol + (Off)k * M.sm, G.in1 ? dn.q : T(0));
        }
    } else if (G.own && (!GEN || (k >= G.kbeg && k < G.kend))) {
        stg_stream(O.p1, M.col + (Off)k * M.sm, dn.p1);
        if (O.q1) stg(O.q1, M.col + (Off)k * M.sm, dn.q);
    }
    if (!GEN || k - 1 >= G.kbeg) {
        T xm = lane_shift(C.d.p1, true);
        T xp = lane_shift(C.d.p1, false);
        if (AL) {
            xm = lane == 0 ? C.ep1 : xm;
            xp = lane == kTileX - 1 ? C.ep1 : xp;
        }
        two2_second_at<T, Off, SA, ZM, AL>(G, M, O, C.d, C.p1m, dn.p1, xm, xp, G.own, k - 1);
    }
    C.p1m = C.d.p1;
    C.d = dn;
    if (AL) C.ep1 = en;
}

template <typename T, typename Off, int PF, int SA, bool ZM, bool AL, bool GEN, bool CHECK,
          int... Us>
__device__ __forceinline__ void two2_group(const Stencil2Args<T>& B, const March<T, Off, 2>& M,
                                           const TwoOut<T, Off>& O, const TwoGeom& G,
                                           const TwoSrc<T>& S, TwoCtx<T, PF>& C, int k,
                                           std::integer_sequence<int, Us...>) {
    ((!CHECK || k + Us <= G.kend
          ? (two2_iter<T, Off, PF, SA, ZM, AL, GEN, Us % PF>(B, M, O, G, S, C, k + Us), 0)
          : 0),
     ...);
}

// The march of one 2D wave: head, steady state, tail as in two_march.
template <typename T, typename Off, int PF, int SA, bool ZM, bool AL>
__device__ __forceinline__ void two2_march(const Stencil2Args<T>& B, const March<T, Off, 2>& M,
                                           const TwoOut<T, Off>& O, const TwoGeom& G,
                                           const TwoSrc<T>& S) {
    static_assert(!AL || PF >= 2, "aligned tiles read the next ring slot");
    TwoCtx<T, PF> C;
    C.pz0 = G.kbeg >= 2 ? ldg(M.pcur, M.col + (Off)(G.kbeg - 2) * M.sm) : T(0);
    C.pz1 = ldg(M.pcur, M.col + (Off)(G.kbeg - 1) * M.sm);
#pragma unroll
    for (int u = 0; u < PF; ++u) {
        const int kk = G.kbeg - 1 + u;
        C.qr[u] = kk <= G.kend ? two2_load<T, Off, SA, ZM, true>(G, M, kk) : NodeLoads<T>{};
        if (AL) C.es[u] = kk <= G.kend ? edge_slot_load2<T, Off, true>(G, M, kk) : EdgeSlot<T>{};
        __builtin_amdgcn_sched_barrier(0);
    }
    C.ez0 = T(0);
    if (AL && G.eln && G.kbeg >= 2)
        C.ez0 = ldg(M.pcur, M.col + (Off)(G.kbeg - 2) * M.sm + (Off)G.dxe);
    C.ep1 = T(0);
    C.d = Mid<T>{};
    C.p1m = T(0);
    C.slot = 0;
    constexpr int H = (2 + PF - 1) / PF * PF;
    int k = G.kbeg - 1;
    two2_group<T, Off, PF, SA, ZM, AL, true, true>(B, M, O, G, S, C, k,
                                                   std::make_integer_sequence<int, H>{});
    k += H;
    const int last = (G.kend < G.nz - 2 ? G.kend : G.nz - 2) - PF;
    // drain the head's memory operations once, as in two_march, so that the
    // steady loop's waits follow its own issue order
    __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0)
    for (; k + PF - 1 <= last; k += PF)
        two2_group<T, Off, PF, SA, ZM, AL, false, false>(B, M, O, G, S, C, k,
                                                         std::make_integer_sequence<int, PF>{});
    for (; k <= G.kend; k += PF)
        two2_group<T, Off, PF, SA, ZM, AL, true, true>(B, M, O, G, S, C, k,
                                                       std::make_integer_sequence<int, PF>{});
}

template <typename T, typename Off, int BY, int PF>
__device__ __forceinline__ void stencil2d_two_block(const Stencil2Args<T>& B, int64_t it0, int lb) {
    __shared__ TwoSrc<T> S;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* lds = reinterpret_cast<T*>(smem_raw);
    const StencilArgs<T>& A = B.s;

    // logical order: x tile fastest, then chunk group, shot
    int rest = lb;
    const int tx = rest % A.tiles_x; rest /= A.tiles_x;
    const int ty = rest % A.tiles_y;
    const int shot = rest / A.tiles_y;
    const int ns = two_stage_sources(B, it0, shot, S);
    if (A.active[1][0] || A.active[1][1]) stage_march_coef(A, 1, lds);
    __syncthreads();
    const int wy = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int chunk = ty * BY + wy;
    if (chunk >= A.nchunks) return;  // no barrier follows
    const int lane = (int)threadIdx.x;
    const int64_t nx = A.nx, ny = A.ny, halo = A.halo;
    int kbeg, kend;
    two_chunk(B, chunk, ny, &kbeg, &kend);
    // any y slab row among the rows this chunk reads (wave-uniform)
    const bool zm = (A.active[1][0] && kbeg - 1 <= halo) || (A.active[1][1] && kend >= ny - halo - 1);
    // aligned 64-node tiles for chunks no y slab reaches (decided per chunk,
    // so that all tiles of a chunk share one tiling)
    const bool al = PF >= 2 && B.align && !zm;
    if (tx >= (al ? B.tiles64 : B.tiles62)) return;  // a tile of the other tiling
    const int64_t i = al ? (int64_t)tx * kTileX + lane : (int64_t)tx * kOwnX + lane;
    const int64_t ic = i < 1 ? 1 : (i > nx - 2 ? nx - 2 : i);
    const bool in1 = i == ic;
    const bool own = in1 && (al || (lane >= 1 && lane <= kOwnX));

    constexpr Off E = (Off)sizeof(T);
    const int64_t fo = (int64_t)shot * A.field_stride;
    March<T, Off, 2> M;
    M.pcur = A.pcur + fo;
    M.pold = A.pold + fo;
    M.pnew = nullptr;
    M.fact = A.fact;
    M.qout = nullptr;
    M.sy = (Off)A.px * E;
    M.sm = M.sy;
    M.col = (Off)ic * E;
    M.i2a = A.inv_d2[0];
    M.i2b = T(0);
    M.i2m = A.inv_d2[1];
    M.ia = A.inv_d[0];
    M.ib = T(0);
    M.imm = A.inv_d[1];
    M.lds = lds;
    M.h1 = (int)halo + 1;
    M.store_p = true;
    const Off hp1 = (Off)halo + 1, hl = (Off)halo;
    // psi_x: (halo+1) x ny, xi_x: halo x ny
    col_slab<T, Off>(A, 0, ic, nx, shot, 0, E, 0, E, hp1 * E, hl * E, M.sa);
    M.sb = M.sa;
    M.sb.in = false;
    // psi_y: px x (halo+1), xi_y: px x halo
    for (int s = 0; s < 2; ++s) M.pm_side[s] = slab_ptrs(A, 1, s, shot);
    M.mshift[0] = 0;
    M.mshift[1] = (int)(ny - halo - 2);
    M.mfirst[0] = 1;
    M.mfirst[1] = right_first(ny, halo);

    TwoOut<T, Off> O;
    O.p1 = B.p1 + fo;
    O.p2 = B.p2 + fo;
    O.q1 = B.q1 ? B.q1 + fo : nullptr;
    O.q2 = B.q2 ? B.q2 + fo : nullptr;
    int mx;
    const int sx = classify(ic, nx, halo, A.active[0], &mx);
    O.xi_x = (sx == 1 ? B.xi_out[0][1] : B.xi_out[0][0]) + shot * A.xi_stride[0];
    O.xi_y = nullptr;
    O.xi_z[0] = B.xi_out[1][0] + shot * A.xi_stride[1];
    O.xi_z[1] = B.xi_out[1][1] + shot * A.xi_stride[1];
    // aligned tiles store whole rows: every lane inside the row pitch
    O.st = al && i < A.px;
    O.in1 = in1;
    O.scol = O.st ? (Off)i * E : M.col;

    // rows holding this shot's sources
    TwoGeom G;
    G.ksrc_lo = 1 << 30;
    G.ksrc_hi = -1;
    G.plane = A.px;
    G.ns = ns;
    for (int e = 0; e < ns; ++e) {
        const int kk = (int)(S.lin[e] / G.plane);
        G.ksrc_lo = kk < G.ksrc_lo ? kk : G.ksrc_lo;
        G.ksrc_hi = kk > G.ksrc_hi ? kk : G.ksrc_hi;
    }
    G.col_el = ic;
    G.wy = wy;
    G.kbeg = kbeg;
    G.kend = kend;
    G.nz = (int)ny;
    G.halo = (int)halo;
    G.act[0] = A.active[1][0];
    G.act[1] = A.active[1][1];
    G.in1 = in1;
    G.own = own;
    G.inx = in1;
    G.iny = true;
    {
        // aligned tiles: lanes 0 and 63 also evaluate their side's node just
        // outside the tile (x = 64*tx-1, 64*tx+64; loads clamped into the
        // interior like every load)
        G.eln = al && (lane == 0 || lane == kTileX - 1);
        const int64_t xe = lane == 0 ? i - 1 : i + 1;
        const int64_t xec = xe < 1 ? 1 : (xe > nx - 2 ? nx - 2 : xe);
        const int64_t xeo = lane == 0 ? xec - 1 : xec + 1;
        G.dxe = (uint32_t)((Off)(xec - ic) * E);
        G.dxo = (uint32_t)((Off)(xeo - ic) * E);
        G.hcol_el = xec;
        G.hin = xe == xec;
        // this edge lane's entry of the edge copy of fact: [y][tile][side]
        G.fedge = B.fact_edge;
        G.fe_off = (uint32_t)(((Off)tx * 2 + (lane == 0 ? 0 : 1)) * E);
        G.fe_step = (uint32_t)((Off)B.tiles64 * 2 * E);
    }

    const bool xs = __any(M.sa.in);
    if constexpr (PF >= 2) {
        if (al) {
            // no y slab in the chunk
            if (xs) two2_march<T, Off, PF, 2, false, true>(B, M, O, G, S);
            else two2_march<T, Off, PF, 0, false, true>(B, M, O, G, S);
            return;
        }
    }
    if (xs) {
        if (zm) two2_march<T, Off, PF, 2, true, false>(B, M, O, G, S);
        else two2_march<T, Off, PF, 2, false, false>(B, M, O, G, S);
    } else {
        if (zm) two2_march<T, Off, PF, 0, true, false>(B, M, O, G, S);
        else two2_march<T, Off, PF, 0, false, false>(B, M, O, G, S);
    }
}

// ---------------------------------------------------------------------
// Entry kernels.  A plain launch covers one plan.  A batch launch covers up
// to kMaxBatch plans (different grids, same precision) in one grid, as the
// single-step batch kernel does: logical block b belongs to the plan p with
// block_start[p] <= b < block_start[p+1], whose arguments are table[p].
// Every plan of a batch launch runs both steps.
// ---------------------------------------------------------------------
template <typename T>
struct Batch2Args {
    const Stencil2Args<T>* table;  // per-plan arguments for this launch's buffer roles
    int64_t it0;                   // 0-based index of step t+1
    int block_start[kMaxBatch + 1];
    int nplans;
    int grid_blocks;
};

template <typename T, typename Off, int W, int PF>
__global__ __launch_bounds__(64 * W) void stencil3d_two(const Stencil2Args<T> B) {
    const int lb = logical_block(blockIdx.x, B.s.grid_blocks);
    if (lb >= B.s.logical_blocks) return;  // whole block
    stencil3d_two_block<T, Off, W, PF>(B, B.it0, lb);
}

template <typename T, typename Off, int BY, int PF>
__global__ __launch_bounds__(64 * BY) void stencil2d_two(const Stencil2Args<T> B) {
    const int lb = logical_block(blockIdx.x, B.s.grid_blocks);
    if (lb >= B.s.logical_blocks) return;  // whole block
    stencil2d_two_block<T, Off, BY, PF>(B, B.it0, lb);
}

template <typename T, typename Off, int W, int PF>
__global__ __launch_bounds__(64 * W) void stencil3d_two_batch(const Batch2Args<T> B) {
    const int lb = logical_block(blockIdx.x, B.grid_blocks);
    if (lb >= B.block_start[B.nplans]) return;  // whole block
    const int p = batch_plan(B.block_start, B.nplans, lb);
    stencil3d_two_block<T, Off, W, PF>(B.table[p], B.it0, lb - B.block_start[p]);
}

template <typename T, typename Off, int BY, int PF>
__global__ __launch_bounds__(64 * BY) void stencil2d_two_batch(const Batch2Args<T> B) {
    const int lb = logical_block(blockIdx.x, B.grid_blocks);
    if (lb >= B.block_start[B.nplans]) return;  // whole block
    const int p = batch_plan(B.block_start, B.nplans, lb);
    stencil2d_two_block<T, Off, BY, PF>(B.table[p], B.it0, lb - B.block_start[p]);
}

}  // namespace awcpml
