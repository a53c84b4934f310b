// stencil_kernels.h -- gfx950 kernels of one CPML acoustic time step.
//
// One launch of stencil2d/stencil3d performs, for every shot of a plan, what
// the reference does with update_psi_{x,y,z}! followed by update_p!
// (src/shared/acoustic_2D_xPU.jl:7-75, src/shared/acoustic_3D_xPU.jl:8-104):
//
//  * every thread owns one (x, y) column (3D) or one x position (2D) and
//    marches along z (3D) or y (2D) through a chunk of the grid; the
//    pressure values below, at and above the current node along the march
//    stay in registers (pm, pc, pp), so pcur is fetched from HBM once;
//    the x/y neighbours are plain loads that hit L1/L2 because the
//    neighbouring lanes and waves fetch the same lines;
//  * the psi recursions are fused into the pressure update.  psi is
//    double-buffered (psi_rd holds step n-1, psi_wr receives step n): the
//    thread at node m recomputes psi[m] and psi[m-1] from psi_rd and the
//    three pressures it holds, then writes psi[m] (and psi[0] when m == 1),
//    so no second launch and no grid-wide ordering are needed and each psi
//    value is produced with exactly the reference's expression;
//  * only threads inside a CPML slab touch psi/xi, so their traffic is paid
//    in the slabs only; a side whose coefficients are a = 0, b_K = 1 (the
//    free top) is skipped entirely;
//  * the arithmetic is written in the reference's evaluation order and the
//    file is compiled with -ffp-contract=off, so the FP64 instantiation
//    reproduces the reference operation for operation.
//
// Blocks are mapped XCD-aware: the hardware hands block b to XCD b % 8, and
// logical_block() renumbers so that each XCD walks one contiguous range of
// tiles (shot fastest, then x tile, y tile, march chunk).  Neighbouring
// tiles, which share halo lines, and the shots of one tile, which share
// `fact`, then run on the same XCD and meet in its L2.
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
    int logical_blocks;
    int grid_blocks;          // launched blocks, a multiple of kNumXcd
};

__device__ __forceinline__ int logical_block(int b, int grid) {
    const int per = grid / kNumXcd;
    return (b % kNumXcd) * per + b / kNumXcd;
}

// One CPML side of one axis at layer m (1 <= m <= halo) of that side's psi
// array.  pm, pc, pp: pressure before, at and after the node along the
// axis; d2: the node's second derivative along the axis; f: fact.
// Returns fact*(dpsi + xi), the reference's damping term
// (acoustic_2D_xPU.jl:45-56 for x-left; the other sides are identical up
// to index shifts).
template <typename T>
__device__ __forceinline__ T cpml_term(const CpmlSide<T>& s, int64_t psi_m,
                                       int64_t psi_m1, int64_t xi_m1, int m,
                                       T pm, T pc, T pp, T d2, T inv_d, T f,
                                       int64_t shot_psi, int64_t shot_xi) {
    const T* psi_rd = s.psi_rd + shot_psi;
    T* psi_wr = s.psi_wr + shot_psi;
    T* xi = s.xi + shot_xi;
    // update_psi_*!: psi[m] = b_K_h[m]*psi[m] + a_h[m]*(p[m+1]-p[m])*_d
    const T psi_here = s.bKh[m] * psi_rd[psi_m] + s.ah[m] * (pp - pc) * inv_d;
    const T psi_prev = s.bKh[m - 1] * psi_rd[psi_m1] + s.ah[m - 1] * (pc - pm) * inv_d;
    psi_wr[psi_m] = psi_here;
    if (m == 1) psi_wr[psi_m1] = psi_prev;
    // update_p!: dpsi, in-place xi, damping
    const T dpsi = (psi_here - psi_prev) * inv_d;
    const T xin = s.bK[m - 1] * xi[xi_m1] + s.a[m - 1] * (d2 + dpsi);
    xi[xi_m1] = xin;
    return f * (dpsi + xin);
}

// Classify a node along one axis: side 0/1 and layer m, or side -1.
// Mirrors `if i <= halo+1 ... elseif i >= nx-halo` (1-based) with the
// 0-based node index q: left if q <= halo, right if q >= n-halo-1, where
// the right layer is m = q - (n-halo-2).  An inactive side still takes its
// branch (so the other side is not tested), as in the reference.
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

// ---------------------------------------------------------------------
// 3D: z-march.  Block (64, BY): lane -> x, threadIdx.y -> y.
// ---------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void stencil3d(const StencilArgs<T> A) {
    const int lb = logical_block(blockIdx.x, A.grid_blocks);
    if (lb >= A.logical_blocks) return;
    int rest = lb;
    const int shot = rest % A.nshots; rest /= A.nshots;
    const int tx = rest % A.tiles_x; rest /= A.tiles_x;
    const int ty = rest % A.tiles_y;
    const int chunk = rest / A.tiles_y;

    const int64_t i = (int64_t)tx * kTileX + threadIdx.x;
    const int64_t j = 1 + (int64_t)ty * blockDim.y + threadIdx.y;
    const int64_t nx = A.nx, ny = A.ny, nz = A.nz, halo = A.halo;
    if (i < 1 || i > nx - 2 || j > ny - 2) return;
    const int64_t kbeg = 1 + (int64_t)chunk * A.march_len;
    const int64_t kend = min(kbeg + (int64_t)A.march_len, nz - 1);
    if (kbeg >= kend) return;

    const int64_t px = A.px, plane = px * ny;
    const int64_t fo = (int64_t)shot * A.field_stride;
    const T* __restrict__ pcur = A.pcur + fo;
    const T* __restrict__ pold = A.pold + fo;
    T* __restrict__ pnew = A.pnew + fo;
    const T* __restrict__ fact = A.fact;
    const T two = T(2);
    const T idx2 = A.inv_d2[0], idy2 = A.inv_d2[1], idz2 = A.inv_d2[2];

    int mx = 0, my = 0;
    const int sx = classify(i, nx, halo, A.active[0], &mx);
    const int sy = classify(j, ny, halo, A.active[1], &my);
    const int64_t hp1 = halo + 1;

    const int64_t col = i + px * j;
    T pm = pcur[col + plane * (kbeg - 1)];
    T pc = pcur[col + plane * kbeg];
    for (int64_t k = kbeg; k < kend; ++k) {
        const int64_t idx = col + plane * k;
        const T pp = pcur[idx + plane];
        const T xm = pcur[idx - 1], xp = pcur[idx + 1];
        const T ym = pcur[idx - px], yp = pcur[idx + px];
        const T po = pold[idx];
        const T f = fact[idx];
        const T d2x = (xp - two * pc + xm) * idx2;
        const T d2y = (yp - two * pc + ym) * idy2;
        const T d2z = (pp - two * pc + pm) * idz2;
        T damp = T(0);
        if (sx >= 0) {
            // psi_x: (halo+1) x ny x nz ; xi_x: halo x ny x nz
            const int64_t r = j + ny * k;
            damp += cpml_term(A.side[0][sx], mx + hp1 * r, (mx - 1) + hp1 * r,
                              (mx - 1) + halo * r, mx, xm, pc, xp, d2x, A.inv_d[0], f,
                              shot * A.psi_stride[0], shot * A.xi_stride[0]);
        }
        if (sy >= 0) {
            // psi_y: nx x (halo+1) x nz ; xi_y: nx x halo x nz (pitch px)
            damp += cpml_term(A.side[1][sy], i + px * (my + hp1 * k),
                              i + px * ((my - 1) + hp1 * k), i + px * ((my - 1) + halo * k),
                              my, ym, pc, yp, d2y, A.inv_d[1], f,
                              shot * A.psi_stride[1], shot * A.xi_stride[1]);
        }
        int mz = 0;
        const int sz = classify(k, nz, halo, A.active[2], &mz);
        if (sz >= 0) {
            // psi_z: nx x ny x (halo+1) ; xi_z: nx x ny x halo (pitch px)
            damp += cpml_term(A.side[2][sz], col + plane * mz, col + plane * (mz - 1),
                              col + plane * (mz - 1), mz, pm, pc, pp, d2z, A.inv_d[2], f,
                              shot * A.psi_stride[2], shot * A.xi_stride[2]);
        }
        pnew[idx] = two * pc - po + f * (d2x + d2y + d2z) + damp;
        pm = pc;
        pc = pp;
    }
}

// ---------------------------------------------------------------------
// 2D: y-march.  Block (64, BY): lane -> x, threadIdx.y -> march chunk.
// ---------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void stencil2d(const StencilArgs<T> A) {
    const int lb = logical_block(blockIdx.x, A.grid_blocks);
    if (lb >= A.logical_blocks) return;
    int rest = lb;
    const int shot = rest % A.nshots; rest /= A.nshots;
    const int tx = rest % A.tiles_x; rest /= A.tiles_x;
    const int ty = rest;  // group of blockDim.y chunks
    const int chunk = ty * blockDim.y + threadIdx.y;

    const int64_t i = (int64_t)tx * kTileX + threadIdx.x;
    const int64_t nx = A.nx, ny = A.ny, halo = A.halo;
    if (i < 1 || i > nx - 2 || chunk >= A.nchunks) return;
    const int64_t jbeg = 1 + (int64_t)chunk * A.march_len;
    const int64_t jend = min(jbeg + (int64_t)A.march_len, ny - 1);
    if (jbeg >= jend) return;

    const int64_t px = A.px;
    const int64_t fo = (int64_t)shot * A.field_stride;
    const T* __restrict__ pcur = A.pcur + fo;
    const T* __restrict__ pold = A.pold + fo;
    T* __restrict__ pnew = A.pnew + fo;
    const T* __restrict__ fact = A.fact;
    const T two = T(2);
    const T idx2 = A.inv_d2[0], idy2 = A.inv_d2[1];

    int mx = 0;
    const int sx = classify(i, nx, halo, A.active[0], &mx);
    const int64_t hp1 = halo + 1;

    T pm = pcur[i + px * (jbeg - 1)];
    T pc = pcur[i + px * jbeg];
    for (int64_t j = jbeg; j < jend; ++j) {
        const int64_t idx = i + px * j;
        const T pp = pcur[idx + px];
        const T xm = pcur[idx - 1], xp = pcur[idx + 1];
        const T po = pold[idx];
        const T f = fact[idx];
        const T d2x = (xp - two * pc + xm) * idx2;
        const T d2y = (pp - two * pc + pm) * idy2;
        T damp = T(0);
        if (sx >= 0) {
            // psi_x: (halo+1) x ny ; xi_x: halo x ny
            damp += cpml_term(A.side[0][sx], mx + hp1 * j, (mx - 1) + hp1 * j,
                              (mx - 1) + halo * j, mx, xm, pc, xp, d2x, A.inv_d[0], f,
                              shot * A.psi_stride[0], shot * A.xi_stride[0]);
        }
        int my = 0;
        const int sy = classify(j, ny, halo, A.active[1], &my);
        if (sy >= 0) {
            // psi_y: nx x (halo+1) ; xi_y: nx x halo (pitch px)
            damp += cpml_term(A.side[1][sy], i + px * my, i + px * (my - 1),
                              i + px * (my - 1), my, pm, pc, pp, d2y, A.inv_d[1], f,
                              shot * A.psi_stride[1], shot * A.xi_stride[1]);
        }
        pnew[idx] = two * pc - po + f * (d2x + d2y) + damp;
        pm = pc;
        pc = pp;
    }
}

// ---------------------------------------------------------------------
// Sources and receivers of step `it0` (0-based), one block per shot:
// inject_sources! then record_receivers! (acoustic_2D_xPU.jl:77-91,126-128).
// Sources of a shot that share a node are added by one thread in source
// order (seq[shot] = 1), otherwise one thread per source.
// ---------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void sources_receivers(
    T* __restrict__ pnew, int64_t field_stride,
    const int64_t* __restrict__ src_lin, const T* __restrict__ src_wav,
    const int* __restrict__ src_start, const int* __restrict__ src_seq,
    const int64_t* __restrict__ rec_lin, T* __restrict__ traces,
    const int* __restrict__ rec_start, int64_t nt, int64_t it0) {
    const int shot = blockIdx.x;
    T* p = pnew + (int64_t)shot * field_stride;
    const int s0 = src_start[shot], s1 = src_start[shot + 1];
    if (src_seq[shot]) {
        if (threadIdx.x == 0)
            for (int e = s0; e < s1; ++e) p[src_lin[e]] += src_wav[(int64_t)e * nt + it0];
    } else {
        for (int e = s0 + threadIdx.x; e < s1; e += blockDim.x)
            p[src_lin[e]] += src_wav[(int64_t)e * nt + it0];
    }
    __syncthreads();
    const int r0 = rec_start[shot], r1 = rec_start[shot + 1];
    for (int r = r0 + threadIdx.x; r < r1; r += blockDim.x)
        traces[(int64_t)r * nt + it0] = p[rec_lin[r]];
}

}  // namespace awcpml
