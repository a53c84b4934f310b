// adjoint_kernels.h -- reverse-time (adjoint) step of the CPML time loop,
// for the FWI gradient with respect to fact = vel^2 dt^2.
//
// Forward step s (interior nodes, per absorbing side with layer index m
// of the node along the side's axis, m = q - off, off = 0 on the low side
// and n - halo - 2 on the high side; slab nodes are 1 <= m <= halo):
//   psi^s[e] = bKh[e] psi^{s-1}[e] + ah[e] (p^{s-1}[off+e+1] - p^{s-1}[off+e]) / d
//   xi^s[m-1] = bK[m-1] xi^{s-1}[m-1] + a[m-1] (d2 p^{s-1} + (psi^s[m] - psi^s[m-1]) / d)
//   p^s = 2 p^{s-1} - p^{s-2} + f (L p^{s-1} + sum (dpsi + xi^s)) + sources
// with traces^s = p^s at the receivers.  For a cotangent tbar of the
// traces, the adjoint variables (lam of p, alpha of psi, beta of xi) obey,
// for s = nt .. 1:
//   lam^s = R^T tbar^s + 2 lam^{s+1} - lam^{s+2} + L(f lam^{s+1})
//           + sum_sides [ (ah[m-1] alpha^{s+1}[m-1] - ah[m] alpha^{s+1}[m]) / d
//                         + D2(z) ],          z = a beta^{s+1} on slab nodes
//   beta^s   = f lam^s + bK beta^{s+1}                     (slab nodes)
//   alpha^{s+1}[e] = bKh[e] alpha^{s+2}[e] + (w[off+e] - w[off+e+1]) / d,
//            w = f lam^{s+1} + a beta^{s+1} on slab nodes, 0 elsewhere
//   fbar += lam^s * q^s,  q^s = L p^{s-1} + sum (dpsi + xi^s)
// (q^s is written by the forward kernel during the checkpoint replay).
// Dirichlet nodes are constants of the forward map, so lam stays zero on
// them.  alpha^{s+1} is recomputed where it is needed from alpha^{s+2} and
// the w values around the node, the same fusion the forward kernel uses
// for psi, so one launch per step reads lam^{s+1}, lam^{s+2}, f, q and
// the gradient and writes lam^s.  Every absorbing side touches the nodes
// 0 <= m <= halo+1 ("extended" range); the plan requires n >= 2 halo + 4
// on every absorbing axis so the two sides' ranges are disjoint.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace awcpml {

template <typename T>
struct AdjSide {
    const T* al_rd;  // alpha^{s+2} (psi layout, halo+1 layers)
    T* al_wr;        // alpha^{s+1}
    const T* be_rd;  // beta^{s+1} (xi layout, halo layers)
    T* be_wr;        // beta^s
    const T* a;
    const T* bK;
    const T* ah;
    const T* bKh;
};

template <typename T>
struct AdjointArgs {
    int64_t n[3];             // n[2] = 1 in 2D
    int64_t px;
    int64_t halo;
    int64_t field_stride;     // elements between two shots' fields
    const T* lam1;            // lam^{s+1}
    const T* lam2;            // lam^{s+2}
    T* lam0;                  // lam^s (written)
    const T* fact;            // shared by the shots
    const T* q;               // q^s
    T* grad;                  // per-shot accumulators of fbar
    T inv_d[3];
    T inv_d2[3];
    AdjSide<T> side[3][2];
    int active[3][2];
    int64_t psi_stride[3];    // elements between two shots' alpha slabs
    int64_t xi_stride[3];
    int ndim;
    int nshots;
    int64_t columns;          // nodes per shot handled by one thread's march
    int march_len, nchunks;
    int64_t total_threads;
};

// Element index of layer L of axis a's slab at node (i, j, k) in a slab
// with `layers` layers along the axis (layouts as the forward's psi/xi).
__device__ __forceinline__ int64_t slab_index(int a, int64_t layers, int64_t L, int64_t i,
                                              int64_t j, int64_t k, int64_t px, int64_t ny) {
    if (a == 0) return L + layers * (j + ny * k);
    if (a == 1) return i + px * (L + layers * k);
    return i + px * (j + ny * L);
}

// One node's adjoint update; (i, j, k) interior, lin its element offset.
template <typename T, int NDIM>
__device__ __forceinline__ void adjoint_node(const AdjointArgs<T>& A, int shot, int64_t i,
                                             int64_t j, int64_t k) {
    const int64_t px = A.px, ny = A.n[1];
    const int64_t fs = (int64_t)shot * A.field_stride;
    const int64_t lin = i + px * (j + ny * k);
    const T* l1 = A.lam1 + fs;
    const T* f = A.fact;
    const int64_t stride[3] = {1, px, px * ny};
    const T fc = f[lin];
    const T lc = l1[lin];
    const T uc = fc * lc;
    T acc = T(2) * lc - A.lam2[fs + lin];
    T um[3], up[3];
#pragma unroll
    for (int a = 0; a < NDIM; ++a) {
        const int64_t sa = stride[a];
        um[a] = f[lin - sa] * l1[lin - sa];
        up[a] = f[lin + sa] * l1[lin + sa];
        acc += (up[a] - T(2) * uc + um[a]) * A.inv_d2[a];
    }
    const int64_t h = A.halo;
    const int64_t c[3] = {i, j, k};
    // slab sides of this node whose beta must be advanced after lam^s
    int bs_axis[3], bs_side[3], bs_layer[3];
    int64_t bs_idx[3];
    T bs_old[3];
    int nbs = 0;
#pragma unroll
    for (int a = 0; a < NDIM; ++a) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            if (!A.active[a][s]) continue;
            const int64_t off = s ? A.n[a] - h - 2 : 0;
            const int64_t m = c[a] - off;
            if (m < 0 || m > h + 1) continue;
            const AdjSide<T>& S = A.side[a][s];
            const T* be = S.be_rd + shot * A.xi_stride[a];
            const T* al = S.al_rd + shot * A.psi_stride[a];
            T* alw = S.al_wr + shot * A.psi_stride[a];
            T z[3], w[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const int64_t mm = m + d - 1;
                const T u = d == 0 ? um[a] : (d == 1 ? uc : up[a]);
                if (mm >= 1 && mm <= h) {
                    const T bv = be[slab_index(a, h, mm - 1, i, j, k, px, ny)];
                    z[d] = S.a[mm - 1] * bv;
                    w[d] = u + z[d];
                    if (d == 1) {
                        bs_axis[nbs] = a;
                        bs_side[nbs] = s;
                        bs_layer[nbs] = (int)(mm - 1);
                        bs_idx[nbs] = slab_index(a, h, mm - 1, i, j, k, px, ny);
                        bs_old[nbs] = bv;
                        ++nbs;
                    }
                } else {
                    z[d] = T(0);
                    w[d] = T(0);
                }
            }
            T y_lo = T(0), y_hi = T(0);
            if (m >= 1) {
                const int64_t e = m - 1;
                const int64_t ie = slab_index(a, h + 1, e, i, j, k, px, ny);
                const T alv = S.bKh[e] * al[ie] + (w[0] - w[1]) * A.inv_d[a];
                y_lo = S.ah[e] * alv;
                if (m == 1 && s == 0) alw[ie] = alv;
            }
            if (m <= h) {
                const int64_t ie = slab_index(a, h + 1, m, i, j, k, px, ny);
                const T alv = S.bKh[m] * al[ie] + (w[1] - w[2]) * A.inv_d[a];
                y_hi = S.ah[m] * alv;
                alw[ie] = alv;
            }
            acc += (y_lo - y_hi) * A.inv_d[a];
            acc += (z[0] - T(2) * z[1] + z[2]) * A.inv_d2[a];
        }
    }
    A.lam0[fs + lin] = acc;
    for (int b = 0; b < nbs; ++b) {
        const int a = bs_axis[b], s = bs_side[b];
        const AdjSide<T>& S = A.side[a][s];
        S.be_wr[shot * A.xi_stride[a] + bs_idx[b]] = S.bK[bs_layer[b]] * bs_old[b] + fc * acc;
    }
    A.grad[fs + lin] += acc * A.q[fs + lin];
}

// One thread per (x, y) column (3D) or x column (2D) of one shot, marching
// over a chunk of march_len nodes along the last axis; x is the fastest
// thread index so a wavefront reads 64 consecutive nodes of a row.
template <typename T, int NDIM>
__global__ __launch_bounds__(256) void adjoint_step(const AdjointArgs<T> A) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.total_threads) return;
    const int64_t nx = A.n[0], ny = A.n[1], nz = A.n[2];
    // t = ((shot * nchunks + chunk) * rows + row) * (nx - 2) + (i - 1)
    const int64_t rowlen = nx - 2;
    const int64_t rows = (NDIM == 3) ? (ny - 2) : 1;
    int64_t r = t;
    const int64_t i = 1 + r % rowlen;
    r /= rowlen;
    const int64_t row = r % rows;
    r /= rows;
    const int chunk = (int)(r % A.nchunks);
    const int shot = (int)(r / A.nchunks);
    const int64_t mext = (NDIM == 3) ? nz : ny;
    const int64_t k0 = 1 + (int64_t)chunk * A.march_len;
    int64_t k1 = k0 + A.march_len;
    if (k1 > mext - 1) k1 = mext - 1;
    if (NDIM == 3) {
        const int64_t j = 1 + row;
        for (int64_t k = k0; k < k1; ++k) adjoint_node<T, 3>(A, shot, i, j, k);
    } else {
        for (int64_t k = k0; k < k1; ++k) adjoint_node<T, 2>(A, shot, i, k, 0);
    }
}

// Trace cotangent injection of step it0 (0-based) after adjoint_step:
// for each distinct interior receiver node of a shot, r = the sum of the
// cotangents of its receivers; lam^s += r, fbar += r q^s, and beta^s +=
// f r on every absorbing slab holding the node (beta^s = f lam^s + ...).
template <typename T>
struct AdjIoArgs {
    T* lam0;
    T* grad;
    const T* q;
    const T* fact;
    int64_t field_stride;
    const int64_t* node_lin;   // distinct interior receiver nodes, all shots
    const int* node_start;     // nshots+1 CSR offsets into node_lin
    const int* rec_start;      // per node: CSR offsets into rec_idx
    const int* rec_idx;        // receiver index (all shots) of each entry
    const T* tbar;             // nt values per receiver
    int64_t nt;
    int64_t n[3];
    int64_t px, halo;
    int ndim;
    AdjSide<T> side[3][2];
    int active[3][2];
    int64_t xi_stride[3];
};

template <typename T>
__global__ __launch_bounds__(256) void adjoint_inject(const AdjIoArgs<T> A, int64_t it0) {
    const int shot = blockIdx.x;
    const int u0 = A.node_start[shot], u1 = A.node_start[shot + 1];
    const int64_t fs = (int64_t)shot * A.field_stride;
    for (int u = u0 + threadIdx.x; u < u1; u += blockDim.x) {
        T r = T(0);
        for (int e = A.rec_start[u]; e < A.rec_start[u + 1]; ++e)
            r += A.tbar[(int64_t)A.rec_idx[e] * A.nt + it0];
        const int64_t lin = A.node_lin[u];
        A.lam0[fs + lin] += r;
        A.grad[fs + lin] += r * A.q[fs + lin];
        const T fr = A.fact[lin] * r;
        const int64_t ny = A.n[1];
        const int64_t c[3] = {lin % A.px, (lin / A.px) % ny, lin / (A.px * ny)};
        const int64_t h = A.halo;
        for (int a = 0; a < A.ndim; ++a)
            for (int s = 0; s < 2; ++s) {
                if (!A.active[a][s]) continue;
                const int64_t m = c[a] - (s ? A.n[a] - h - 2 : 0);
                if (m < 1 || m > h) continue;
                const int64_t idx = slab_index(a, h, m - 1, c[0], c[1], c[2], A.px, ny);
                A.side[a][s].be_wr[shot * A.xi_stride[a] + idx] += fr;
            }
    }
}

// Sum of the per-shot gradients into shot 0's accumulator.
template <typename T>
__global__ __launch_bounds__(256) void sum_shots(T* grad, int64_t field_stride, int nshots) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= field_stride) return;
    T acc = grad[t];
    for (int s = 1; s < nshots; ++s) acc += grad[(int64_t)s * field_stride + t];
    grad[t] = acc;
}

}  // namespace awcpml
