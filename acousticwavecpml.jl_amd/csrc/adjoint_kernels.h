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

#include "stencil_kernels.h"  // kTileX, logical_block

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
    int march_len, nchunks;   // march (z in 3D, y in 2D) chunking
    int tiles_x, tiles_y;     // 64-wide x tiles; 3D: row groups, 2D: chunk groups
    int logical_blocks, grid_blocks;
};

// Element index of layer L of axis a's slab at node (i, j, k) in a slab
// with `layers` layers along the axis (layouts as the forward's psi/xi).
__device__ __forceinline__ int64_t slab_index(int a, int64_t layers, int64_t L, int64_t i,
                                              int64_t j, int64_t k, int64_t px, int64_t ny) {
    if (a == 0) return L + layers * (j + ny * k);
    if (a == 1) return i + px * (L + layers * k);
    return i + px * (j + ny * L);
}

// x-axis CPML data of one lane (constant along the march): whether the
// lane's column lies in the range 0 <= m <= halo+1 of an active x side,
// the coefficients of the layers and edges it touches, and the offsets of
// its alpha/beta rows.  Only the slab values themselves are loaded per node.
template <typename T>
struct AdjX {
    bool ext;
    int m;
    bool v0, v1, v2;           // mm = m-1, m, m+1 are slab layers (1..halo)
    T a0, a1, a2;              // a[mm-1] of those layers
    T bK1;                     // bK[m-1] (the node's own beta)
    bool e_lo, e_hi, w_lo;     // edges m-1 / m exist; this lane writes alpha[m-1]
    T bKh_lo, ah_lo, bKh_hi, ah_hi;
    const T* be_rd;
    T* be_wr;
    const T* al_rd;
    T* al_wr;
    int64_t be0, al0;          // element offset of layer 0 of the row at march index 0
    int64_t be_step, al_step;  // per march step
};

template <typename T, int NDIM>
__device__ __forceinline__ AdjX<T> adj_xslab(const AdjointArgs<T>& A, int shot, int64_t i,
                                             int64_t j) {
    AdjX<T> X;
    X.ext = false;
    X.m = 0;
    X.v0 = X.v1 = X.v2 = X.e_lo = X.e_hi = X.w_lo = false;
    X.a0 = X.a1 = X.a2 = X.bK1 = X.bKh_lo = X.ah_lo = X.bKh_hi = X.ah_hi = T(0);
    X.be_rd = X.al_rd = nullptr;
    X.be_wr = X.al_wr = nullptr;
    X.be0 = X.al0 = X.be_step = X.al_step = 0;
    const int64_t h = A.halo, n = A.n[0], ny = A.n[1];
    int side = -1;
    int64_t m = 0;
    for (int s = 0; s < 2 && side < 0; ++s) {
        if (!A.active[0][s]) continue;
        const int64_t mm = i - (s ? n - h - 2 : 0);
        if (mm >= 0 && mm <= h + 1) {
            side = s;
            m = mm;
        }
    }
    if (side < 0) return X;
    const AdjSide<T>& S = A.side[0][side];
    X.ext = true;
    X.m = (int)m;
    X.v0 = m - 1 >= 1 && m - 1 <= h;
    X.v1 = m >= 1 && m <= h;
    X.v2 = m + 1 >= 1 && m + 1 <= h;
    if (X.v0) X.a0 = S.a[m - 2];
    if (X.v1) {
        X.a1 = S.a[m - 1];
        X.bK1 = S.bK[m - 1];
    }
    if (X.v2) X.a2 = S.a[m];
    X.e_lo = m >= 1;
    X.e_hi = m <= h;
    X.w_lo = m == 1 && side == 0;
    if (X.e_lo) {
        X.bKh_lo = S.bKh[m - 1];
        X.ah_lo = S.ah[m - 1];
    }
    if (X.e_hi) {
        X.bKh_hi = S.bKh[m];
        X.ah_hi = S.ah[m];
    }
    X.be_rd = S.be_rd + shot * A.xi_stride[0];
    X.be_wr = S.be_wr + shot * A.xi_stride[0];
    X.al_rd = S.al_rd + shot * A.psi_stride[0];
    X.al_wr = S.al_wr + shot * A.psi_stride[0];
    // slab_index(0, L, ...) = L + layers * (j + ny * k): 3D march over k,
    // 2D march over the second index (passed as j, with k = 0)
    if (NDIM == 3) {
        X.be0 = h * j;
        X.al0 = (h + 1) * j;
        X.be_step = h * ny;
        X.al_step = (h + 1) * ny;
    } else {
        X.be_step = h;
        X.al_step = h + 1;
    }
    return X;
}

// One node's adjoint update.  (i, j, k) interior (2D: j is the march index
// and k = 0), lin its element offset in the shot's field; fc, lc: fact and
// lam^{s+1} at the node; um/up: f lam^{s+1} one node below/above along each
// axis; l2, qv, gv: lam^{s+2}, q^s and the gradient at the node.
template <typename T, int NDIM>
__device__ __forceinline__ void adjoint_core(const AdjointArgs<T>& A, const AdjX<T>& X, int64_t km,
                                             int shot, int64_t i,
                                             int64_t j, int64_t k, int64_t lin, int64_t fs,
                                             T fc, T lc, const T* um, const T* up, T l2, T qv,
                                             T gv) {
    const int64_t px = A.px, ny = A.n[1];
    const T uc = fc * lc;
    T acc = T(2) * lc - l2;
#pragma unroll
    for (int a = 0; a < NDIM; ++a) acc += (up[a] - T(2) * uc + um[a]) * A.inv_d2[a];
    const int64_t h = A.halo;
    const int64_t c[3] = {i, j, k};
    // x axis from the lane's precomputed data (km: march index)
    T xb1 = T(0);
    int64_t xbo = 0;
    if (X.ext) {
        xbo = X.be0 + km * X.be_step;
        const int64_t xao = X.al0 + km * X.al_step;
        const int mm = X.m;
        const T b0 = X.v0 ? X.be_rd[xbo + mm - 2] : T(0);
        xb1 = X.v1 ? X.be_rd[xbo + mm - 1] : T(0);
        const T b2 = X.v2 ? X.be_rd[xbo + mm] : T(0);
        const T z0 = X.v0 ? X.a0 * b0 : T(0);
        const T z1 = X.v1 ? X.a1 * xb1 : T(0);
        const T z2 = X.v2 ? X.a2 * b2 : T(0);
        const T w0 = X.v0 ? um[0] + z0 : T(0);
        const T w1 = X.v1 ? uc + z1 : T(0);
        const T w2 = X.v2 ? up[0] + z2 : T(0);
        T y_lo = T(0), y_hi = T(0);
        if (X.e_lo) {
            const T alv = X.bKh_lo * X.al_rd[xao + mm - 1] + (w0 - w1) * A.inv_d[0];
            y_lo = X.ah_lo * alv;
            if (X.w_lo) X.al_wr[xao + mm - 1] = alv;
        }
        if (X.e_hi) {
            const T alv = X.bKh_hi * X.al_rd[xao + mm] + (w1 - w2) * A.inv_d[0];
            y_hi = X.ah_hi * alv;
            X.al_wr[xao + mm] = alv;
        }
        acc += (y_lo - y_hi) * A.inv_d[0];
        acc += (z0 - T(2) * z1 + z2) * A.inv_d2[0];
    }
    // slab sides (y, z) of this node whose beta must be advanced after lam^s
    int bs_axis[3], bs_side[3], bs_layer[3];
    int64_t bs_idx[3];
    T bs_old[3];
    int nbs = 0;
#pragma unroll
    for (int a = 1; a < NDIM; ++a) {
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
    if (X.ext && X.v1) X.be_wr[xbo + X.m - 1] = X.bK1 * xb1 + fc * acc;
    for (int b = 0; b < nbs; ++b) {
        const int a = bs_axis[b], s = bs_side[b];
        const AdjSide<T>& S = A.side[a][s];
        S.be_wr[shot * A.xi_stride[a] + bs_idx[b]] = S.bK[bs_layer[b]] * bs_old[b] + fc * acc;
    }
    A.grad[fs + lin] = gv + acc * qv;
}

// Loads of one node that are not carried along the march: the in-plane
// neighbours of lam^{s+1} and fact, lam^{s+2}, q^s and the gradient.
template <typename T>
struct AdjNear {
    T lxm, lxp, fxm, fxp, lym, lyp, fym, fyp, l2, qv, gv;
};

template <typename T, int NDIM>
__device__ __forceinline__ void adj_load_near(const T* l1, const T* f, const T* l2, const T* q,
                                              const T* g, int64_t lin, int64_t sy, AdjNear<T>& n) {
    n.lxm = l1[lin - 1];
    n.lxp = l1[lin + 1];
    n.fxm = f[lin - 1];
    n.fxp = f[lin + 1];
    if (NDIM == 3) {
        n.lym = l1[lin - sy];
        n.lyp = l1[lin + sy];
        n.fym = f[lin - sy];
        n.fyp = f[lin + sy];
    }
    n.l2 = l2[lin];
    n.qv = q[lin];
    n.gv = g[lin];
}

// One reverse step.  Block (64, BY) as the forward kernel: lane -> x,
// threadIdx.y -> y row (3D) or march chunk (2D); each thread marches a
// chunk of march_len nodes along the last axis.  f lam^{s+1} along the
// march is carried in registers (one new plane per node), and every load
// of node k+1 is issued before node k is computed.  Logical blocks: x tile
// fastest, then row/chunk group, chunk (3D), shot, renumbered XCD-aware.
template <typename T, int NDIM>
__global__ __launch_bounds__(256) void adjoint_step(const AdjointArgs<T> A) {
    const int lb = logical_block(blockIdx.x, A.grid_blocks);
    if (lb >= A.logical_blocks) return;
    const int64_t nx = A.n[0], ny = A.n[1], nz = A.n[2], px = A.px;
    int rest = lb;
    const int tx = rest % A.tiles_x;
    rest /= A.tiles_x;
    const int ty = rest % A.tiles_y;
    rest /= A.tiles_y;
    const int wy = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    int chunk, shot;
    int64_t j = 0;
    if (NDIM == 3) {
        chunk = rest % A.nchunks;
        shot = rest / A.nchunks;
        j = 1 + (int64_t)ty * blockDim.y + wy;
        if (j > ny - 2) return;
    } else {
        shot = rest;
        chunk = ty * (int)blockDim.y + wy;
        if (chunk >= A.nchunks) return;
    }
    const int64_t i = (int64_t)tx * kTileX + threadIdx.x;
    if (i < 1 || i > nx - 2) return;
    const int64_t mext = NDIM == 3 ? nz : ny;
    const int64_t k0 = 1 + (int64_t)chunk * A.march_len;
    const int64_t k1 = k0 + A.march_len < mext - 1 ? k0 + A.march_len : mext - 1;
    if (k0 >= k1) return;
    const int64_t ms = NDIM == 3 ? px * ny : px;  // march stride
    const int64_t fs = (int64_t)shot * A.field_stride;
    const T* l1 = A.lam1 + fs;
    const T* l2 = A.lam2 + fs;
    const T* qq = A.q + fs;
    const T* gg = A.grad + fs;
    const T* f = A.fact;
    const int64_t base = NDIM == 3 ? i + px * j : i;
    int64_t lin = base + k0 * ms;
    T u_m = f[lin - ms] * l1[lin - ms];
    T lc = l1[lin], fc = f[lin];
    T lp = l1[lin + ms], fp = f[lin + ms];
    AdjNear<T> nr;
    adj_load_near<T, NDIM>(l1, f, l2, qq, gg, lin, px, nr);
    const AdjX<T> X = adj_xslab<T, NDIM>(A, shot, i, j);
    for (int64_t k = k0; k < k1; ++k, lin += ms) {
        // next node's loads (planes stay inside the field at the chunk end)
        const int64_t kk = k + 2 < mext ? k + 2 : mext - 1;
        const T lpp = l1[base + kk * ms], fpp = f[base + kk * ms];
        AdjNear<T> nn;
        adj_load_near<T, NDIM>(l1, f, l2, qq, gg, lin + ms, px, nn);
        // keep the next node's loads ahead of this node's arithmetic
        __builtin_amdgcn_sched_barrier(0);
        T um[3], up[3];
        um[0] = nr.fxm * nr.lxm;
        up[0] = nr.fxp * nr.lxp;
        if (NDIM == 3) {
            um[1] = nr.fym * nr.lym;
            up[1] = nr.fyp * nr.lyp;
        }
        um[NDIM - 1] = u_m;
        up[NDIM - 1] = fp * lp;
        if (NDIM == 3)
            adjoint_core<T, 3>(A, X, k, shot, i, j, k, lin, fs, fc, lc, um, up, nr.l2, nr.qv, nr.gv);
        else
            adjoint_core<T, 2>(A, X, k, shot, i, k, 0, lin, fs, fc, lc, um, up, nr.l2, nr.qv, nr.gv);
        u_m = fc * lc;
        lc = lp;
        fc = fp;
        lp = lpp;
        fp = fpp;
        nr = nn;
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
