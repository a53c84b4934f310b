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
// for psi, so one launch per step reads lam^{s+1}, lam^{s+2} and q of
// every shot and writes lam^s; f and the gradient accumulator are read
// once per group of shots.  Every absorbing side touches the nodes
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
    T* grad;                  // accumulators of fbar, one per shot group
    T inv_d[3];
    T inv_d2[3];
    AdjSide<T> side[3][2];
    int active[3][2];
    int64_t psi_stride[3];    // elements between two shots' alpha slabs
    int64_t xi_stride[3];
    int ndim;
    int nshots;
    int group;                // shots per thread (one gradient accumulator per group)
    int tiles_x, tiles_y;     // 64-wide x tiles, row groups
    // block classes: per axis (x tiles, row groups, 3D planes) the tiles
    // none of whose nodes lies in the CPML range of an absorbing side,
    // [int_lo, int_hi) of ntiles; class c (bit a set: the tiles of axis a
    // outside that range) holds the launch's logical blocks
    // [cls_start[c], cls_start[c+1])
    int ntiles[3], int_lo[3], int_hi[3];
    int cls_start[9];
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

// CPML data of one node on one axis.  The node lies at m = c - off in the
// range 0 <= m <= halo+1 of at most one active side; v0/v1/v2 say whether
// the nodes m-1, m, m+1 are slab layers (1..halo), e_lo/e_hi whether the
// alpha edges m-1 and m exist, w_lo whether this node stores edge m-1
// (edge 0 of the low side, which no other node owns).
template <typename T>
struct AdjCoef {
    bool ext, v0, v1, v2, e_lo, e_hi, w_lo;
    int side, m;
    T a0, a1, a2;      // a of layers m-1, m, m+1
    T bK1;             // bK of layer m (the node's own beta)
    T bKh_lo, ah_lo;   // edge m-1
    T bKh_hi, ah_hi;   // edge m
};

template <typename T>
__device__ __forceinline__ AdjCoef<T> adj_coef(const AdjointArgs<T>& A, int a, int64_t c) {
    AdjCoef<T> C;
    C.ext = C.v0 = C.v1 = C.v2 = C.e_lo = C.e_hi = C.w_lo = false;
    C.side = 0;
    C.m = 0;
    C.a0 = C.a1 = C.a2 = C.bK1 = C.bKh_lo = C.ah_lo = C.bKh_hi = C.ah_hi = T(0);
    const int64_t h = A.halo;
    int64_t m = -1;
    int side = 0;
    if (A.active[a][0] && c <= h + 1) {
        m = c;
    } else if (A.active[a][1] && c >= A.n[a] - h - 2) {
        m = c - (A.n[a] - h - 2);
        side = 1;
    }
    if (m < 0 || m > h + 1) return C;
    const T* av = side ? A.side[a][1].a : A.side[a][0].a;
    const T* bK = side ? A.side[a][1].bK : A.side[a][0].bK;
    const T* bKh = side ? A.side[a][1].bKh : A.side[a][0].bKh;
    const T* ah = side ? A.side[a][1].ah : A.side[a][0].ah;
    C.ext = true;
    C.side = side;
    C.m = (int)m;
    C.v0 = m >= 2 && m - 1 <= h;
    C.v1 = m >= 1 && m <= h;
    C.v2 = m + 1 <= h;
    if (C.v0) C.a0 = av[m - 2];
    if (C.v1) {
        C.a1 = av[m - 1];
        C.bK1 = bK[m - 1];
    }
    if (C.v2) C.a2 = av[m];
    C.e_lo = m >= 1;
    C.e_hi = m <= h;
    C.w_lo = m == 1 && side == 0;
    if (C.e_lo) {
        C.bKh_lo = bKh[m - 1];
        C.ah_lo = ah[m - 1];
    }
    if (C.e_hi) {
        C.bKh_hi = bKh[m];
        C.ah_hi = ah[m];
    }
    return C;
}

// The slab values one node reads on one axis: beta^{s+1} of layers m-1,
// m, m+1 and alpha^{s+2} of edges m-1, m.
template <typename T>
struct AdjVals {
    T b0, b1, b2, alo, ahi;
};

// ob: offset of the node's own beta (layer m); oa: offset of alpha edge
// m-1; ls: offset between neighbouring layers/edges.
template <typename T>
__device__ __forceinline__ void adj_vals(const AdjCoef<T>& C, const T* be, const T* al,
                                         int64_t ob, int64_t oa, int64_t ls, AdjVals<T>& V) {
    V.b0 = C.v0 ? be[ob - ls] : T(0);
    V.b1 = C.v1 ? be[ob] : T(0);
    V.b2 = C.v2 ? be[ob + ls] : T(0);
    V.alo = C.e_lo ? al[oa] : T(0);
    V.ahi = C.e_hi ? al[oa + ls] : T(0);
}

// One axis's slab terms of lam^s (added to acc) and its alpha^{s+1} stores.
// um/uc/up: f lam^{s+1} one node below, at and above the node on the axis.
template <typename T>
__device__ __forceinline__ void adj_term(const AdjCoef<T>& C, const AdjVals<T>& V, T um, T uc,
                                         T up, T inv_d, T inv_d2, T* alw, int64_t oa,
                                         int64_t ls, T& acc) {
    const T z0 = C.v0 ? C.a0 * V.b0 : T(0);
    const T z1 = C.v1 ? C.a1 * V.b1 : T(0);
    const T z2 = C.v2 ? C.a2 * V.b2 : T(0);
    const T w0 = C.v0 ? um + z0 : T(0);
    const T w1 = C.v1 ? uc + z1 : T(0);
    const T w2 = C.v2 ? up + z2 : T(0);
    T y_lo = T(0), y_hi = T(0);
    if (C.e_lo) {
        const T alv = C.bKh_lo * V.alo + (w0 - w1) * inv_d;
        y_lo = C.ah_lo * alv;
        if (C.w_lo) alw[oa] = alv;
    }
    if (C.e_hi) {
        const T alv = C.bKh_hi * V.ahi + (w1 - w2) * inv_d;
        y_hi = C.ah_hi * alv;
        alw[oa + ls] = alv;
    }
    acc += (y_lo - y_hi) * inv_d;
    acc += (z0 - T(2) * z1 + z2) * inv_d2;
}

// CPML data of one node on one axis with its slab offsets: ob of the
// node's own beta layer m-1, oa of alpha edge m-1, ls between neighbouring
// layers/edges (offsets follow slab_index).
template <typename T>
struct AdjAxis {
    AdjCoef<T> C;
    int64_t ob, oa, ls;
};

template <typename T>
__device__ __forceinline__ AdjAxis<T> adj_axis(const AdjointArgs<T>& A, int a, int64_t i,
                                               int64_t j, int64_t k) {
    AdjAxis<T> X;
    const int64_t c = a == 0 ? i : (a == 1 ? j : k);
    X.C = adj_coef<T>(A, a, c);
    const int64_t h = A.halo, ny = A.n[1], px = A.px, L = X.C.m - 1;
    X.ob = slab_index(a, h, L, i, j, k, px, ny);
    X.oa = slab_index(a, h + 1, L, i, j, k, px, ny);
    X.ls = a == 0 ? 1 : (a == 1 ? px : px * ny);
    return X;
}

// Everything one shot reads at one node: lam^{s+1} at the node and its
// axis neighbours, lam^{s+2}, q^s and (CP) the slab values of every axis
// whose CPML range holds the node.
template <typename T, int NDIM, bool CP>
struct AdjShot {
    T lc, lm[NDIM], lp[NDIM], l2, q;
    AdjVals<T> v[CP ? NDIM : 1];
};

template <typename T, int NDIM, bool CP>
__device__ __forceinline__ void adj_load_shot(const AdjointArgs<T>& A, const AdjAxis<T>* X,
                                              int g, int64_t lin, const int64_t* so,
                                              AdjShot<T, NDIM, CP>& S) {
    const int64_t fs = (int64_t)g * A.field_stride;
    const T* l1 = A.lam1 + fs;
    S.lc = l1[lin];
#pragma unroll
    for (int a = 0; a < NDIM; ++a) {
        S.lm[a] = l1[lin - so[a]];
        S.lp[a] = l1[lin + so[a]];
    }
    S.l2 = A.lam2[fs + lin];
    S.q = A.q[fs + lin];
    if constexpr (CP) {
#pragma unroll
        for (int a = 0; a < NDIM; ++a) {
            if (!X[a].C.ext) continue;
            const bool hi = X[a].C.side != 0;
            const T* be = hi ? A.side[a][1].be_rd : A.side[a][0].be_rd;
            const T* al = hi ? A.side[a][1].al_rd : A.side[a][0].al_rd;
            adj_vals(X[a].C, be + g * A.xi_stride[a], al + g * A.psi_stride[a], X[a].ob,
                     X[a].oa, X[a].ls, S.v[a]);
        }
    }
}

// Shot g's reverse step at the node from its loads S: stores lam^s (and,
// CP, the slab values) and returns lam^s q^s, the node's gradient term.
template <typename T, int NDIM, bool CP>
__device__ __forceinline__ T adj_shot(const AdjointArgs<T>& A, const AdjAxis<T>* X, int g,
                                      int64_t lin, T fc, const T* fm, const T* fp,
                                      const AdjShot<T, NDIM, CP>& S) {
    const int64_t fs = (int64_t)g * A.field_stride;
    const T uc = fc * S.lc;
    T um[NDIM], up[NDIM];
    T acc = T(2) * S.lc - S.l2;
#pragma unroll
    for (int a = 0; a < NDIM; ++a) {
        um[a] = fm[a] * S.lm[a];
        up[a] = fp[a] * S.lp[a];
        acc += (up[a] - T(2) * uc + um[a]) * A.inv_d2[a];
    }
    if constexpr (CP) {
#pragma unroll
        for (int a = 0; a < NDIM; ++a) {
            if (!X[a].C.ext) continue;
            T* alw = (X[a].C.side ? A.side[a][1].al_wr : A.side[a][0].al_wr) +
                     g * A.psi_stride[a];
            adj_term(X[a].C, S.v[a], um[a], uc, up[a], A.inv_d[a], A.inv_d2[a], alw, X[a].oa,
                     X[a].ls, acc);
        }
    }
    A.lam0[fs + lin] = acc;
    if constexpr (CP) {
        // beta^s = bK beta^{s+1} + f lam^s on every slab holding the node
#pragma unroll
        for (int a = 0; a < NDIM; ++a) {
            if (!(X[a].C.ext && X[a].C.v1)) continue;
            T* bew = (X[a].C.side ? A.side[a][1].be_wr : A.side[a][0].be_wr) +
                     g * A.xi_stride[a];
            bew[X[a].ob] = X[a].C.bK1 * S.v[a].b1 + fc * acc;
        }
    }
    return acc * S.q;
}

// One shot of the three-slot pipeline: refill slot F with shot g+2's
// loads, then compute shot g from slot C (loaded two shots earlier).
template <typename T, int NDIM, bool CP>
__device__ __forceinline__ void adj_pipe(const AdjointArgs<T>& A, const AdjAxis<T>* X, int g,
                                         int g1, int64_t lin, const int64_t* so, T fc,
                                         const T* fm, const T* fp, AdjShot<T, NDIM, CP>& C,
                                         AdjShot<T, NDIM, CP>& F, T& gsum) {
    if (g >= g1) return;
    if (g + 2 < g1) adj_load_shot<T, NDIM, CP>(A, X, g + 2, lin, so, F);
    // keep the later shots' loads ahead of this shot's arithmetic
    __builtin_amdgcn_sched_barrier(0);
    gsum += adj_shot<T, NDIM, CP>(A, X, g, lin, fc, fm, fp, C);
}

// Tiles of axis a in block class c, and the u-th of them.
__device__ __forceinline__ int adj_class_tiles(int c, int n, int lo, int hi, int a) {
    return ((c >> a) & 1) ? n - (hi - lo) : hi - lo;
}
__device__ __forceinline__ int adj_class_tile(int c, int lo, int hi, int a, int u) {
    if (!((c >> a) & 1)) return lo + u;
    return u < lo ? u : hi + (u - lo);
}

// One reverse step.  Block (64, BY): lane -> x, threadIdx.y -> y row; a
// thread owns one interior node (plane k in 3D) and walks the shots of
// its shot group, so fact, the CPML coefficients and the gradient
// accumulator of the node are read once per group and the group's
// gradient terms are summed in registers (one read-modify-write of the
// group's accumulator per step).  Shots run through a three-slot pipeline:
// the loads of shots g+1 and g+2 are in flight while shot g is computed.
// CP = false: the launch covers block class 0 (no node of the block in a
// CPML range) and carries no slab code, so it runs with the registers of
// the plain stencil; CP = true: classes 1 and up.  Logical blocks within a
// class: x tile fastest, then row group, plane (3D), shot group,
// renumbered XCD-aware.
template <typename T, int NDIM, bool CP>
__global__ __launch_bounds__(256) void adjoint_step(const AdjointArgs<T> A) {
    const int lb = logical_block(blockIdx.x, A.grid_blocks);
    if (lb >= A.logical_blocks) return;
    const int64_t nx = A.n[0], ny = A.n[1], px = A.px;
    int cls = 0;
    if (CP)
        for (int c = 1; c < 8; ++c)
            if (lb >= A.cls_start[c]) cls = c;
    int rest = lb - A.cls_start[cls];
    int t[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int na = adj_class_tiles(cls, A.ntiles[a], A.int_lo[a], A.int_hi[a], a);
        t[a] = adj_class_tile(cls, A.int_lo[a], A.int_hi[a], a, rest % na);
        rest /= na;
    }
    const int grp = rest;
    const int64_t k = NDIM == 3 ? 1 + t[2] : 0;
    const int64_t i = (int64_t)t[0] * kTileX + threadIdx.x;
    // one row per wavefront: j (and the y axis's CPML data) is wave-uniform
    const int wy = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int64_t j = 1 + (int64_t)t[1] * blockDim.y + wy;
    if (i < 1 || i > nx - 2 || j > ny - 2) return;
    const int64_t so[3] = {1, px, px * ny};
    const int64_t lin = i + px * j + so[2] * k;
    const T* f = A.fact;
    const T fc = f[lin];
    T fm[NDIM], fp[NDIM];
#pragma unroll
    for (int a = 0; a < NDIM; ++a) {
        fm[a] = f[lin - so[a]];
        fp[a] = f[lin + so[a]];
    }
    AdjAxis<T> X[CP ? NDIM : 1];
    if constexpr (CP) {
#pragma unroll
        for (int a = 0; a < NDIM; ++a) X[a] = adj_axis<T>(A, a, i, j, k);
    }
    const int g0 = grp * A.group;
    const int g1 = g0 + A.group < A.nshots ? g0 + A.group : A.nshots;
    T gsum = T(0);
    AdjShot<T, NDIM, CP> s0, s1, s2;
    adj_load_shot<T, NDIM, CP>(A, X, g0, lin, so, s0);
    if (g0 + 1 < g1) adj_load_shot<T, NDIM, CP>(A, X, g0 + 1, lin, so, s1);
    for (int g = g0; g < g1; g += 3) {
        adj_pipe<T, NDIM, CP>(A, X, g, g1, lin, so, fc, fm, fp, s0, s2, gsum);
        adj_pipe<T, NDIM, CP>(A, X, g + 1, g1, lin, so, fc, fm, fp, s1, s0, gsum);
        adj_pipe<T, NDIM, CP>(A, X, g + 2, g1, lin, so, fc, fm, fp, s2, s1, gsum);
    }
    T* gr = A.grad + (int64_t)grp * A.field_stride;
    gr[lin] += gsum;
}

// Trace cotangent injection of step it0 (0-based) after adjoint_step:
// for each distinct interior receiver node of a shot, r = the sum of the
// cotangents of its receivers; lam^s += r, beta^s += f r on every
// absorbing slab holding the node (beta^s = f lam^s + ...), and r q^s,
// the node's extra gradient term, is added to the node's own entry of rq
// (entry u of the receiver node list), which fold_receivers adds to the
// gradient.  One block per shot; entries of different shots never share
// memory, so the shots run in parallel and the sums stay in a fixed order.
template <typename T>
struct AdjIoArgs {
    T* lam0;
    T* rq;                     // per receiver node entry: sum of r q^s
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
        A.rq[u] += r * A.q[fs + lin];
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

// Add the receiver gradient terms rq to the gradient total and clear them:
// for each distinct receiver node v over all shots, total[lin[v]] += the
// sum of its entries rq[idx[e]], start[v] <= e < start[v+1], in shot
// order.
template <typename T, typename A>
__global__ __launch_bounds__(256) void fold_receivers(T* rq, A* total, const int64_t* lin,
                                                      const int* start, const int* idx, int count) {
    const int v = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (v >= count) return;
    A acc = A(0);
    for (int e = start[v]; e < start[v + 1]; ++e) {
        acc += (A)rq[idx[e]];
        rq[idx[e]] = T(0);
    }
    total[lin[v]] += acc;
}

// Fold the shot groups' partial gradients of the last reverse steps into
// the FP64 total: total += sum over groups (in group order), partials = 0.
// The partial sums then run over at most one checkpoint segment and one
// group of shots in the plan's precision, and the long sum over time and
// groups is FP64.
template <typename T>
__global__ __launch_bounds__(256) void fold_gradient(T* grad, double* total, int64_t field_stride,
                                                     int ngroups) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= field_stride) return;
    double acc = total[t];
    for (int s = 0; s < ngroups; ++s) {
        T& g = grad[(int64_t)s * field_stride + t];
        acc += (double)g;
        g = T(0);
    }
    total[t] = acc;
}

// Sum of the shot groups' gradients into group 0's accumulator.
template <typename T>
__global__ __launch_bounds__(256) void sum_shots(T* grad, int64_t field_stride, int ngroups) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= field_stride) return;
    T acc = grad[t];
    for (int s = 1; s < ngroups; ++s) acc += grad[(int64_t)s * field_stride + t];
    grad[t] = acc;
}

}  // namespace awcpml
