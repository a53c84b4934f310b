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

// An axis across the march (x, and y in 3D): its CPML data is the same for
// every node of a thread's march, and the slab offsets advance by a fixed
// stride per march step.
template <typename T>
struct AdjAxis {
    AdjCoef<T> C;
    const T* be_rd;
    T* be_wr;
    const T* al_rd;
    T* al_wr;
    int64_t ob0, oa0;   // offsets (own beta layer, alpha edge m-1) at march index 0
    int64_t ls;         // layer stride
    int64_t sb, sa;     // beta / alpha offset per march step
};

// Offsets follow slab_index: x slabs are [march][row][layer] (layer
// fastest), y slabs (3D) are [z][layer][x].
template <typename T, int NDIM>
__device__ __forceinline__ AdjAxis<T> adj_axis(const AdjointArgs<T>& A, int a, int shot,
                                               int64_t i, int64_t j) {
    AdjAxis<T> X;
    X.C = adj_coef<T>(A, a, a == 0 ? i : j);
    const int s = X.C.side;
    X.be_rd = (s ? A.side[a][1].be_rd : A.side[a][0].be_rd) + shot * A.xi_stride[a];
    X.be_wr = (s ? A.side[a][1].be_wr : A.side[a][0].be_wr) + shot * A.xi_stride[a];
    X.al_rd = (s ? A.side[a][1].al_rd : A.side[a][0].al_rd) + shot * A.psi_stride[a];
    X.al_wr = (s ? A.side[a][1].al_wr : A.side[a][0].al_wr) + shot * A.psi_stride[a];
    const int64_t h = A.halo, ny = A.n[1], px = A.px, L = X.C.m - 1;
    if (a == 0) {
        X.ls = 1;
        X.ob0 = L + (NDIM == 3 ? h * j : 0);
        X.oa0 = L + (NDIM == 3 ? (h + 1) * j : 0);
        X.sb = NDIM == 3 ? h * ny : h;
        X.sa = NDIM == 3 ? (h + 1) * ny : h + 1;
    } else {
        X.ls = px;
        X.ob0 = i + px * L;
        X.oa0 = X.ob0;
        X.sb = px * h;
        X.sa = px * (h + 1);
    }
    return X;
}

template <typename T>
__device__ __forceinline__ void adj_axis_vals(const AdjAxis<T>& X, int64_t k, AdjVals<T>& V) {
    adj_vals(X.C, X.be_rd, X.al_rd, X.ob0 + k * X.sb, X.oa0 + k * X.sa, X.ls, V);
}

// The march axis: its slab layers are whole planes, layer L of a thread's
// column at base + L * ms (ms: march stride), as its field nodes.
template <typename T>
__device__ __forceinline__ void adj_march_vals(const AdjointArgs<T>& A, const AdjCoef<T>& C,
                                               int a, int shot, int64_t base, int64_t ms,
                                               AdjVals<T>& V) {
    const int s = C.side;
    const T* be = (s ? A.side[a][1].be_rd : A.side[a][0].be_rd) + shot * A.xi_stride[a];
    const T* al = (s ? A.side[a][1].al_rd : A.side[a][0].al_rd) + shot * A.psi_stride[a];
    const int64_t o = base + (int64_t)(C.m - 1) * ms;
    adj_vals(C, be, al, o, o, ms, V);
}

// One node's adjoint update.  (i, j, k) interior (2D: j is the march index
// and k = 0), lin its element offset in the shot's field, km its march
// index; fc, lc: fact and lam^{s+1} at the node; um/up: f lam^{s+1} one
// node below/above along each axis; l2, qv, gv: lam^{s+2}, q^s and the
// gradient at the node; X, Y, Z: CPML data of the x, y (3D) and march
// axes with their slab values xv, yv, zv.
template <typename T, int NDIM>
__device__ __forceinline__ void adjoint_core(const AdjointArgs<T>& A, const AdjAxis<T>& X,
                                             const AdjAxis<T>& Y, const AdjCoef<T>& Z,
                                             const AdjVals<T>& xv, const AdjVals<T>& yv,
                                             const AdjVals<T>& zv, int64_t km, int shot,
                                             int64_t base, int64_t ms, int64_t lin, int64_t fs,
                                             T fc, T lc, const T* um, const T* up, T l2, T qv,
                                             T gv) {
    const T uc = fc * lc;
    T acc = T(2) * lc - l2;
#pragma unroll
    for (int a = 0; a < NDIM; ++a) acc += (up[a] - T(2) * uc + um[a]) * A.inv_d2[a];
    const int M = NDIM - 1;
    const int64_t xo = X.oa0 + km * X.sa;
    if (X.C.ext) adj_term(X.C, xv, um[0], uc, up[0], A.inv_d[0], A.inv_d2[0], X.al_wr, xo, X.ls, acc);
    int64_t yo = 0;
    if (NDIM == 3) {
        yo = Y.oa0 + km * Y.sa;
        if (Y.C.ext)
            adj_term(Y.C, yv, um[1], uc, up[1], A.inv_d[1], A.inv_d2[1], Y.al_wr, yo, Y.ls, acc);
    }
    const int64_t zo = base + (int64_t)(Z.m - 1) * ms;
    T* zal = nullptr;
    T* zbe = nullptr;
    if (Z.ext) {
        const int s = Z.side;
        zal = (s ? A.side[M][1].al_wr : A.side[M][0].al_wr) + shot * A.psi_stride[M];
        zbe = (s ? A.side[M][1].be_wr : A.side[M][0].be_wr) + shot * A.xi_stride[M];
        adj_term(Z, zv, um[M], uc, up[M], A.inv_d[M], A.inv_d2[M], zal, zo, ms, acc);
    }
    A.lam0[fs + lin] = acc;
    // beta^s = bK beta^{s+1} + f lam^s on every slab holding the node
    if (X.C.ext && X.C.v1) X.be_wr[X.ob0 + km * X.sb] = X.C.bK1 * xv.b1 + fc * acc;
    if (NDIM == 3 && Y.C.ext && Y.C.v1) Y.be_wr[Y.ob0 + km * Y.sb] = Y.C.bK1 * yv.b1 + fc * acc;
    if (Z.ext && Z.v1) zbe[zo] = Z.bK1 * zv.b1 + fc * acc;
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
// of node k+1, its CPML slab values included, is issued before node k is
// computed.  Logical blocks: x tile fastest, then row/chunk group, chunk
// (3D), shot, renumbered XCD-aware.
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
    const int M = NDIM - 1;
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
    const AdjAxis<T> X = adj_axis<T, NDIM>(A, 0, shot, i, j);
    AdjAxis<T> Y;
    if (NDIM == 3) Y = adj_axis<T, NDIM>(A, 1, shot, i, j);
    else Y.C.ext = false;
    AdjCoef<T> Z = adj_coef<T>(A, M, k0);
    AdjVals<T> xv = {}, yv = {}, zv = {};
    if (X.C.ext) adj_axis_vals(X, k0, xv);
    if (NDIM == 3 && Y.C.ext) adj_axis_vals(Y, k0, yv);
    if (Z.ext) adj_march_vals(A, Z, M, shot, base, ms, zv);
    for (int64_t k = k0; k < k1; ++k, lin += ms) {
        // next node's loads (planes stay inside the field at the chunk end)
        const int64_t kk = k + 2 < mext ? k + 2 : mext - 1;
        const T lpp = l1[base + kk * ms], fpp = f[base + kk * ms];
        AdjNear<T> nn;
        adj_load_near<T, NDIM>(l1, f, l2, qq, gg, lin + ms, px, nn);
        const AdjCoef<T> Zn = adj_coef<T>(A, M, k + 1);
        AdjVals<T> xn = {}, yn = {}, zn = {};
        if (X.C.ext) adj_axis_vals(X, k + 1, xn);
        if (NDIM == 3 && Y.C.ext) adj_axis_vals(Y, k + 1, yn);
        if (Zn.ext) adj_march_vals(A, Zn, M, shot, base, ms, zn);
        // keep the next node's loads ahead of this node's arithmetic
        __builtin_amdgcn_sched_barrier(0);
        T um[3], up[3];
        um[0] = nr.fxm * nr.lxm;
        up[0] = nr.fxp * nr.lxp;
        if (NDIM == 3) {
            um[1] = nr.fym * nr.lym;
            up[1] = nr.fyp * nr.lyp;
        }
        um[M] = u_m;
        up[M] = fp * lp;
        adjoint_core<T, NDIM>(A, X, Y, Z, xv, yv, zv, k, shot, base, ms, lin, fs, fc, lc, um, up,
                              nr.l2, nr.qv, nr.gv);
        u_m = fc * lc;
        lc = lp;
        fc = fp;
        lp = lpp;
        fp = fpp;
        nr = nn;
        Z = Zn;
        xv = xn;
        yv = yn;
        zv = zn;
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

// Fold the per-shot partial gradients of the last reverse steps into the
// FP64 total: total += sum over shots (in shot order), partials = 0.  The
// partial sums then run over at most one checkpoint segment in the plan's
// precision, and the long sum over time and shots is FP64.
template <typename T>
__global__ __launch_bounds__(256) void fold_gradient(T* grad, double* total, int64_t field_stride,
                                                     int nshots) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= field_stride) return;
    double acc = total[t];
    for (int s = 0; s < nshots; ++s) {
        T& g = grad[(int64_t)s * field_stride + t];
        acc += (double)g;
        g = T(0);
    }
    total[t] = acc;
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
