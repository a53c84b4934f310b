// awcpml.hip -- plan management and launches behind include/awcpml.h.
//
// Device layout of a plan (per shot; shots are stacked back to back):
//   fields pold/pcur/pnew: px x ny x nz, px = nx rounded up to 128 bytes,
//     x fastest (the reference's column-major order, rows aligned);
//   fact: one px x ny x nz array shared by all shots;
//   psi_x: (halo+1) x ny x nz, psi_y: px x (halo+1) x nz,
//   psi_z: px x ny x (halo+1), two copies each (step n-1 / step n);
//   xi_*: the same with halo layers, one copy (updated in place);
//   CPML coefficient vectors, wavelets (nt per source), traces (nt per
//     receiver), all in the plan's precision.
// The three field buffers rotate (pold, pcur, pnew) <- (pcur, pnew, pold)
// after every step, as the reference's time loop does
// (src/shared/acoustic_2D_xPU.jl:332); Dirichlet nodes are never written
// by the stencil, so a source on such a node accumulates per buffer exactly
// as in the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "adjoint_kernels.h"
#include "awcpml.h"
#include "launch.h"
#include "stencil_kernels.h"

using awcpml::StencilArgs;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                              \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess)                                                      \
            return fail(e_ == hipErrorOutOfMemory ? AWCPML_ERR_NOMEM : AWCPML_ERR_HIP, \
                        "%s failed: %s", #expr, hipGetErrorString(e_));            \
    } while (0)

int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

// Which physical buffer plays which role: the field rotation (p[rot] =
// pold ...), the psi parity, and the spare buffers of two-step launches,
// which trade places with the buffers they replace.  A captured graph
// holds physical addresses, so it is replayed only from the layout it was
// captured in.
struct BufLayout {
    void* p[4];           // p[0..2], spare
    void* psi[3][2][3];   // psi[a][s][0..1], spare
    void* xi[3][2][2];    // xi[a][s], spare
    int rot, par;
    bool operator==(const BufLayout& o) const {
        if (rot != o.rot || par != o.par) return false;
        for (int b = 0; b < 4; ++b)
            if (p[b] != o.p[b]) return false;
        for (int a = 0; a < 3; ++a)
            for (int s = 0; s < 2; ++s) {
                for (int c = 0; c < 3; ++c)
                    if (psi[a][s][c] != o.psi[a][s][c]) return false;
                for (int c = 0; c < 2; ++c)
                    if (xi[a][s][c] != o.xi[a][s][c]) return false;
            }
        return true;
    }
};

struct GraphEntry {
    int64_t it0, n;
    hipGraphExec_t exec;
    BufLayout before, after;            // plans only
    int roles_before, roles_after;      // batches only
};

void drop_graphs(std::vector<GraphEntry>& g) {
    for (auto& e : g) (void)hipGraphExecDestroy(e.exec);
    g.clear();
}

}  // namespace

struct awcpml_plan {
    int ndim = 0, dtype = 0, nshots = 0, halo = 0;
    int64_t n[3] = {1, 1, 1};
    int64_t px = 0;
    size_t esize = 0;
    int64_t field_elems = 0;
    void* p[3] = {nullptr, nullptr, nullptr};
    void* fact = nullptr;
    // fact at the aligned two-step tiles' edge nodes (x = 64*t-1, 64*t+64,
    // clamped into the interior), rows of a tile edge contiguous:
    // 3D [z][t][side][y], 2D [y][t][side]
    void* fact_edge = nullptr;
    int64_t fe_elems = 0;
    void* psi[3][2][2] = {};
    void* xi[3][2] = {};
    int64_t psi_elems[3] = {0, 0, 0};
    int64_t xi_elems[3] = {0, 0, 0};
    void* coef[3][8] = {};
    double inv_d[3] = {0, 0, 0}, inv_d2[3] = {0, 0, 0};
    int active[3][2] = {};
    bool model_set = false;
    bool cpml_set[3] = {false, false, false};

    // sources / receivers (host staging, per shot)
    std::vector<std::vector<int64_t>> src_lin, rec_lin;
    std::vector<std::vector<double>> src_wav;  // nt per source
    std::vector<bool> src_set, rec_set;
    int64_t nt = -1;
    bool io_dirty = true;
    void* d_src_lin = nullptr;
    void* d_src_wav = nullptr;
    void* d_src_start = nullptr;
    void* d_src_seq = nullptr;
    void* d_rec_lin = nullptr;
    void* d_rec_start = nullptr;
    void* d_traces = nullptr;
    int64_t n_src_total = 0, n_rec_total = 0;
    bool has_io = false;

    int64_t it = 0;   // steps done
    int rot = 0;      // p[rot] = pold, p[rot+1] = pcur, p[rot+2] = pnew
    int par = 0;      // psi[..][par] = previous step

    hipStream_t stream = nullptr;
    // captured step sequences (hipGraph), keyed by first step and length
    bool use_graph = false;
    std::vector<GraphEntry> graphs;
    int block_y = 0, march_len = 0;
    int prefetch = 2;  // software-pipeline depth of the march (1..3)
    int shot_group = 1;  // shots per block sharing rows (divides block_y)
    int device = 0;

    // adjoint: forward-state checkpoints every ck_every steps (state after
    // c*ck_every steps: the three field buffers, psi of the current parity,
    // xi), valid until the model, CPML or sources/receivers change
    void* ck = nullptr;
    size_t ck_bytes = 0;  // per checkpoint
    int64_t ck_every = 0, ck_count = 0;
    size_t ck_alloc = 0;
    bool ck_valid = false;
    std::vector<int> ck_rot, ck_par;
    // reverse steps whose gradient terms are summed in the plan's precision
    // before they are folded into the FP64 total (0: one checkpoint
    // segment; -1: never, the whole sum stays in the plan's precision)
    int64_t grad_fold = 0;
    int adj_group = 0;      // shots per thread of the reverse step, 0: automatic
    void* q_cur = nullptr;  // d(pnew)/d(fact) output of the next stencil launch
    void* q_next = nullptr; // and of its second step (two-step launches)

    // two time steps per launch (3D): step t+2's field, psi and xi go to
    // spare buffers, which then trade places with the buffers they replace
    int time_block = 1;
    int tb_rows = 0, tb_march = 0;   // waves per block, chunk length (0: default)
    int tb_prefetch = 0;             // planes of loads in flight (0: default)
    int tb_align = 1;                // 3D: aligned 64-lane tiles where no CPML slab reaches (FP32 default)
    void* p_spare = nullptr;
    void* psi_spare[3][2] = {};
    void* xi_spare[3][2] = {};
    BufLayout home = {};             // the layout a reset restores
    bool src_on_boundary = false;    // some source lies on a Dirichlet node
    int src_per_shot_max = 0;
    bool boundary_dirty = false;     // such a source was injected since the reset
    bool wide = false;               // 64-bit offsets forced (awcpml_set_wide_offsets)

    // configuration generation: bumped whenever something a batch copies
    // into its argument tables changes (source/receiver buffers and nt,
    // inverse spacings, active CPML sides, launch shape, spare buffers), so
    // that a batch built earlier notices and rebuilds its tables
    uint64_t gen = 0;
};

namespace {

int64_t slab_elems(const awcpml_plan* P, int axis, int64_t layers) {
    // layouts documented at the top of the file
    if (axis == 0) return layers * P->n[1] * P->n[2];
    if (axis == 1) return P->px * layers * P->n[2];
    return P->px * P->n[1] * layers;
}

BufLayout get_layout(const awcpml_plan* P) {
    BufLayout L = {};
    for (int b = 0; b < 3; ++b) L.p[b] = P->p[b];
    L.p[3] = P->p_spare;
    for (int a = 0; a < 3; ++a)
        for (int s = 0; s < 2; ++s) {
            L.psi[a][s][0] = P->psi[a][s][0];
            L.psi[a][s][1] = P->psi[a][s][1];
            L.psi[a][s][2] = P->psi_spare[a][s];
            L.xi[a][s][0] = P->xi[a][s];
            L.xi[a][s][1] = P->xi_spare[a][s];
        }
    L.rot = P->rot;
    L.par = P->par;
    return L;
}

void set_layout(awcpml_plan* P, const BufLayout& L) {
    for (int b = 0; b < 3; ++b) P->p[b] = L.p[b];
    P->p_spare = L.p[3];
    for (int a = 0; a < 3; ++a)
        for (int s = 0; s < 2; ++s) {
            P->psi[a][s][0] = L.psi[a][s][0];
            P->psi[a][s][1] = L.psi[a][s][1];
            P->psi_spare[a][s] = L.psi[a][s][2];
            P->xi[a][s] = L.xi[a][s][0];
            P->xi_spare[a][s] = L.xi[a][s][1];
        }
    P->rot = L.rot;
    P->par = L.par;
}

// The plan's buffers play their home roles (no two-step launch swapped
// them since the last reset); batches precompute arguments for that layout.
bool at_home(const awcpml_plan* P) {
    BufLayout now = get_layout(P);
    now.rot = P->home.rot;
    now.par = P->home.par;
    return now == P->home;
}

int dalloc(void** ptr, size_t bytes) {
    if (bytes == 0) bytes = 256;
    HIP_TRY(hipMalloc(ptr, bytes));
    return AWCPML_OK;
}

void dfree(void*& ptr) {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
}

// host FP64 values -> device buffer of the plan's precision
int upload(awcpml_plan* P, void* dst, const double* src, size_t count) {
    if (count == 0) return AWCPML_OK;
    if (P->dtype == AWCPML_F64) {
        HIP_TRY(hipMemcpyAsync(dst, src, count * 8, hipMemcpyHostToDevice, P->stream));
    } else {
        std::vector<float> tmp(count);
        for (size_t i = 0; i < count; ++i) tmp[i] = (float)src[i];
        HIP_TRY(hipMemcpyAsync(dst, tmp.data(), count * 4, hipMemcpyHostToDevice, P->stream));
        HIP_TRY(hipStreamSynchronize(P->stream));
    }
    return AWCPML_OK;
}

int check_plan(const awcpml_plan* P) {
    if (!P) return fail(AWCPML_ERR_ARG, "plan is NULL");
    return AWCPML_OK;
}

int default_tiling(awcpml_plan* P) {
    if (P->block_y <= 0) P->block_y = 4;
    if (P->march_len <= 0) {
        if (P->ndim == 3) {
            // enough blocks to fill 256 CUs several times over
            const int64_t tiles = round_up(P->n[0], awcpml::kTileX) / awcpml::kTileX *
                                  ((P->n[1] - 2 + P->block_y - 1) / P->block_y) * P->nshots;
            int64_t interior = P->n[2] - 2;
            int64_t want_chunks = std::max<int64_t>(1, (4096 + tiles - 1) / tiles);
            int64_t len = (interior + want_chunks - 1) / want_chunks;
            P->march_len = (int)std::max<int64_t>(16, std::min<int64_t>(len, 128));
        } else {
            int64_t interior = P->n[1] - 2;
            int64_t len = std::max<int64_t>(8, std::min<int64_t>(64, interior / 16));
            P->march_len = (int)len;
        }
    }
    return AWCPML_OK;
}

// Kernel arguments for rotation state (rot, par): p[rot] = pold,
// p[rot+1] = pcur, p[rot+2] = pnew; psi[par] = previous step.
template <typename T>
StencilArgs<T> make_args(const awcpml_plan* P, int rot, int par) {
    StencilArgs<T> A;
    memset(&A, 0, sizeof A);
    A.nx = P->n[0];
    A.ny = P->n[1];
    A.nz = P->n[2];
    A.px = P->px;
    A.halo = P->halo;
    A.field_stride = P->field_elems;
    A.pold = (const T*)P->p[rot % 3];
    A.pcur = (const T*)P->p[(rot + 1) % 3];
    A.pnew = (T*)P->p[(rot + 2) % 3];
    A.fact = (const T*)P->fact;
    A.qout = (T*)P->q_cur;
    for (int a = 0; a < 3; ++a) {
        A.inv_d[a] = (T)P->inv_d[a];
        A.inv_d2[a] = (T)P->inv_d2[a];
        A.psi_stride[a] = P->psi_elems[a];
        A.xi_stride[a] = P->xi_elems[a];
        for (int s = 0; s < 2; ++s) {
            A.active[a][s] = (a < P->ndim) ? P->active[a][s] : 0;
            awcpml::CpmlSide<T>& S = A.side[a][s];
            S.psi_rd = (const T*)P->psi[a][s][par];
            S.psi_wr = (T*)P->psi[a][s][par ^ 1];
            S.xi = (T*)P->xi[a][s];
            S.a = (const T*)P->coef[a][0 + s];
            S.bK = (const T*)P->coef[a][2 + s];
            S.ah = (const T*)P->coef[a][4 + s];
            S.bKh = (const T*)P->coef[a][6 + s];
        }
    }
    const int G = P->shot_group;
    const int rows = P->block_y / G;  // rows (2D: chunks) per block
    A.march_len = P->march_len;
    A.nshots = P->nshots;
    A.shot_group = G;
    A.nsgroups = (P->nshots + G - 1) / G;
    A.tiles_x = (int)(round_up(P->n[0], awcpml::kTileX) / awcpml::kTileX);
    {
        // lane-shuffle x slabs: each active side's psi layers 0..halo in one
        // x tile, and the two sides' lane ranges disjoint
        const int64_t nx = P->n[0], h = P->halo, t = awcpml::kTileX;
        bool fast = h > 0 && nx >= 2 * h + 3;
        if (P->active[0][0]) fast = fast && h < t;
        if (P->active[0][1]) fast = fast && (nx - h - 2) / t == (nx - 2) / t;
        A.x_fast = fast ? 1 : 0;
    }
    if (P->ndim == 3) {
        A.tiles_y = (int)((P->n[1] - 2 + rows - 1) / rows);
        A.nchunks = (int)((P->n[2] - 2 + P->march_len - 1) / P->march_len);
        A.logical_blocks = A.nsgroups * A.tiles_x * A.tiles_y * A.nchunks;
    } else {
        A.nchunks = (int)((P->n[1] - 2 + P->march_len - 1) / P->march_len);
        A.tiles_y = (A.nchunks + rows - 1) / rows;
        A.logical_blocks = A.nsgroups * A.tiles_x * A.tiles_y;
    }
    A.grid_blocks = (int)round_up(A.logical_blocks, awcpml::kNumXcd);
    return A;
}

template <typename T>
int launch_stencil_t(awcpml_plan* P) {
    StencilArgs<T> A = make_args<T>(P, P->rot, P->par);
    awcpml::LaunchShape L;
    L.grid = dim3(A.grid_blocks);
    L.block = dim3(awcpml::kTileX, P->block_y);
    // march-axis CPML coefficients staged in LDS: 8 vectors of halo+1
    L.lds = (size_t)8 * (P->halo + 1) * sizeof(T);
    L.stream = P->stream;
    // 32-bit byte offsets when one shot's field fits in 4 GiB (the slabs
    // are smaller than a field)
    const bool off64 = P->wide || (uint64_t)P->field_elems * sizeof(T) >= (1ull << 32);
    if (P->ndim == 3) awcpml::launch_plan_stencil<T, 3>(A, L, P->prefetch, off64);
    else awcpml::launch_plan_stencil<T, 2>(A, L, P->prefetch, off64);
    HIP_TRY(hipGetLastError());
    return AWCPML_OK;
}

int launch_stencil(awcpml_plan* P) {
    return P->dtype == AWCPML_F64 ? launch_stencil_t<double>(P) : launch_stencil_t<float>(P);
}

template <typename T>
awcpml::IoArgs<T> make_io(const awcpml_plan* P, int rot) {
    awcpml::IoArgs<T> A;
    A.pnew = (T*)P->p[(rot + 2) % 3];
    A.field_stride = P->field_elems;
    A.src_lin = (const int64_t*)P->d_src_lin;
    A.src_wav = (const T*)P->d_src_wav;
    A.src_start = (const int*)P->d_src_start;
    A.src_seq = (const int*)P->d_src_seq;
    A.rec_lin = (const int64_t*)P->d_rec_lin;
    A.traces = (T*)P->d_traces;
    A.rec_start = (const int*)P->d_rec_start;
    A.nt = P->nt;
    A.inject = 1;
    return A;
}

template <typename T>
int launch_io_t(awcpml_plan* P, int64_t it0) {
    awcpml::sources_receivers<T><<<dim3(P->nshots), dim3(256), 0, P->stream>>>(make_io<T>(P, P->rot), it0);
    HIP_TRY(hipGetLastError());
    return AWCPML_OK;
}

int launch_io(awcpml_plan* P, int64_t it0) {
    return P->dtype == AWCPML_F64 ? launch_io_t<double>(P, it0) : launch_io_t<float>(P, it0);
}

// Upload the per-shot source/receiver lists as CSR arrays.
int sync_io(awcpml_plan* P) {
    if (!P->io_dirty) return AWCPML_OK;
    for (int s = 0; s < P->nshots; ++s) {
        if (!P->src_set[s]) return fail(AWCPML_ERR_STATE, "sources of shot %d not set", s);
        if (!P->rec_set[s]) return fail(AWCPML_ERR_STATE, "receivers of shot %d not set", s);
    }
    std::vector<int64_t> slin, rlin;
    std::vector<int> sstart(1, 0), rstart(1, 0), seq;
    std::vector<double> wav;
    for (int s = 0; s < P->nshots; ++s) {
        const auto& L = P->src_lin[s];
        slin.insert(slin.end(), L.begin(), L.end());
        wav.insert(wav.end(), P->src_wav[s].begin(), P->src_wav[s].end());
        sstart.push_back((int)slin.size());
        std::vector<int64_t> sorted(L);
        std::sort(sorted.begin(), sorted.end());
        seq.push_back(std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end() ? 1 : 0);
        rlin.insert(rlin.end(), P->rec_lin[s].begin(), P->rec_lin[s].end());
        rstart.push_back((int)rlin.size());
    }
    drop_graphs(P->graphs);  // they reference the buffers below
    ++P->gen;                // and so do the argument tables of batches
    dfree(P->d_src_lin);
    dfree(P->d_src_wav);
    dfree(P->d_src_start);
    dfree(P->d_src_seq);
    dfree(P->d_rec_lin);
    dfree(P->d_rec_start);
    dfree(P->d_traces);
    int st;
    if ((st = dalloc(&P->d_src_lin, slin.size() * 8))) return st;
    if ((st = dalloc(&P->d_src_wav, wav.size() * P->esize))) return st;
    if ((st = dalloc(&P->d_src_start, sstart.size() * 4))) return st;
    if ((st = dalloc(&P->d_src_seq, seq.size() * 4))) return st;
    if ((st = dalloc(&P->d_rec_lin, rlin.size() * 8))) return st;
    if ((st = dalloc(&P->d_rec_start, rstart.size() * 4))) return st;
    if ((st = dalloc(&P->d_traces, (size_t)rlin.size() * (size_t)std::max<int64_t>(P->nt, 1) * P->esize)))
        return st;
    if (!slin.empty())
        HIP_TRY(hipMemcpyAsync(P->d_src_lin, slin.data(), slin.size() * 8, hipMemcpyHostToDevice, P->stream));
    if ((st = upload(P, P->d_src_wav, wav.data(), wav.size()))) return st;
    HIP_TRY(hipMemcpyAsync(P->d_src_start, sstart.data(), sstart.size() * 4, hipMemcpyHostToDevice, P->stream));
    HIP_TRY(hipMemcpyAsync(P->d_src_seq, seq.data(), seq.size() * 4, hipMemcpyHostToDevice, P->stream));
    if (!rlin.empty())
        HIP_TRY(hipMemcpyAsync(P->d_rec_lin, rlin.data(), rlin.size() * 8, hipMemcpyHostToDevice, P->stream));
    HIP_TRY(hipMemcpyAsync(P->d_rec_start, rstart.data(), rstart.size() * 4, hipMemcpyHostToDevice, P->stream));
    if (!rlin.empty() && P->nt > 0)
        HIP_TRY(hipMemsetAsync(P->d_traces, 0, (size_t)rlin.size() * P->nt * P->esize, P->stream));
    HIP_TRY(hipStreamSynchronize(P->stream));
    P->n_src_total = (int64_t)slin.size();
    P->n_rec_total = (int64_t)rlin.size();
    P->src_on_boundary = false;
    P->src_per_shot_max = 0;
    for (int s = 0; s < P->nshots; ++s)
        P->src_per_shot_max = std::max(P->src_per_shot_max, sstart[s + 1] - sstart[s]);
    for (int64_t lin : slin) {
        const int64_t i = lin % P->px, j = (lin / P->px) % P->n[1], k = lin / (P->px * P->n[1]);
        bool b = i == 0 || i == P->n[0] - 1 || j == 0 || j == P->n[1] - 1;
        if (P->ndim == 3) b = b || k == 0 || k == P->n[2] - 1;
        P->src_on_boundary = P->src_on_boundary || b;
    }
    P->has_io = (P->n_src_total + P->n_rec_total) > 0;
    P->io_dirty = false;
    return AWCPML_OK;
}

int ready(awcpml_plan* P) {
    if (!P->model_set) return fail(AWCPML_ERR_STATE, "model not set (awcpml_set_model)");
    for (int a = 0; a < P->ndim; ++a)
        if (P->halo > 0 && !P->cpml_set[a])
            return fail(AWCPML_ERR_STATE, "CPML coefficients of axis %d not set", a);
    int st = sync_io(P);
    if (st) return st;
    return default_tiling(P);
}

// one time step with explicit 0-based index it0, then rotate
int do_step(awcpml_plan* P, int64_t it0) {
    int st = launch_stencil(P);
    if (st) return st;
    if (P->has_io && (st = launch_io(P, it0))) return st;
    if (P->has_io && P->src_on_boundary) P->boundary_dirty = true;
    P->rot = (P->rot + 1) % 3;
    P->par ^= 1;
    return AWCPML_OK;
}

// Two-step launches can run: spare buffers allocated, every field buffer
// zero on the Dirichlet nodes, at most kMaxSrc2 sources per shot (staged
// in LDS), 32-bit byte offsets, and a d(pnew)/d(fact) output (if any) for
// both steps.  two_ok: they also were requested for the plan.
bool two_applies(const awcpml_plan* P) {
    return P->p_spare && (!P->q_cur || P->q_next) && !P->src_on_boundary &&
           P->src_per_shot_max <= awcpml::kMaxSrc2 && !P->boundary_dirty && !P->wide &&
           (uint64_t)P->field_elems * P->esize < (1ull << 32);
}

bool two_ok(const awcpml_plan* P) { return P->time_block == 2 && two_applies(P); }

int two_rows_default(int ndim, int dtype) { return ndim == 2 ? 4 : (dtype == AWCPML_F64 ? 8 : 16); }

// The plan's chunk length of two-step launches with W waves per block.
int two_march(const awcpml_plan* P, int W) {
    if (P->tb_march > 0) return P->tb_march;
    if (P->ndim == 2) return (int)std::max<int64_t>(8, std::min<int64_t>(64, (P->n[1] - 2) / 16));
    // enough blocks to fill 256 CUs several times over
    const int64_t tiles = ((P->n[0] - 2 + awcpml::kOwnX - 1) / awcpml::kOwnX) *
                          ((P->n[1] - 2 + W - 3) / (W - 2)) * P->nshots;
    const int64_t interior = P->n[2] - 2;
    const int64_t want = std::max<int64_t>(1, (2048 + tiles - 1) / tiles);
    return (int)std::max<int64_t>(16, std::min<int64_t>((interior + want - 1) / want, 128));
}

void two_tiling(const awcpml_plan* P, int* rows, int* march, int* prefetch) {
    const int W = P->tb_rows > 0 ? P->tb_rows : two_rows_default(P->ndim, P->dtype);
    *rows = W;
    *march = two_march(P, W);
    *prefetch = P->tb_prefetch > 0 ? P->tb_prefetch : 2;
}

// March chunks of a two-step launch (chunk_start[0..count]), or 0 for
// uniform chunks of ml planes.  A chunk reads one plane beyond each end
// (step t+1 there), so the CPML layers 1..halo of each absorbing side of
// the march axis, with the plane next to them, form chunks of their own:
// [1, halo+2) and [n-halo-2, n-1).  The interior chunks between then take
// the march without slab terms.  Each part is cut into near-equal chunks of
// at most ml planes.
int two_chunk_list(const awcpml_plan* P, int ml, int* starts) {
    const int a = P->ndim - 1;
    const int64_t n = P->n[a], h = P->halo;
    if (h == 0 || (!P->active[a][0] && !P->active[a][1])) return 0;
    const int64_t lo = P->active[a][0] ? h + 2 : 1;
    const int64_t hi = P->active[a][1] ? n - h - 2 : n - 1;
    if (lo >= hi) return 0;  // the layers fill the axis: uniform chunks
    const int64_t bounds[4] = {1, lo, hi, n - 1};
    int count = 0;
    starts[0] = 1;
    for (int part = 0; part < 3; ++part) {
        const int64_t b0 = bounds[part], b1 = bounds[part + 1];
        if (b1 <= b0) continue;
        const int64_t pieces = (b1 - b0 + ml - 1) / ml;
        if (count + pieces > awcpml::kMaxChunks2) return 0;
        for (int64_t q = 1; q <= pieces; ++q) starts[++count] = (int)(b0 + (b1 - b0) * q / pieces);
    }
    return count;
}

// Aligned 64-lane tiles need the plain stencil at the nodes just outside
// each tile (x = 64*tx-1 and 64*tx+64): none of them may be an interior
// node of an active x slab.
bool two_align_ok(const awcpml_plan* P) {
    const int64_t nx = P->n[0], h = P->halo;
    const int64_t tiles = (nx + awcpml::kTileX - 1) / awcpml::kTileX;
    for (int64_t t = 0; t < tiles; ++t)
        for (int64_t x : {t * awcpml::kTileX - 1, t * awcpml::kTileX + awcpml::kTileX}) {
            if (x < 1 || x > nx - 2) continue;
            if ((P->active[0][0] && x <= h) || (P->active[0][1] && x >= nx - h - 1)) return false;
        }
    return true;
}

// Arguments of a two-step launch from the current layout: step t+1 reads
// p[rot] (pold), p[rot+1] (pcur), psi[par] and xi and writes p[rot+2];
// step t+2 writes the spare field, psi and xi buffers.
template <typename T>
awcpml::Stencil2Args<T> make_args2(const awcpml_plan* P, int64_t it0, int W, int ml) {
    awcpml::Stencil2Args<T> B;
    memset(&B, 0, sizeof B);
    B.s = make_args<T>(P, P->rot, P->par);
    StencilArgs<T>& A = B.s;
    for (int a = 0; a < 3; ++a)
        for (int s = 0; s < 2; ++s) {
            A.side[a][s].psi_wr = (T*)P->psi_spare[a][s];
            B.xi_out[a][s] = (T*)P->xi_spare[a][s];
        }
    A.qout = nullptr;
    A.shot_group = 1;
    A.nsgroups = P->nshots;
    A.tiles_x = (int)((P->n[0] - 2 + awcpml::kOwnX - 1) / awcpml::kOwnX);
    A.march_len = ml;
    B.nchunk_list = two_chunk_list(P, ml, B.chunk_start);
    B.tiles62 = A.tiles_x;
    B.tiles64 = (int)((P->n[0] + awcpml::kTileX - 1) / awcpml::kTileX);
    B.fact_edge = (const T*)P->fact_edge;
    B.align = P->tb_align && two_align_ok(P);
    if (B.align) A.tiles_x = std::max(B.tiles62, B.tiles64);
    if (P->ndim == 3) {
        A.tiles_y = (int)((P->n[1] - 2 + W - 3) / (W - 2));
        A.nchunks = B.nchunk_list > 0 ? B.nchunk_list : (int)((P->n[2] - 2 + ml - 1) / ml);
        A.logical_blocks = A.tiles_x * A.tiles_y * A.nchunks * P->nshots;
    } else {
        // a block's W waves march W consecutive chunks of rows
        A.nchunks = B.nchunk_list > 0 ? B.nchunk_list : (int)((P->n[1] - 2 + ml - 1) / ml);
        A.tiles_y = (A.nchunks + W - 1) / W;
        A.logical_blocks = A.tiles_x * A.tiles_y * P->nshots;
    }
    A.grid_blocks = (int)round_up(A.logical_blocks, awcpml::kNumXcd);
    B.p1 = (T*)P->p[(P->rot + 2) % 3];
    B.p2 = (T*)P->p_spare;
    B.q1 = (T*)P->q_cur;
    B.q2 = (T*)P->q_next;
    B.src_lin = (const int64_t*)P->d_src_lin;
    B.src_wav = (const T*)P->d_src_wav;
    B.src_start = (const int*)P->d_src_start;
    B.nt = P->nt;
    B.it0 = it0;
    return B;
}

template <typename T>
int launch_stencil2_t(awcpml_plan* P, int64_t it0) {
    int W, ml, pf;
    two_tiling(P, &W, &ml, &pf);
    const awcpml::Stencil2Args<T> B = make_args2<T>(P, it0, W, ml);
    awcpml::LaunchShape L;
    L.grid = dim3(B.s.grid_blocks);
    L.block = dim3(awcpml::kTileX, W);
    L.lds = (size_t)8 * (P->halo + 1) * sizeof(T);
    L.stream = P->stream;
    if (P->ndim == 3) awcpml::launch_plan_stencil2<T, 3>(B, L, W, pf);
    else awcpml::launch_plan_stencil2<T, 2>(B, L, W, pf);
    HIP_TRY(hipGetLastError());
    return AWCPML_OK;
}

// After a two-step launch: the spares hold step t+2 and take the places of
// p[rot] (the reference's pnew of step t+2), psi[par] and xi.
void two_swap(awcpml_plan* P) {
    std::swap(P->p[P->rot % 3], P->p_spare);
    for (int a = 0; a < 3; ++a)
        for (int s = 0; s < 2; ++s) {
            std::swap(P->psi[a][s][P->par], P->psi_spare[a][s]);
            std::swap(P->xi[a][s], P->xi_spare[a][s]);
        }
    P->rot = (P->rot + 2) % 3;
}

template <typename T>
int io_two_t(awcpml_plan* P, int64_t it0, int rot, bool inject) {
    awcpml::IoArgs<T> A = make_io<T>(P, rot);
    A.inject = inject ? 1 : 0;
    awcpml::sources_receivers<T><<<dim3(P->nshots), dim3(256), 0, P->stream>>>(A, it0);
    HIP_TRY(hipGetLastError());
    return AWCPML_OK;
}

// Steps it0 and it0+1 in one stencil launch: the kernel injects step
// it0's sources; the receivers of step it0 are then read from p^{t+1},
// and step it0+1's sources and receivers handled on p^{t+2} as usual.
int do_step2(awcpml_plan* P, int64_t it0) {
    const bool f64 = P->dtype == AWCPML_F64;
    int st = f64 ? launch_stencil2_t<double>(P, it0) : launch_stencil2_t<float>(P, it0);
    if (st) return st;
    if (P->has_io) {
        st = f64 ? io_two_t<double>(P, it0, P->rot, false) : io_two_t<float>(P, it0, P->rot, false);
        if (st) return st;
    }
    const int rot1 = (P->rot + 1) % 3;
    two_swap(P);
    if (P->has_io) {
        st = f64 ? io_two_t<double>(P, it0 + 1, rot1, true) : io_two_t<float>(P, it0 + 1, rot1, true);
        if (st) return st;
    }
    return AWCPML_OK;
}

// Enqueue the steps it0 .. it0+n-1 (pairs in two-step launches where they
// apply), advancing the host's layout state as they are enqueued.
int enqueue_steps(awcpml_plan* P, int64_t it0, int64_t n) {
    const bool two = two_ok(P);
    for (int64_t q = 0; q < n;) {
        int st;
        if (two && n - q >= 2) {
            st = do_step2(P, it0 + q);
            q += 2;
        } else {
            st = do_step(P, it0 + q);
            q += 1;
        }
        if (st) return st;
    }
    return AWCPML_OK;
}

// Capture `n` steps starting at step it0 on `stream` into an executable
// graph.  step(q) enqueues step it0+q.  No memset is recorded.
template <typename StepFn>
int capture_steps(hipStream_t stream, int64_t n, StepFn step, hipGraphExec_t* out) {
    HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    int st = AWCPML_OK;
    for (int64_t q = 0; q < n && !st; ++q) st = step(q);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(stream, &graph);
    if (st) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    if (e != hipSuccess) return fail(AWCPML_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(AWCPML_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
    return AWCPML_OK;
}

// Capture whatever fn() enqueues on `stream` into an executable graph.
template <typename Fn>
int capture_seq(hipStream_t stream, Fn fn, hipGraphExec_t* out) {
    return capture_steps(stream, 1, [&](int64_t) { return fn(); }, out);
}

// The plan's graph for n steps from step it0 and layout L, captured on
// first use.  The host state (layout, boundary flag) after the capture is
// the state after the sequence; it is recorded and the plan restored to L.
int plan_graph(awcpml_plan* P, int64_t n, GraphEntry** out) {
    const BufLayout now = get_layout(P);
    const int64_t it0 = P->it;
    auto& G = P->graphs;
    for (size_t q = 0; q < G.size(); ++q)
        if (G[q].it0 == it0 && G[q].n == n) {
            if (G[q].before == now) {
                *out = &G[q];
                return AWCPML_OK;
            }
            (void)hipGraphExecDestroy(G[q].exec);
            G.erase(G.begin() + q);
            break;
        }
    if (G.size() >= 4) {
        (void)hipGraphExecDestroy(G.front().exec);
        G.erase(G.begin());
    }
    hipGraphExec_t exec;
    int st = capture_seq(P->stream, [&]() { return enqueue_steps(P, it0, n); }, &exec);
    const BufLayout after = get_layout(P);
    set_layout(P, now);
    if (st) return st;
    GraphEntry e = {};
    e.it0 = it0;
    e.n = n;
    e.exec = exec;
    e.before = now;
    e.after = after;
    G.push_back(e);
    *out = &G.back();
    return AWCPML_OK;
}

// Enqueue n steps from the plan's current state (no synchronisation).  With
// graphs enabled the sequence is captured once per (first step, length,
// buffer layout) and replayed afterwards.
int run_steps(awcpml_plan* P, int64_t n) {
    if (n <= 0) return AWCPML_OK;
    if (P->use_graph) {
        GraphEntry* g = nullptr;
        int st = plan_graph(P, n, &g);
        if (st) return st;
        HIP_TRY(hipGraphLaunch(g->exec, P->stream));
        P->it += n;
        set_layout(P, g->after);
        if (P->has_io && P->src_on_boundary) P->boundary_dirty = true;
        return AWCPML_OK;
    }
    int st = enqueue_steps(P, P->it, n);
    if (st) return st;
    P->it += n;
    return AWCPML_OK;
}

int64_t lin_index(const awcpml_plan* P, const int64_t* pos, int count, int item) {
    int64_t i = pos[item] - 1;
    int64_t j = pos[(int64_t)count + item] - 1;
    int64_t k = (P->ndim == 3) ? pos[2 * (int64_t)count + item] - 1 : 0;
    return i + P->px * (j + P->n[1] * k);
}

int check_positions(const awcpml_plan* P, const int64_t* pos, int count, const char* what) {
    for (int r = 0; r < count; ++r)
        for (int a = 0; a < P->ndim; ++a) {
            int64_t v = pos[(int64_t)a * count + r];
            if (v < 1 || v > P->n[a])
                return fail(AWCPML_ERR_RANGE, "At least one %s is not inside the model! (%s %d axis %d index %lld)",
                            what, what, r, a, (long long)v);
        }
    return AWCPML_OK;
}

}  // namespace

// ---------------------------------------------------------------------
// Batches: several plans (different grids, same ndim and precision)
// advanced by one stencil launch and one source/receiver launch per step,
// or per two steps with two-step launches.
// ---------------------------------------------------------------------
namespace {

// Buffer roles of a batch's plans relative to their home layouts: field
// slot b (0..2: the rotation, 3: the spare) holds home buffer p[b], psi
// slot c (0..1: the parities, 2: the spare) home psi[c], xi slot c (1: the
// spare) home xi[c].  A two-step launch permutes them the same way in every
// plan of the batch.
struct Roles {
    int p[4], psi[3], xi[2];
    bool operator==(const Roles& o) const {
        for (int b = 0; b < 4; ++b)
            if (p[b] != o.p[b]) return false;
        for (int c = 0; c < 3; ++c)
            if (psi[c] != o.psi[c]) return false;
        return xi[0] == o.xi[0] && xi[1] == o.xi[1];
    }
};

Roles home_roles() { return Roles{{0, 1, 2, 3}, {0, 1, 2}, {0, 1}}; }

// The roles after a two-step launch of steps it0, it0+1 (two_swap).
Roles roles_after_two(Roles R, int64_t it0) {
    std::swap(R.p[it0 % 3], R.p[3]);
    std::swap(R.psi[it0 % 2], R.psi[2]);
    std::swap(R.xi[0], R.xi[1]);
    return R;
}

// The layout of plan P after `it` steps with buffer roles R.
BufLayout roles_layout(const awcpml_plan* P, const Roles& R, int64_t it) {
    BufLayout L = P->home;
    for (int b = 0; b < 4; ++b) L.p[b] = P->home.p[R.p[b]];
    for (int a = 0; a < 3; ++a)
        for (int s = 0; s < 2; ++s) {
            for (int c = 0; c < 3; ++c) L.psi[a][s][c] = P->home.psi[a][s][R.psi[c]];
            for (int c = 0; c < 2; ++c) L.xi[a][s][c] = P->home.xi[a][s][R.xi[c]];
        }
    L.rot = (int)(it % 3);
    L.par = (int)(it % 2);
    return L;
}

// Device argument tables of every plan for one set of roles.
struct RoleTables {
    Roles r;
    void* d_args = nullptr;   // 6 (step mod 6) x nplans single-step arguments
    void* d_args2 = nullptr;  // 6 (step mod 6) x nplans two-step arguments, or null
    void* d_io = nullptr;     // 3 (rotation) x 2 (inject) x nplans io arguments
};

constexpr int kMaxRoles = 288;  // 4! * 3! * 2!: every permutation of the roles

}  // namespace

struct awcpml_batch {
    std::vector<awcpml_plan*> plans;
    int ndim = 0, dtype = 0, block_y = 4, prefetch = 2, halo_max = 0;
    bool use32 = true, has_io = false;
    hipStream_t stream = nullptr;
    std::vector<RoleTables> roles;   // built on first use, kept until destroy
    int cur = 0;                     // roles[cur]: the plans' roles now
    int time_block = 1, tb_rows = 0, tb_prefetch = 0;
    std::vector<int> block_start, block_start2, shot_start;
    int grid_blocks = 0, grid_blocks2 = 0;
    int64_t it = 0, nt_max = 0, nt_min = 0;
    bool use_graph = false;
    std::vector<GraphEntry> graphs;
    std::vector<uint64_t> gens;      // each plan's configuration generation the tables hold
};

namespace {

void batch_two_shape(const awcpml_batch* B, int* rows, int* prefetch) {
    *rows = B->tb_rows > 0 ? B->tb_rows : two_rows_default(B->ndim, B->dtype);
    *prefetch = B->tb_prefetch > 0 ? B->tb_prefetch : 2;
}

// Two-step batch launches apply (roles tables without two-step arguments
// were built while some plan had no spare buffers).
bool batch_two_ok(const awcpml_batch* B) {
    if (B->time_block != 2 || !B->roles[B->cur].d_args2) return false;
    for (auto* P : B->plans)
        if (!two_applies(P)) return false;
    return true;
}

template <typename T>
int batch_tables_t(awcpml_batch* B, const Roles& R, RoleTables* out) {
    const int np = (int)B->plans.size();
    bool two = true;
    for (auto* P : B->plans) two = two && P->p_spare;
    int W, pf;
    batch_two_shape(B, &W, &pf);
    std::vector<StencilArgs<T>> args((size_t)6 * np);
    std::vector<awcpml::Stencil2Args<T>> args2(two ? (size_t)6 * np : 0);
    std::vector<awcpml::IoArgs<T>> io((size_t)6 * np);
    for (int q = 0; q < np; ++q) {
        awcpml_plan* P = B->plans[q];
        const BufLayout keep = get_layout(P);
        set_layout(P, roles_layout(P, R, 0));
        for (int v = 0; v < 6; ++v) args[(size_t)v * np + q] = make_args<T>(P, v % 3, v % 2);
        if (two) {
            const int ml = two_march(P, W);
            for (int v = 0; v < 6; ++v) {
                P->rot = v % 3;
                P->par = v % 2;
                args2[(size_t)v * np + q] = make_args2<T>(P, 0, W, ml);
            }
        }
        for (int r = 0; r < 3; ++r)
            for (int inj = 0; inj < 2; ++inj) {
                awcpml::IoArgs<T> A = make_io<T>(P, r);
                A.inject = inj;
                io[(size_t)(r * 2 + inj) * np + q] = A;
            }
        set_layout(P, keep);
    }
    if (two) {
        B->block_start2.assign(1, 0);
        for (int q = 0; q < np; ++q) B->block_start2.push_back(B->block_start2.back() + args2[q].s.logical_blocks);
        B->grid_blocks2 = (int)round_up(B->block_start2.back(), awcpml::kNumXcd);
    }
    RoleTables t;
    t.r = R;
    int st;
    if ((st = dalloc(&t.d_args, args.size() * sizeof(StencilArgs<T>)))) return st;
    if (two && (st = dalloc(&t.d_args2, args2.size() * sizeof(awcpml::Stencil2Args<T>)))) {
        dfree(t.d_args);
        return st;
    }
    if ((st = dalloc(&t.d_io, io.size() * sizeof(awcpml::IoArgs<T>)))) {
        dfree(t.d_args);
        dfree(t.d_args2);
        return st;
    }
    *out = t;
    HIP_TRY(hipMemcpyAsync(t.d_args, args.data(), args.size() * sizeof(StencilArgs<T>),
                           hipMemcpyHostToDevice, B->stream));
    if (two)
        HIP_TRY(hipMemcpyAsync(t.d_args2, args2.data(), args2.size() * sizeof(awcpml::Stencil2Args<T>),
                               hipMemcpyHostToDevice, B->stream));
    HIP_TRY(hipMemcpyAsync(t.d_io, io.data(), io.size() * sizeof(awcpml::IoArgs<T>),
                           hipMemcpyHostToDevice, B->stream));
    HIP_TRY(hipStreamSynchronize(B->stream));
    return AWCPML_OK;
}

void free_tables(RoleTables& t) {
    dfree(t.d_args);
    dfree(t.d_args2);
    dfree(t.d_io);
}

// Index of the tables for roles R, built on first use (allocates: call
// outside stream capture).
int batch_roles(awcpml_batch* B, const Roles& R, int* idx) {
    for (size_t q = 0; q < B->roles.size(); ++q)
        if (B->roles[q].r == R) {
            *idx = (int)q;
            return AWCPML_OK;
        }
    if ((int)B->roles.size() >= kMaxRoles) return fail(AWCPML_ERR_STATE, "batch roles table full");
    RoleTables t;
    int st = B->dtype == AWCPML_F64 ? batch_tables_t<double>(B, R, &t) : batch_tables_t<float>(B, R, &t);
    if (st) {
        free_tables(t);
        return st;
    }
    B->roles.push_back(t);
    *idx = (int)B->roles.size() - 1;
    return AWCPML_OK;
}

// Drop every table and graph and rebuild the tables of the current roles
// (the home roles when there are none).
int batch_rebuild(awcpml_batch* B) {
    HIP_TRY(hipStreamSynchronize(B->stream));
    drop_graphs(B->graphs);
    const Roles R = B->roles.empty() ? home_roles() : B->roles[B->cur].r;
    for (auto& t : B->roles) free_tables(t);
    B->roles.clear();
    B->cur = 0;
    B->gens.clear();
    for (auto* P : B->plans) B->gens.push_back(P->gen);
    return batch_roles(B, R, &B->cur);
}

// Some plan was reconfigured since the tables were built: its source and
// receiver buffers, nt, spacings, active sides, launch shape or spare
// buffers may differ from what the tables hold.
bool batch_stale(const awcpml_batch* B) {
    if (B->gens.size() != B->plans.size()) return true;
    for (size_t q = 0; q < B->plans.size(); ++q)
        if (B->plans[q]->io_dirty || B->plans[q]->gen != B->gens[q]) return true;
    return false;
}

int batch_stale_error() {
    return fail(AWCPML_ERR_STATE,
                "a plan of the batch was reconfigured (sources, receivers, model, CPML, launch "
                "shape) since the batch was built or reset: call awcpml_batch_reset");
}

// The batch's values derived from its plans' configuration.  The plans must
// still share block_y and prefetch.
int batch_scan(awcpml_batch* B) {
    const awcpml_plan* P0 = B->plans[0];
    for (size_t q = 0; q < B->plans.size(); ++q)
        if (B->plans[q]->block_y != P0->block_y || B->plans[q]->prefetch != P0->prefetch)
            return fail(AWCPML_ERR_ARG, "plans of a batch must share block_y and prefetch (plan %d differs)",
                        (int)q);
    B->block_y = P0->block_y;
    B->prefetch = P0->prefetch;
    B->halo_max = 0;
    B->nt_max = 0;
    B->nt_min = P0->nt;
    B->use32 = true;
    B->has_io = false;
    for (auto* P : B->plans) {
        B->halo_max = std::max(B->halo_max, P->halo);
        B->nt_max = std::max(B->nt_max, P->nt);
        B->nt_min = std::min(B->nt_min, P->nt);
        B->use32 = B->use32 && !P->wide && (uint64_t)P->field_elems * P->esize < (1ull << 32);
        B->has_io = B->has_io || P->has_io;
    }
    return AWCPML_OK;
}

int batch_build(awcpml_batch* B) {
    const int np = (int)B->plans.size();
    B->block_start.assign(1, 0);
    B->shot_start.assign(1, 0);
    for (int q = 0; q < np; ++q) {
        const awcpml_plan* P = B->plans[q];
        int lb;
        if (B->dtype == AWCPML_F64) lb = make_args<double>(P, 0, 0).logical_blocks;
        else lb = make_args<float>(P, 0, 0).logical_blocks;
        B->block_start.push_back(B->block_start.back() + lb);
        B->shot_start.push_back(B->shot_start.back() + P->nshots);
    }
    B->grid_blocks = (int)round_up(B->block_start.back(), awcpml::kNumXcd);
    return batch_rebuild(B);
}

template <typename T>
int batch_io_t(awcpml_batch* B, const RoleTables& R, int rot, int inject, int64_t it0) {
    const int np = (int)B->plans.size();
    awcpml::IoBatchArgs<T> I;
    memset(&I, 0, sizeof I);
    I.table = (const awcpml::IoArgs<T>*)R.d_io + (size_t)(rot * 2 + inject) * np;
    I.it0 = it0;
    I.nplans = np;
    for (int q = 0; q <= np; ++q) I.shot_start[q] = B->shot_start[q];
    awcpml::sources_receivers_batch<T><<<dim3(B->shot_start[np]), dim3(256), 0, B->stream>>>(I);
    HIP_TRY(hipGetLastError());
    return AWCPML_OK;
}

template <typename T>
int batch_step_t(awcpml_batch* B, int64_t it0, bool with_io) {
    const int np = (int)B->plans.size();
    const RoleTables& R = B->roles[B->cur];
    awcpml::BatchArgs<T> A;
    memset(&A, 0, sizeof A);
    A.table = (const StencilArgs<T>*)R.d_args + (size_t)(it0 % 6) * np;
    A.it0 = it0;
    A.nplans = np;
    A.grid_blocks = B->grid_blocks;
    for (int q = 0; q < np; ++q) A.nt[q] = B->plans[q]->nt;
    for (int q = 0; q <= np; ++q) A.block_start[q] = B->block_start[q];
    awcpml::LaunchShape L;
    L.grid = dim3(B->grid_blocks);
    L.block = dim3(awcpml::kTileX, B->block_y);
    L.lds = (size_t)8 * (B->halo_max + 1) * sizeof(T);
    L.stream = B->stream;
    if (B->ndim == 3) awcpml::launch_batch_stencil<T, 3>(A, L, B->prefetch, !B->use32);
    else awcpml::launch_batch_stencil<T, 2>(A, L, B->prefetch, !B->use32);
    HIP_TRY(hipGetLastError());
    if (with_io && B->has_io) return batch_io_t<T>(B, R, (int)(it0 % 3), 1, it0);
    return AWCPML_OK;
}

int batch_step(awcpml_batch* B, int64_t it0, bool with_io = true) {
    return B->dtype == AWCPML_F64 ? batch_step_t<double>(B, it0, with_io)
                                  : batch_step_t<float>(B, it0, with_io);
}

template <typename T>
int batch_launch2_t(awcpml_batch* B, const RoleTables& R, int64_t it0) {
    const int np = (int)B->plans.size();
    if (!R.d_args2) return fail(AWCPML_ERR_STATE, "batch roles without two-step arguments");
    int W, pf;
    batch_two_shape(B, &W, &pf);
    awcpml::Batch2Args<T> A;
    memset(&A, 0, sizeof A);
    A.table = (const awcpml::Stencil2Args<T>*)R.d_args2 + (size_t)(it0 % 6) * np;
    A.it0 = it0;
    A.nplans = np;
    A.grid_blocks = B->grid_blocks2;
    for (int q = 0; q <= np; ++q) A.block_start[q] = B->block_start2[q];
    awcpml::LaunchShape L;
    L.grid = dim3(B->grid_blocks2);
    L.block = dim3(awcpml::kTileX, W);
    L.lds = (size_t)8 * (B->halo_max + 1) * sizeof(T);
    L.stream = B->stream;
    if (B->ndim == 3) awcpml::launch_batch_stencil2<T, 3>(A, L, W, pf);
    else awcpml::launch_batch_stencil2<T, 2>(A, L, W, pf);
    HIP_TRY(hipGetLastError());
    return AWCPML_OK;
}

// Steps it0 and it0+1 of every plan in one launch, then as do_step2: the
// receivers of step it0 read p^{t+1}, and step it0+1's sources and
// receivers are handled on p^{t+2} under the new roles (whose tables must
// exist when this is captured).
template <typename T>
int batch_step2_t(awcpml_batch* B, int64_t it0) {
    const int before = B->cur;
    int st = batch_launch2_t<T>(B, B->roles[before], it0);
    if (st) return st;
    if (B->has_io && (st = batch_io_t<T>(B, B->roles[before], (int)(it0 % 3), 0, it0))) return st;
    if ((st = batch_roles(B, roles_after_two(B->roles[before].r, it0), &B->cur))) return st;
    if (B->has_io) return batch_io_t<T>(B, B->roles[B->cur], (int)((it0 + 1) % 3), 1, it0 + 1);
    return AWCPML_OK;
}

// Pairs of steps run in two-step launches while every plan has two steps
// left, so that the roles change only while the whole batch runs.
bool batch_pair(const awcpml_batch* B, bool two, int64_t it0, int64_t left) {
    return two && left >= 2 && it0 + 2 <= B->nt_min;
}

// Enqueue steps it0 .. it0+n-1, advancing B->cur as they are enqueued.
int batch_enqueue(awcpml_batch* B, int64_t it0, int64_t n) {
    const bool two = batch_two_ok(B);
    for (int64_t q = 0; q < n;) {
        int st;
        if (batch_pair(B, two, it0 + q, n - q)) {
            st = B->dtype == AWCPML_F64 ? batch_step2_t<double>(B, it0 + q) : batch_step2_t<float>(B, it0 + q);
            q += 2;
        } else {
            st = batch_step(B, it0 + q);
            q += 1;
        }
        if (st) return st;
    }
    return AWCPML_OK;
}

// Build the tables of every roles entry that steps it0 .. it0+n-1 reach, so
// that capturing them allocates nothing.
int batch_prepare(awcpml_batch* B, int64_t it0, int64_t n) {
    if (!batch_two_ok(B)) return AWCPML_OK;
    Roles R = B->roles[B->cur].r;
    for (int64_t q = 0; q < n;) {
        if (batch_pair(B, true, it0 + q, n - q)) {
            R = roles_after_two(R, it0 + q);
            int idx, st = batch_roles(B, R, &idx);
            if (st) return st;
            q += 2;
        } else {
            q += 1;
        }
    }
    return AWCPML_OK;
}

// After the batch ran to step B->it, every plan is at min(it, its nt), in
// the batch's roles (which change only while every plan runs).
void batch_sync_plans(awcpml_batch* B) {
    for (auto* P : B->plans) {
        P->it = std::min(B->it, P->nt);
        set_layout(P, roles_layout(P, B->roles[B->cur].r, P->it));
    }
}

// The batch's graph for n steps from its current step and roles, captured
// on first use.
int batch_graph(awcpml_batch* B, int64_t n, GraphEntry** out) {
    const int64_t it0 = B->it;
    const int before = B->cur;
    for (auto& e : B->graphs)
        if (e.it0 == it0 && e.n == n && e.roles_before == before) {
            *out = &e;
            return AWCPML_OK;
        }
    int st = batch_prepare(B, it0, n);
    if (st) return st;
    if (B->graphs.size() >= 4) {
        (void)hipGraphExecDestroy(B->graphs.front().exec);
        B->graphs.erase(B->graphs.begin());
    }
    hipGraphExec_t exec;
    st = capture_seq(B->stream, [&]() { return batch_enqueue(B, it0, n); }, &exec);
    const int after = B->cur;
    B->cur = before;
    if (st) return st;
    GraphEntry e = {};
    e.it0 = it0;
    e.n = n;
    e.exec = exec;
    e.roles_before = before;
    e.roles_after = after;
    B->graphs.push_back(e);
    *out = &B->graphs.back();
    return AWCPML_OK;
}

int batch_run_steps(awcpml_batch* B, int64_t n) {
    if (n <= 0) return AWCPML_OK;
    if (B->use_graph) {
        GraphEntry* g = nullptr;
        int st = batch_graph(B, n, &g);
        if (st) return st;
        HIP_TRY(hipGraphLaunch(g->exec, B->stream));
        B->it += n;
        B->cur = g->roles_after;
        return AWCPML_OK;
    }
    int st = batch_enqueue(B, B->it, n);
    if (st) return st;
    B->it += n;
    return AWCPML_OK;
}

}  // namespace

extern "C" {

const char* awcpml_last_error(void) { return g_err.c_str(); }

const char* awcpml_version(void) { return "awcpml 0.1.0 (gfx950)"; }

int awcpml_plan_create(int ndim, int64_t nx, int64_t ny, int64_t nz, int halo, int dtype,
                       int nshots, awcpml_plan** out) {
    g_err.clear();
    if (!out) return fail(AWCPML_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (ndim != 2 && ndim != 3) return fail(AWCPML_ERR_ARG, "ndim must be 2 or 3 (got %d)", ndim);
    if (dtype != AWCPML_F64 && dtype != AWCPML_F32) return fail(AWCPML_ERR_ARG, "unknown dtype %d", dtype);
    if (nshots < 1) return fail(AWCPML_ERR_ARG, "nshots must be >= 1");
    if (halo < 0) return fail(AWCPML_ERR_ARG, "halo must be >= 0");
    if (ndim == 2) nz = 1;
    int64_t n[3] = {nx, ny, nz};
    for (int a = 0; a < ndim; ++a) {
        if (n[a] < 3) return fail(AWCPML_ERR_ARG, "axis %d has %lld points (need >= 3)", a, (long long)n[a]);
        // the right psi layers start at node n-halo-2 (0-based) and must lie
        // in the grid, as in update_psi_*! (acoustic_2D_xPU.jl:7-31); below
        // 2*halo+2 points the two slabs overlap and the left side wins
        if (n[a] < (int64_t)halo + 2)
            return fail(AWCPML_ERR_ARG, "axis %d has %lld points: halo %d needs >= %d", a,
                        (long long)n[a], halo, halo + 2);
    }
    awcpml_plan* P = new (std::nothrow) awcpml_plan();
    if (!P) return fail(AWCPML_ERR_NOMEM, "host allocation failed");
    P->ndim = ndim;
    P->dtype = dtype;
    P->nshots = nshots;
    P->halo = halo;
    for (int a = 0; a < 3; ++a) P->n[a] = n[a];
    P->esize = dtype == AWCPML_F64 ? 8 : 4;
    // aligned two-step tiles: on for FP32 (whole-line stores pay for the
    // edge nodes' extra loads); FP64 measured faster with 62-node tiles
    P->tb_align = dtype == AWCPML_F32;
    P->px = round_up(nx, (int64_t)(128 / P->esize));
    P->field_elems = P->px * ny * nz;
    P->src_lin.resize(nshots);
    P->rec_lin.resize(nshots);
    P->src_wav.resize(nshots);
    P->src_set.assign(nshots, false);
    P->rec_set.assign(nshots, false);
    int st = AWCPML_OK;
    hipError_t he = hipGetDevice(&P->device);
    if (he != hipSuccess) {
        delete P;
        return fail(AWCPML_ERR_HIP, "hipGetDevice failed: %s", hipGetErrorString(he));
    }
    he = hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking);
    if (he != hipSuccess) {
        delete P;
        return fail(AWCPML_ERR_HIP, "hipStreamCreate failed: %s", hipGetErrorString(he));
    }
    const size_t fbytes = (size_t)P->field_elems * nshots * P->esize;
    for (int b = 0; b < 3 && !st; ++b) st = dalloc(&P->p[b], fbytes);
    if (!st) st = dalloc(&P->fact, (size_t)P->field_elems * P->esize);
    {
        const int64_t tiles = (nx + awcpml::kTileX - 1) / awcpml::kTileX;
        P->fe_elems = ndim == 3 ? nz * tiles * 2 * ny : ny * tiles * 2;
        if (!st) st = dalloc(&P->fact_edge, (size_t)P->fe_elems * P->esize);
    }
    for (int a = 0; a < ndim && !st; ++a) {
        P->psi_elems[a] = slab_elems(P, a, halo + 1);
        P->xi_elems[a] = slab_elems(P, a, halo);
        for (int s = 0; s < 2 && !st; ++s) {
            st = dalloc(&P->psi[a][s][0], (size_t)P->psi_elems[a] * nshots * P->esize);
            if (!st) st = dalloc(&P->psi[a][s][1], (size_t)P->psi_elems[a] * nshots * P->esize);
            if (!st) st = dalloc(&P->xi[a][s], (size_t)P->xi_elems[a] * nshots * P->esize);
        }
        for (int c = 0; c < 8 && !st; ++c) st = dalloc(&P->coef[a][c], (size_t)(halo + 1) * P->esize);
    }
    if (!st) {
        hipError_t e = hipMemsetAsync(P->fact, 0, (size_t)P->field_elems * P->esize, P->stream);
        if (e == hipSuccess)
            e = hipMemsetAsync(P->fact_edge, 0, (size_t)P->fe_elems * P->esize, P->stream);
        if (e != hipSuccess) st = fail(AWCPML_ERR_HIP, "hipMemset failed: %s", hipGetErrorString(e));
    }
    P->home = get_layout(P);
    if (!st) st = awcpml_reset(P);
    if (!st) st = default_tiling(P);
    if (st) {
        std::string keep = g_err;
        awcpml_plan_destroy(P);
        g_err = keep;
        return st;
    }
    *out = P;
    return AWCPML_OK;
}

int awcpml_plan_destroy(awcpml_plan* P) {
    if (!P) return AWCPML_OK;
    if (P->stream) (void)hipStreamSynchronize(P->stream);
    drop_graphs(P->graphs);
    for (int b = 0; b < 3; ++b) dfree(P->p[b]);
    dfree(P->p_spare);
    dfree(P->fact);
    dfree(P->fact_edge);
    for (int a = 0; a < 3; ++a) {
        for (int s = 0; s < 2; ++s) {
            dfree(P->psi[a][s][0]);
            dfree(P->psi[a][s][1]);
            dfree(P->psi_spare[a][s]);
            dfree(P->xi[a][s]);
            dfree(P->xi_spare[a][s]);
        }
        for (int c = 0; c < 8; ++c) dfree(P->coef[a][c]);
    }
    dfree(P->d_src_lin);
    dfree(P->d_src_wav);
    dfree(P->d_src_start);
    dfree(P->d_src_seq);
    dfree(P->d_rec_lin);
    dfree(P->d_rec_start);
    dfree(P->d_traces);
    dfree(P->ck);
    if (P->stream) (void)hipStreamDestroy(P->stream);
    delete P;
    return AWCPML_OK;
}

int awcpml_set_model(awcpml_plan* P, const double* fact, const double* inv_d, const double* inv_d2) {
    g_err.clear();
    int st = check_plan(P);
    if (st) return st;
    if (!fact || !inv_d || !inv_d2) return fail(AWCPML_ERR_ARG, "NULL argument");
    P->ck_valid = false;
    for (int a = 0; a < P->ndim; ++a) {
        P->inv_d[a] = inv_d[a];
        P->inv_d2[a] = inv_d2[a];
    }
    const int64_t nx = P->n[0], rows = P->n[1] * P->n[2];
    // the aligned tiles' edge-node copy: same values, rows of an edge adjacent
    const int64_t ny = P->n[1], nz = P->ndim == 3 ? P->n[2] : 1;
    const int64_t tiles = (nx + awcpml::kTileX - 1) / awcpml::kTileX;
    std::vector<double> fe((size_t)P->fe_elems);
    for (int64_t k = 0; k < (P->ndim == 3 ? nz : ny); ++k)
        for (int64_t t = 0; t < tiles; ++t)
            for (int s = 0; s < 2; ++s) {
                int64_t x = s == 0 ? t * awcpml::kTileX - 1 : t * awcpml::kTileX + awcpml::kTileX;
                x = x < 1 ? 1 : (x > nx - 2 ? nx - 2 : x);
                if (P->ndim == 3) {
                    for (int64_t j = 0; j < ny; ++j)
                        fe[(size_t)(((k * tiles + t) * 2 + s) * ny + j)] = fact[x + nx * (j + ny * k)];
                } else {
                    fe[(size_t)((k * tiles + t) * 2 + s)] = fact[x + nx * k];
                }
            }
    if (P->dtype == AWCPML_F64) {
        HIP_TRY(hipMemcpy2DAsync(P->fact, P->px * 8, fact, nx * 8, nx * 8, rows, hipMemcpyHostToDevice,
                                 P->stream));
        HIP_TRY(hipMemcpyAsync(P->fact_edge, fe.data(), fe.size() * 8, hipMemcpyHostToDevice, P->stream));
    } else {
        std::vector<float> tmp((size_t)nx * rows);
        for (size_t q = 0; q < tmp.size(); ++q) tmp[q] = (float)fact[q];
        HIP_TRY(hipMemcpy2DAsync(P->fact, P->px * 4, tmp.data(), nx * 4, nx * 4, rows, hipMemcpyHostToDevice,
                                 P->stream));
        std::vector<float> fe32(fe.size());
        for (size_t q = 0; q < fe.size(); ++q) fe32[q] = (float)fe[q];
        HIP_TRY(hipMemcpyAsync(P->fact_edge, fe32.data(), fe32.size() * 4, hipMemcpyHostToDevice, P->stream));
    }
    HIP_TRY(hipStreamSynchronize(P->stream));
    P->model_set = true;
    ++P->gen;
    return AWCPML_OK;
}

int awcpml_set_cpml(awcpml_plan* P, int axis, const double* a_l, const double* a_r, const double* b_K_l,
                    const double* b_K_r, const double* a_hl, const double* a_hr, const double* b_K_hl,
                    const double* b_K_hr) {
    g_err.clear();
    int st = check_plan(P);
    if (st) return st;
    if (axis < 0 || axis >= P->ndim) return fail(AWCPML_ERR_ARG, "axis %d out of range", axis);
    const double* v[8] = {a_l, a_r, b_K_l, b_K_r, a_hl, a_hr, b_K_hl, b_K_hr};
    const int h = P->halo;
    for (int c = 0; c < 8; ++c) {
        const size_t len = (size_t)h + (c >= 4 ? 1 : 0);
        if (len && !v[c]) return fail(AWCPML_ERR_ARG, "coefficient vector %d is NULL", c);
        if ((st = upload(P, P->coef[axis][c], v[c], len))) return st;
    }
    HIP_TRY(hipStreamSynchronize(P->stream));
    for (int s = 0; s < 2; ++s) {
        bool idle = true;
        for (int q = 0; q < h && idle; ++q) idle = v[0 + s][q] == 0.0 && v[2 + s][q] == 1.0;
        for (int q = 0; q < h + 1 && idle; ++q) idle = v[4 + s][q] == 0.0 && v[6 + s][q] == 1.0;
        P->active[axis][s] = (h > 0 && !idle) ? 1 : 0;
    }
    P->cpml_set[axis] = true;
    P->ck_valid = false;
    drop_graphs(P->graphs);
    ++P->gen;
    return AWCPML_OK;
}

int awcpml_set_sources(awcpml_plan* P, int shot, int nsrcs, const int64_t* pos, int64_t nt,
                       const double* dt2srctf) {
    g_err.clear();
    int st = check_plan(P);
    if (st) return st;
    if (shot < 0 || shot >= P->nshots) return fail(AWCPML_ERR_ARG, "shot %d out of range", shot);
    if (nsrcs < 0 || nt < 0) return fail(AWCPML_ERR_ARG, "negative count");
    if (nsrcs > 0 && (!pos || (nt > 0 && !dt2srctf))) return fail(AWCPML_ERR_ARG, "NULL argument");
    if (P->nt >= 0 && nt != P->nt) {
        bool other = false;
        for (int s = 0; s < P->nshots; ++s) other |= (s != shot && P->src_set[s]);
        if (other) return fail(AWCPML_ERR_ARG, "nt %lld differs from the plan's nt %lld", (long long)nt, (long long)P->nt);
    }
    if ((st = check_positions(P, pos, nsrcs, "source"))) return st;
    if (nt != P->nt) {
        // traces of other shots are sized by nt: they are rebuilt on upload
        P->nt = nt;
    }
    P->src_lin[shot].resize(nsrcs);
    for (int s = 0; s < nsrcs; ++s) P->src_lin[shot][s] = lin_index(P, pos, nsrcs, s);
    P->src_wav[shot].assign(dt2srctf, dt2srctf + (size_t)nt * nsrcs);
    P->src_set[shot] = true;
    P->io_dirty = true;
    P->ck_valid = false;
    return AWCPML_OK;
}

int awcpml_set_receivers(awcpml_plan* P, int shot, int nrecs, const int64_t* pos) {
    g_err.clear();
    int st = check_plan(P);
    if (st) return st;
    if (shot < 0 || shot >= P->nshots) return fail(AWCPML_ERR_ARG, "shot %d out of range", shot);
    if (nrecs < 0) return fail(AWCPML_ERR_ARG, "negative count");
    if (nrecs > 0 && !pos) return fail(AWCPML_ERR_ARG, "NULL argument");
    if ((st = check_positions(P, pos, nrecs, "receiver"))) return st;
    P->rec_lin[shot].resize(nrecs);
    for (int r = 0; r < nrecs; ++r) P->rec_lin[shot][r] = lin_index(P, pos, nrecs, r);
    P->rec_set[shot] = true;
    P->io_dirty = true;
    P->ck_valid = false;
    return AWCPML_OK;
}

int awcpml_reset(awcpml_plan* P) {
    int st = check_plan(P);
    if (st) return st;
    set_layout(P, P->home);
    const size_t fbytes = (size_t)P->field_elems * P->nshots * P->esize;
    for (int b = 0; b < 3; ++b) HIP_TRY(hipMemsetAsync(P->p[b], 0, fbytes, P->stream));
    if (P->p_spare) HIP_TRY(hipMemsetAsync(P->p_spare, 0, fbytes, P->stream));
    for (int a = 0; a < P->ndim; ++a)
        for (int s = 0; s < 2; ++s) {
            const size_t pb = (size_t)P->psi_elems[a] * P->nshots * P->esize;
            const size_t xb = (size_t)P->xi_elems[a] * P->nshots * P->esize;
            HIP_TRY(hipMemsetAsync(P->psi[a][s][0], 0, pb, P->stream));
            HIP_TRY(hipMemsetAsync(P->psi[a][s][1], 0, pb, P->stream));
            if (P->psi_spare[a][s]) HIP_TRY(hipMemsetAsync(P->psi_spare[a][s], 0, pb, P->stream));
            if (xb) HIP_TRY(hipMemsetAsync(P->xi[a][s], 0, xb, P->stream));
            if (xb && P->xi_spare[a][s]) HIP_TRY(hipMemsetAsync(P->xi_spare[a][s], 0, xb, P->stream));
        }
    P->boundary_dirty = false;
    if (P->d_traces && P->n_rec_total > 0 && P->nt > 0)
        HIP_TRY(hipMemsetAsync(P->d_traces, 0, (size_t)P->n_rec_total * P->nt * P->esize, P->stream));
    HIP_TRY(hipStreamSynchronize(P->stream));
    P->it = 0;
    P->rot = 0;
    P->par = 0;
    return AWCPML_OK;
}

int awcpml_advance(awcpml_plan* P, int64_t nsteps) {
    g_err.clear();
    int st = check_plan(P);
    if (st) return st;
    if (nsteps < 0) return fail(AWCPML_ERR_ARG, "nsteps < 0");
    if ((st = ready(P))) return st;
    if (P->it + nsteps > P->nt)
        return fail(AWCPML_ERR_STATE, "advance past nt: %lld + %lld > %lld", (long long)P->it,
                    (long long)nsteps, (long long)P->nt);
    if ((st = run_steps(P, nsteps))) return st;
    HIP_TRY(hipStreamSynchronize(P->stream));
    return AWCPML_OK;
}

int awcpml_forward(awcpml_plan* P, int64_t it) {
    g_err.clear();
    int st = check_plan(P);
    if (st) return st;
    if ((st = ready(P))) return st;
    if (it < 1 || it > P->nt) return fail(AWCPML_ERR_ARG, "it %lld outside 1..nt", (long long)it);
    if ((st = do_step(P, it - 1))) return st;
    HIP_TRY(hipStreamSynchronize(P->stream));
    return AWCPML_OK;
}

int awcpml_steps_done(const awcpml_plan* P, int64_t* out) {
    if (!P || !out) return fail(AWCPML_ERR_ARG, "NULL argument");
    *out = P->it;
    return AWCPML_OK;
}

int awcpml_get_traces(awcpml_plan* P, int shot, double* out) {
    g_err.clear();
    int st = check_plan(P);
    if (st) return st;
    if (shot < 0 || shot >= P->nshots) return fail(AWCPML_ERR_ARG, "shot %d out of range", shot);
    if ((st = sync_io(P))) return st;
    const int64_t nrec = (int64_t)P->rec_lin[shot].size();
    if (nrec == 0 || P->nt <= 0) return AWCPML_OK;
    if (!out) return fail(AWCPML_ERR_ARG, "out is NULL");
    int64_t first = 0;
    for (int s = 0; s < shot; ++s) first += (int64_t)P->rec_lin[s].size();
    const size_t count = (size_t)nrec * P->nt;
    if (P->dtype == AWCPML_F64) {
        HIP_TRY(hipMemcpyAsync(out, (char*)P->d_traces + (size_t)first * P->nt * 8, count * 8,
                               hipMemcpyDeviceToHost, P->stream));
        HIP_TRY(hipStreamSynchronize(P->stream));
    } else {
        std::vector<float> tmp(count);
        HIP_TRY(hipMemcpyAsync(tmp.data(), (char*)P->d_traces + (size_t)first * P->nt * 4, count * 4,
                               hipMemcpyDeviceToHost, P->stream));
        HIP_TRY(hipStreamSynchronize(P->stream));
        for (size_t q = 0; q < count; ++q) out[q] = tmp[q];
    }
    return AWCPML_OK;
}

int awcpml_get_field(awcpml_plan* P, int shot, int which, double* out) {
    g_err.clear();
    int st = check_plan(P);
    if (st) return st;
    if (shot < 0 || shot >= P->nshots) return fail(AWCPML_ERR_ARG, "shot %d out of range", shot);
    if (which < 0 || which > 2) return fail(AWCPML_ERR_ARG, "which must be 0, 1 or 2");
    if (!out) return fail(AWCPML_ERR_ARG, "out is NULL");
    const char* src = (const char*)P->p[(P->rot + which) % 3] + (size_t)shot * P->field_elems * P->esize;
    const int64_t nx = P->n[0], rows = P->n[1] * P->n[2];
    if (P->dtype == AWCPML_F64) {
        HIP_TRY(hipMemcpy2DAsync(out, nx * 8, src, P->px * 8, nx * 8, rows, hipMemcpyDeviceToHost, P->stream));
        HIP_TRY(hipStreamSynchronize(P->stream));
    } else {
        std::vector<float> tmp((size_t)nx * rows);
        HIP_TRY(hipMemcpy2DAsync(tmp.data(), nx * 4, src, P->px * 4, nx * 4, rows, hipMemcpyDeviceToHost,
                                 P->stream));
        HIP_TRY(hipStreamSynchronize(P->stream));
        for (size_t q = 0; q < tmp.size(); ++q) out[q] = tmp[q];
    }
    return AWCPML_OK;
}

int awcpml_run(awcpml_plan* P, double* traces, double* pend) {
    int st = check_plan(P);
    if (st) return st;
    if (P->nshots != 1) return fail(AWCPML_ERR_ARG, "awcpml_run needs a one-shot plan");
    if ((st = awcpml_reset(P))) return st;
    if ((st = awcpml_advance(P, P->nt))) return st;
    if (traces && (st = awcpml_get_traces(P, 0, traces))) return st;
    if (pend && (st = awcpml_get_field(P, 0, 1, pend))) return st;
    return AWCPML_OK;
}

int awcpml_time_stencil(awcpml_plan* P, int nlaunches, double* ms_total) {
    g_err.clear();
    int st = check_plan(P);
    if (st) return st;
    if (!ms_total || nlaunches < 1) return fail(AWCPML_ERR_ARG, "bad argument");
    if ((st = ready(P))) return st;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, P->stream));
    for (int q = 0; q < nlaunches; ++q) {
        if ((st = launch_stencil(P))) break;
        P->rot = (P->rot + 1) % 3;
        P->par ^= 1;
    }
    HIP_TRY(hipEventRecord(e1, P->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_total = ms;
    return st;
}

int awcpml_time_advance(awcpml_plan* P, int64_t nsteps, double* ms_total) {
    g_err.clear();
    int st = check_plan(P);
    if (st) return st;
    if (!ms_total || nsteps < 0) return fail(AWCPML_ERR_ARG, "bad argument");
    if ((st = ready(P))) return st;
    if (P->it + nsteps > P->nt) return fail(AWCPML_ERR_STATE, "advance past nt");
    if (P->use_graph && nsteps > 0) {
        // capture outside the timed region
        GraphEntry* g = nullptr;
        if ((st = plan_graph(P, nsteps, &g))) return st;
    }
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, P->stream));
    st = run_steps(P, nsteps);
    HIP_TRY(hipEventRecord(e1, P->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_total = ms;
    return st;
}

int awcpml_set_graph(awcpml_plan* P, int enable) {
    g_err.clear();
    int st = check_plan(P);
    if (st) return st;
    P->use_graph = enable != 0;
    if (!P->use_graph) drop_graphs(P->graphs);
    return AWCPML_OK;
}

int awcpml_set_tiling(awcpml_plan* P, int block_y, int march_len, int prefetch) {
    g_err.clear();
    int st = check_plan(P);
    if (st) return st;
    if (block_y != 0 && block_y != 1 && block_y != 2 && block_y != 4 && block_y != 8)
        return fail(AWCPML_ERR_ARG, "block_y must be 1, 2, 4 or 8 (got %d)", block_y);
    if (march_len < 0) return fail(AWCPML_ERR_ARG, "march_len < 0");
    if (prefetch < 0 || prefetch > 5) return fail(AWCPML_ERR_ARG, "prefetch must be 0..5 (got %d)", prefetch);
    P->block_y = block_y;
    P->march_len = march_len;
    P->prefetch = prefetch ? prefetch : 2;
    drop_graphs(P->graphs);
    ++P->gen;
    int st2 = default_tiling(P);
    if (P->block_y % P->shot_group) P->shot_group = 1;
    return st2;
}

int awcpml_set_shot_group(awcpml_plan* P, int shot_group) {
    g_err.clear();
    int st = check_plan(P);
    if (st) return st;
    if (shot_group != 1 && shot_group != 2 && shot_group != 4 && shot_group != 8)
        return fail(AWCPML_ERR_ARG, "shot_group must be 1, 2, 4 or 8 (got %d)", shot_group);
    if (P->block_y % shot_group)
        return fail(AWCPML_ERR_ARG, "shot_group %d does not divide block_y %d", shot_group, P->block_y);
    P->shot_group = shot_group;
    drop_graphs(P->graphs);
    ++P->gen;
    return AWCPML_OK;
}

int awcpml_get_shot_group(const awcpml_plan* P, int* shot_group) {
    if (!P || !shot_group) return fail(AWCPML_ERR_ARG, "NULL argument");
    *shot_group = P->shot_group;
    return AWCPML_OK;
}

int awcpml_set_time_block(awcpml_plan* P, int steps, int block_rows, int march_len, int prefetch) {
    g_err.clear();
    int st = check_plan(P);
    if (st) return st;
    if (steps != 1 && steps != 2) return fail(AWCPML_ERR_ARG, "steps must be 1 or 2 (got %d)", steps);
    if (P->ndim == 3 && block_rows != 0 && block_rows != 4 && block_rows != 8 && block_rows != 16)
        return fail(AWCPML_ERR_ARG, "block_rows must be 0, 4, 8 or 16 for 3D plans (got %d)", block_rows);
    if (P->ndim == 2 && block_rows != 0 && block_rows != 4 && block_rows != 8)
        return fail(AWCPML_ERR_ARG, "block_rows must be 0, 4 or 8 for 2D plans (got %d)", block_rows);
    if (march_len < 0) return fail(AWCPML_ERR_ARG, "march_len < 0");
    if (prefetch < 0 || prefetch > 3) return fail(AWCPML_ERR_ARG, "prefetch must be 0..3 (got %d)", prefetch);
    if (steps == 2 && !P->p_spare) {
        // spares start as zeros, like every buffer after a reset; a spare is
        // allocated before any two-step launch, so the layout is still home
        const size_t fbytes = (size_t)P->field_elems * P->nshots * P->esize;
        if ((st = dalloc(&P->p_spare, fbytes))) return st;
        HIP_TRY(hipMemsetAsync(P->p_spare, 0, fbytes, P->stream));
        for (int a = 0; a < P->ndim; ++a)
            for (int s = 0; s < 2; ++s) {
                const size_t pb = (size_t)P->psi_elems[a] * P->nshots * P->esize;
                const size_t xb = (size_t)P->xi_elems[a] * P->nshots * P->esize;
                if ((st = dalloc(&P->psi_spare[a][s], pb))) return st;
                if ((st = dalloc(&P->xi_spare[a][s], xb))) return st;
                HIP_TRY(hipMemsetAsync(P->psi_spare[a][s], 0, pb, P->stream));
                if (xb) HIP_TRY(hipMemsetAsync(P->xi_spare[a][s], 0, xb, P->stream));
            }
        HIP_TRY(hipStreamSynchronize(P->stream));
        P->home.p[3] = P->p_spare;
        for (int a = 0; a < 3; ++a)
            for (int s = 0; s < 2; ++s) {
                P->home.psi[a][s][2] = P->psi_spare[a][s];
                P->home.xi[a][s][1] = P->xi_spare[a][s];
            }
    }
    P->time_block = steps;
    P->tb_rows = block_rows;
    P->tb_march = march_len;
    P->tb_prefetch = prefetch;
    drop_graphs(P->graphs);
    ++P->gen;
    return AWCPML_OK;
}

int awcpml_get_time_block(const awcpml_plan* P, int* steps, int* block_rows, int* march_len,
                          int* prefetch) {
    if (!P || !steps || !block_rows || !march_len || !prefetch)
        return fail(AWCPML_ERR_ARG, "NULL argument");
    int W, ml, pf;
    two_tiling(P, &W, &ml, &pf);
    // the steps a launch advances now (1 where two-step launches do not apply)
    *steps = two_ok(P) ? 2 : 1;
    *block_rows = W;
    *march_len = ml;
    *prefetch = pf;
    return AWCPML_OK;
}

int awcpml_set_two_step_align(awcpml_plan* P, int enable) {
    g_err.clear();
    int st = check_plan(P);
    if (st) return st;
    P->tb_align = enable != 0;
    drop_graphs(P->graphs);
    ++P->gen;
    return AWCPML_OK;
}

int awcpml_get_two_step_align(const awcpml_plan* P, int* enable) {
    if (!P || !enable) return fail(AWCPML_ERR_ARG, "NULL argument");
    *enable = P->tb_align;
    return AWCPML_OK;
}

int awcpml_time_stencil2(awcpml_plan* P, int nlaunches, double* ms_total) {
    g_err.clear();
    int st = check_plan(P);
    if (st) return st;
    if (!ms_total || nlaunches < 1) return fail(AWCPML_ERR_ARG, "bad argument");
    if ((st = ready(P))) return st;
    if (!two_ok(P)) return fail(AWCPML_ERR_STATE, "two-step launches do not apply to this plan");
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, P->stream));
    for (int q = 0; q < nlaunches; ++q) {
        st = P->dtype == AWCPML_F64 ? launch_stencil2_t<double>(P, 0) : launch_stencil2_t<float>(P, 0);
        if (st) break;
        two_swap(P);
    }
    HIP_TRY(hipEventRecord(e1, P->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_total = ms;
    return st;
}

int awcpml_batch_create(awcpml_plan* const* plans, int count, awcpml_batch** out) {
    g_err.clear();
    if (!out || !plans) return fail(AWCPML_ERR_ARG, "NULL argument");
    *out = nullptr;
    if (count < 1 || count > awcpml::kMaxBatch)
        return fail(AWCPML_ERR_ARG, "a batch holds 1..%d plans (got %d)", awcpml::kMaxBatch, count);
    for (int q = 0; q < count; ++q) {
        if (!plans[q]) return fail(AWCPML_ERR_ARG, "plan %d is NULL", q);
        for (int r = 0; r < q; ++r)
            if (plans[r] == plans[q]) return fail(AWCPML_ERR_ARG, "plan %d appears twice", q);
        if (plans[q]->ndim != plans[0]->ndim || plans[q]->dtype != plans[0]->dtype)
            return fail(AWCPML_ERR_ARG, "plans of a batch must share ndim and dtype (plan %d differs)", q);
        int st = ready(plans[q]);
        if (st) return st;
        if (!at_home(plans[q]))
            return fail(AWCPML_ERR_STATE,
                        "plan %d's buffers traded roles in two-step launches (alone or in a batch) "
                        "since its last reset: reset it before batching", q);
    }
    awcpml_batch* B = new (std::nothrow) awcpml_batch();
    if (!B) return fail(AWCPML_ERR_NOMEM, "host allocation failed");
    B->plans.assign(plans, plans + count);
    B->ndim = plans[0]->ndim;
    B->dtype = plans[0]->dtype;
    {
        int st = batch_scan(B);
        if (st) {
            delete B;
            return st;
        }
    }
    // two-step launches when every plan asks for them with one launch shape
    bool two = true;
    for (auto* P : B->plans)
        two = two && P->time_block == 2 && P->tb_rows == plans[0]->tb_rows &&
              P->tb_prefetch == plans[0]->tb_prefetch;
    if (two) {
        B->time_block = 2;
        B->tb_rows = plans[0]->tb_rows;
        B->tb_prefetch = plans[0]->tb_prefetch;
    }
    hipError_t e = hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete B;
        return fail(AWCPML_ERR_HIP, "hipStreamCreate failed: %s", hipGetErrorString(e));
    }
    int st = batch_build(B);
    if (st) {
        std::string keep = g_err;
        awcpml_batch_destroy(B);
        g_err = keep;
        return st;
    }
    *out = B;
    return AWCPML_OK;
}

int awcpml_batch_destroy(awcpml_batch* B) {
    if (!B) return AWCPML_OK;
    if (B->stream) (void)hipStreamSynchronize(B->stream);
    drop_graphs(B->graphs);
    for (auto& t : B->roles) free_tables(t);
    if (B->stream) (void)hipStreamDestroy(B->stream);
    delete B;
    return AWCPML_OK;
}

int awcpml_batch_reset(awcpml_batch* B) {
    g_err.clear();
    if (!B) return fail(AWCPML_ERR_ARG, "batch is NULL");
    for (auto* P : B->plans) {
        int st = awcpml_reset(P);
        if (st) return st;
        if ((st = ready(P))) return st;
    }
    B->it = 0;
    if (batch_stale(B)) {
        // a plan was reconfigured: rebuild every table from the plans' new
        // buffers and values (the plans are at home after their reset)
        HIP_TRY(hipStreamSynchronize(B->stream));
        drop_graphs(B->graphs);
        for (auto& t : B->roles) free_tables(t);
        B->roles.clear();
        B->cur = 0;
        int st = batch_scan(B);
        if (st) return st;
        return batch_build(B);
    }
    return batch_roles(B, home_roles(), &B->cur);
}

int awcpml_batch_advance(awcpml_batch* B, int64_t nsteps) {
    g_err.clear();
    if (!B) return fail(AWCPML_ERR_ARG, "batch is NULL");
    if (nsteps < 0) return fail(AWCPML_ERR_ARG, "nsteps < 0");
    if (batch_stale(B)) return batch_stale_error();
    for (auto* P : B->plans)
        if (P->it != std::min(B->it, P->nt) ||
            !(get_layout(P) == roles_layout(P, B->roles[B->cur].r, P->it)))
            return fail(AWCPML_ERR_STATE, "a plan of the batch was stepped or changed on its own; "
                                          "rebuild the batch or call awcpml_batch_reset");
    if (B->it + nsteps > B->nt_max)
        return fail(AWCPML_ERR_STATE, "advance past the longest run: %lld + %lld > %lld",
                    (long long)B->it, (long long)nsteps, (long long)B->nt_max);
    int st = batch_run_steps(B, nsteps);
    if (!st) {
        hipError_t e = hipStreamSynchronize(B->stream);
        if (e != hipSuccess) st = fail(AWCPML_ERR_HIP, "hipStreamSynchronize failed: %s", hipGetErrorString(e));
    }
    batch_sync_plans(B);
    return st;
}

int awcpml_batch_set_graph(awcpml_batch* B, int enable) {
    g_err.clear();
    if (!B) return fail(AWCPML_ERR_ARG, "batch is NULL");
    B->use_graph = enable != 0;
    if (!B->use_graph) drop_graphs(B->graphs);
    return AWCPML_OK;
}

int awcpml_batch_time_advance(awcpml_batch* B, int64_t nsteps, double* ms_total) {
    g_err.clear();
    if (!B || !ms_total || nsteps < 0) return fail(AWCPML_ERR_ARG, "bad argument");
    if (batch_stale(B)) return batch_stale_error();
    if (B->it + nsteps > B->nt_max) return fail(AWCPML_ERR_STATE, "advance past the longest run");
    int st;
    if (nsteps > 0 && B->use_graph) {  // capture outside the timed region
        GraphEntry* g = nullptr;
        if ((st = batch_graph(B, nsteps, &g))) return st;
    } else if ((st = batch_prepare(B, B->it, nsteps))) {
        return st;
    }
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, B->stream));
    st = batch_run_steps(B, nsteps);
    HIP_TRY(hipEventRecord(e1, B->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_total = ms;
    batch_sync_plans(B);
    return st;
}

int awcpml_batch_time_stencil(awcpml_batch* B, int nlaunches, double* ms_total) {
    g_err.clear();
    if (!B || !ms_total || nlaunches < 1) return fail(AWCPML_ERR_ARG, "bad argument");
    if (batch_stale(B)) return batch_stale_error();
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, B->stream));
    int st = AWCPML_OK;
    for (int q = 0; q < nlaunches && !st; ++q) st = batch_step(B, q % 6, false);
    HIP_TRY(hipEventRecord(e1, B->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_total = ms;
    return st;
}

int awcpml_batch_set_time_block(awcpml_batch* B, int steps, int block_rows, int prefetch) {
    g_err.clear();
    if (!B) return fail(AWCPML_ERR_ARG, "batch is NULL");
    if (steps != 1 && steps != 2) return fail(AWCPML_ERR_ARG, "steps must be 1 or 2 (got %d)", steps);
    if (B->ndim == 3 && block_rows != 0 && block_rows != 4 && block_rows != 8 && block_rows != 16)
        return fail(AWCPML_ERR_ARG, "block_rows must be 0, 4, 8 or 16 for 3D plans (got %d)", block_rows);
    if (B->ndim == 2 && block_rows != 0 && block_rows != 4 && block_rows != 8)
        return fail(AWCPML_ERR_ARG, "block_rows must be 0, 4 or 8 for 2D plans (got %d)", block_rows);
    if (prefetch < 0 || prefetch > 3) return fail(AWCPML_ERR_ARG, "prefetch must be 0..3 (got %d)", prefetch);
    HIP_TRY(hipStreamSynchronize(B->stream));
    // each plan keeps its own chunk length; spares are allocated where missing
    for (auto* P : B->plans) {
        int st = awcpml_set_time_block(P, steps, block_rows, P->tb_march, prefetch);
        if (st) return st;
    }
    B->time_block = steps;
    B->tb_rows = block_rows;
    B->tb_prefetch = prefetch;
    return batch_rebuild(B);
}

int awcpml_batch_get_time_block(const awcpml_batch* B, int* steps, int* block_rows, int* prefetch) {
    if (!B || !steps || !block_rows || !prefetch) return fail(AWCPML_ERR_ARG, "NULL argument");
    *steps = batch_two_ok(B) ? 2 : 1;
    batch_two_shape(B, block_rows, prefetch);
    return AWCPML_OK;
}

int awcpml_batch_time_stencil2(awcpml_batch* B, int nlaunches, double* ms_total) {
    g_err.clear();
    if (!B || !ms_total || nlaunches < 1) return fail(AWCPML_ERR_ARG, "bad argument");
    if (batch_stale(B)) return batch_stale_error();
    if (!batch_two_ok(B)) return fail(AWCPML_ERR_STATE, "two-step launches do not apply to this batch");
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, B->stream));
    int st = AWCPML_OK;
    const RoleTables& R = B->roles[B->cur];
    for (int q = 0; q < nlaunches && !st; ++q)
        st = B->dtype == AWCPML_F64 ? batch_launch2_t<double>(B, R, 0) : batch_launch2_t<float>(B, R, 0);
    HIP_TRY(hipEventRecord(e1, B->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_total = ms;
    return st;
}

int awcpml_set_wide_offsets(awcpml_plan* P, int enable) {
    g_err.clear();
    int st = check_plan(P);
    if (st) return st;
    P->wide = enable != 0;
    drop_graphs(P->graphs);
    ++P->gen;
    return AWCPML_OK;
}

int awcpml_get_tiling(const awcpml_plan* P, int* block_y, int* march_len, int* prefetch) {
    if (!P || !block_y || !march_len || !prefetch) return fail(AWCPML_ERR_ARG, "NULL argument");
    *block_y = P->block_y;
    *march_len = P->march_len;
    *prefetch = P->prefetch;
    return AWCPML_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------
// Adjoint (FWI gradient): checkpointed forward, then per checkpoint
// segment a forward replay that stores q^s = d(pnew)/d(fact) for every
// step of the segment, followed by the reverse-time steps of the segment
// (adjoint_kernels.h).  Memory: ck_count checkpoints of 3 fields + slabs,
// ck_every q fields, 3 adjoint fields, 2 x slabs, one gradient per shot
// group.
// ---------------------------------------------------------------------
namespace {

constexpr int kAdjBlockY = 4;  // wavefronts per block of the reverse step
// Blocks one shot group of the reverse step should at least fill: 8
// blocks of 4 wavefronts per CU on 256 CUs.
constexpr int64_t kAdjTargetBlocks = 2048;

// Launch shape of the reverse step: blocks per shot group (x tiles x row
// groups x interior planes) and the shots per thread.  Fewer groups mean
// fewer gradient accumulators and fewer passes over fact; the automatic
// choice splits the shots only as far as the grid needs to fill the chip.
struct AdjShape {
    int tiles_x, tiles_y;
    int64_t blocks_per_group;
    int group, ngroups;
    // blocks with CPML nodes (classes 1..7) walk fewer shots per thread:
    // they are few and carry the slab loads, so the launch needs more of
    // them in flight; the gradient has ngroups_cp >= ngroups accumulators
    int group_cp, ngroups_cp;
    // block classes (bit a: the tiles of axis a holding CPML nodes); per
    // axis the tiles and the range of those without CPML nodes
    int ntiles[3], int_lo[3], int_hi[3];
    int64_t class_blocks[8];  // all shot groups
};

// Interior tile range [*lo, *hi) of one axis from per-tile flags (set: the
// tile holds CPML nodes).  The flagged tiles must lie before and after the
// range; otherwise the whole axis counts as flagged (lo = hi = 0), which
// runs every tile with the CPML terms.
void interior_range(const std::vector<char>& slab, int* lo, int* hi) {
    const int n = (int)slab.size();
    int a = 0;
    while (a < n && slab[a]) ++a;
    int b = n;
    while (b > a && slab[b - 1]) --b;
    for (int t = a; t < b; ++t)
        if (slab[t]) {
            *lo = *hi = 0;
            return;
        }
    *lo = a;
    *hi = b;
}

// Whether node c of axis a lies in the range 0 <= m <= halo+1 of an
// absorbing side, where the reverse step has slab terms (adj_coef).
bool adj_node_ext(const awcpml_plan* P, int a, int64_t c) {
    const int64_t h = P->halo, n = P->n[a];
    return (P->active[a][0] && c <= h + 1) || (P->active[a][1] && c >= n - h - 2);
}

AdjShape adj_shape(const awcpml_plan* P) {
    AdjShape S;
    S.tiles_x = (int)(round_up(P->n[0], awcpml::kTileX) / awcpml::kTileX);
    S.tiles_y = (int)((P->n[1] - 2 + kAdjBlockY - 1) / kAdjBlockY);
    S.blocks_per_group = (int64_t)S.tiles_x * S.tiles_y * (P->ndim == 3 ? P->n[2] - 2 : 1);
    int g = P->adj_group;
    if (g <= 0) {
        const int64_t want = (kAdjTargetBlocks + S.blocks_per_group - 1) / S.blocks_per_group;
        const int ng = (int)std::max<int64_t>(1, std::min<int64_t>(P->nshots, want));
        g = (P->nshots + ng - 1) / ng;
    }
    S.group = std::min(g, P->nshots);
    S.ngroups = (P->nshots + S.group - 1) / S.group;
    S.group_cp = P->adj_group > 0 ? S.group : std::max(1, (S.group + 3) / 4);
    S.ngroups_cp = (P->nshots + S.group_cp - 1) / S.group_cp;
    // tiles: x tiles of kTileX lanes (nodes 1..nx-2), row groups of
    // kAdjBlockY rows from row 1, planes 1..nz-2 (3D)
    S.ntiles[0] = S.tiles_x;
    S.ntiles[1] = S.tiles_y;
    S.ntiles[2] = P->ndim == 3 ? (int)(P->n[2] - 2) : 1;
    for (int a = 0; a < 3; ++a) {
        std::vector<char> flag(S.ntiles[a], 0);
        if (a < P->ndim)
            for (int t = 0; t < S.ntiles[a]; ++t) {
                const int64_t c0 = a == 0 ? (int64_t)t * awcpml::kTileX
                                          : (a == 1 ? 1 + (int64_t)t * kAdjBlockY : 1 + t);
                const int64_t c1 = a == 0 ? c0 + awcpml::kTileX : (a == 1 ? c0 + kAdjBlockY : c0 + 1);
                for (int64_t c = std::max<int64_t>(c0, 1); c < c1 && c <= P->n[a] - 2 && !flag[t]; ++c)
                    flag[t] = adj_node_ext(P, a, c);
            }
        interior_range(flag, &S.int_lo[a], &S.int_hi[a]);
    }
    for (int c = 0; c < 8; ++c) {
        int64_t nb = c ? S.ngroups_cp : S.ngroups;
        for (int a = 0; a < 3; ++a) {
            const int in = S.int_hi[a] - S.int_lo[a];
            nb *= ((c >> a) & 1) ? S.ntiles[a] - in : in;
        }
        S.class_blocks[c] = nb;
    }
    return S;
}

struct CkLayout {
    size_t field;            // bytes of one field buffer (all shots)
    size_t psi[3], xi[3];    // bytes of one side's slab (all shots)
    size_t total;
};

CkLayout ck_layout(const awcpml_plan* P) {
    CkLayout L{};
    L.field = (size_t)P->field_elems * P->nshots * P->esize;
    L.total = 3 * L.field;
    for (int a = 0; a < P->ndim; ++a) {
        L.psi[a] = (size_t)P->psi_elems[a] * P->nshots * P->esize;
        L.xi[a] = (size_t)P->xi_elems[a] * P->nshots * P->esize;
        L.total += 2 * (L.psi[a] + L.xi[a]);
    }
    return L;
}

// copy the plan state to (save) or from (restore) checkpoint c
int ck_copy(awcpml_plan* P, int64_t c, bool save) {
    const CkLayout L = ck_layout(P);
    char* base = (char*)P->ck + (size_t)c * L.total;
    auto mv = [&](void* state, size_t bytes) -> int {
        if (bytes == 0) return AWCPML_OK;
        void* dst = save ? (void*)base : state;
        const void* src = save ? (const void*)state : (const void*)base;
        HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, P->stream));
        base += bytes;
        return AWCPML_OK;
    };
    int st;
    const int par = save ? P->par : P->ck_par[c];
    for (int b = 0; b < 3; ++b)
        if ((st = mv(P->p[b], L.field))) return st;
    for (int a = 0; a < P->ndim; ++a)
        for (int s = 0; s < 2; ++s) {
            if ((st = mv(P->psi[a][s][par], L.psi[a]))) return st;
            if ((st = mv(P->xi[a][s], L.xi[a]))) return st;
        }
    if (save) {
        P->ck_rot[c] = P->rot;
        P->ck_par[c] = P->par;
    } else {
        P->rot = P->ck_rot[c];
        P->par = P->ck_par[c];
        P->it = c * P->ck_every;
    }
    return AWCPML_OK;
}

int check_adjoint_shape(const awcpml_plan* P) {
    for (int a = 0; a < P->ndim; ++a)
        if ((P->active[a][0] || P->active[a][1]) && P->n[a] < 2 * (int64_t)P->halo + 4)
            return fail(AWCPML_ERR_ARG,
                        "adjoint needs >= 2*halo+4 = %d points on absorbing axis %d (has %lld)",
                        2 * P->halo + 4, a, (long long)P->n[a]);
    return AWCPML_OK;
}

// device buffers of one adjoint call
struct AdjBuffers {
    void* lam[3] = {};
    void* al[3][2][2] = {};
    void* be[3][2][2] = {};
    void* q = nullptr;
    void* grad = nullptr;
    void* grad64 = nullptr;
    void* tbar = nullptr;
    void* node_lin = nullptr;
    void* node_start = nullptr;
    void* rec_start = nullptr;
    void* rec_idx = nullptr;
    void* rq = nullptr;        // per receiver node entry: its gradient terms
    void* rf_lin = nullptr;    // distinct receiver nodes over all shots
    void* rf_start = nullptr;  // CSR offsets into rf_idx
    void* rf_idx = nullptr;    // entries of each distinct node, in shot order
    ~AdjBuffers() {
        for (auto& v : lam) dfree(v);
        for (int a = 0; a < 3; ++a)
            for (int s = 0; s < 2; ++s)
                for (int b = 0; b < 2; ++b) {
                    dfree(al[a][s][b]);
                    dfree(be[a][s][b]);
                }
        dfree(q);
        dfree(grad);
        dfree(grad64);
        dfree(tbar);
        dfree(node_lin);
        dfree(node_start);
        dfree(rec_start);
        dfree(rec_idx);
        dfree(rq);
        dfree(rf_lin);
        dfree(rf_start);
        dfree(rf_idx);
    }
};

// cp = false: the launch over block class 0 (no CPML nodes); true: over
// classes 1..7.
template <typename T>
awcpml::AdjointArgs<T> make_adj_args(const awcpml_plan* P, const AdjBuffers& Bf, const AdjShape& S,
                                     int lr, int par, const void* q, bool cp) {
    awcpml::AdjointArgs<T> A;
    memset(&A, 0, sizeof A);
    for (int a = 0; a < 3; ++a) A.n[a] = P->n[a];
    A.px = P->px;
    A.halo = P->halo;
    A.field_stride = P->field_elems;
    A.lam0 = (T*)Bf.lam[lr];
    A.lam1 = (const T*)Bf.lam[(lr + 1) % 3];
    A.lam2 = (const T*)Bf.lam[(lr + 2) % 3];
    A.fact = (const T*)P->fact;
    A.q = (const T*)q;
    A.grad = (T*)Bf.grad;
    for (int a = 0; a < 3; ++a) {
        A.inv_d[a] = (T)P->inv_d[a];
        A.inv_d2[a] = (T)P->inv_d2[a];
        A.psi_stride[a] = P->psi_elems[a];
        A.xi_stride[a] = P->xi_elems[a];
        for (int s = 0; s < 2; ++s) {
            A.active[a][s] = (a < P->ndim) ? P->active[a][s] : 0;
            if (a >= P->ndim) continue;
            awcpml::AdjSide<T>& S = A.side[a][s];
            S.al_rd = (const T*)Bf.al[a][s][par];
            S.al_wr = (T*)Bf.al[a][s][par ^ 1];
            S.be_rd = (const T*)Bf.be[a][s][par];
            S.be_wr = (T*)Bf.be[a][s][par ^ 1];
            S.a = (const T*)P->coef[a][0 + s];
            S.bK = (const T*)P->coef[a][2 + s];
            S.ah = (const T*)P->coef[a][4 + s];
            S.bKh = (const T*)P->coef[a][6 + s];
        }
    }
    A.ndim = P->ndim;
    A.nshots = P->nshots;
    A.group = cp ? S.group_cp : S.group;
    A.tiles_x = S.tiles_x;
    A.tiles_y = S.tiles_y;
    for (int a = 0; a < 3; ++a) {
        A.ntiles[a] = S.ntiles[a];
        A.int_lo[a] = S.int_lo[a];
        A.int_hi[a] = S.int_hi[a];
    }
    int64_t total = 0;
    A.cls_start[0] = 0;
    for (int c = 0; c < 8; ++c) {
        if (cp == (c > 0)) total += S.class_blocks[c];
        A.cls_start[c + 1] = cp ? (int)total : 0;
    }
    A.logical_blocks = (int)total;
    A.grid_blocks = (int)round_up(A.logical_blocks, awcpml::kNumXcd);
    return A;
}

template <typename T>
int adjoint_t(awcpml_plan* P, const double* tbar, double* fbar) {
    const int64_t nt = P->nt, K = P->ck_every;
    const size_t fb = (size_t)P->field_elems * P->nshots * sizeof(T);
    AdjBuffers Bf;
    int st;
    for (int b = 0; b < 3; ++b) {
        if ((st = dalloc(&Bf.lam[b], fb))) return st;
        HIP_TRY(hipMemsetAsync(Bf.lam[b], 0, fb, P->stream));
    }
    for (int a = 0; a < P->ndim; ++a)
        for (int s = 0; s < 2; ++s)
            for (int b = 0; b < 2; ++b) {
                const size_t pb = (size_t)P->psi_elems[a] * P->nshots * sizeof(T);
                const size_t xb = (size_t)P->xi_elems[a] * P->nshots * sizeof(T);
                if ((st = dalloc(&Bf.al[a][s][b], pb))) return st;
                if ((st = dalloc(&Bf.be[a][s][b], xb))) return st;
                HIP_TRY(hipMemsetAsync(Bf.al[a][s][b], 0, pb, P->stream));
                if (xb) HIP_TRY(hipMemsetAsync(Bf.be[a][s][b], 0, xb, P->stream));
            }
    if ((st = dalloc(&Bf.q, fb * (size_t)K))) return st;
    const AdjShape shape = adj_shape(P);
    if ((uint64_t)shape.blocks_per_group * shape.ngroups_cp > (uint64_t)INT32_MAX)
        return fail(AWCPML_ERR_ARG, "grid too large for the reverse step");
    const size_t gb = (size_t)P->field_elems * shape.ngroups_cp * sizeof(T);
    if ((st = dalloc(&Bf.grad, gb))) return st;
    HIP_TRY(hipMemsetAsync(Bf.grad, 0, gb, P->stream));
    const bool fold = P->grad_fold >= 0;
    const int64_t fold_every = P->grad_fold > 0 ? P->grad_fold : K;
    const int64_t nb_fold = (P->field_elems + 255) / 256;
    if (fold) {
        if ((st = dalloc(&Bf.grad64, (size_t)P->field_elems * 8))) return st;
        HIP_TRY(hipMemsetAsync(Bf.grad64, 0, (size_t)P->field_elems * 8, P->stream));
    }
    int rf_count = 0;  // distinct receiver nodes (set below)
    auto fold_rq = [&](auto* total) -> int {
        if (rf_count == 0) return AWCPML_OK;
        awcpml::fold_receivers<<<dim3((unsigned)((rf_count + 255) / 256)), dim3(256), 0, P->stream>>>(
            (T*)Bf.rq, total, (const int64_t*)Bf.rf_lin, (const int*)Bf.rf_start,
            (const int*)Bf.rf_idx, rf_count);
        HIP_TRY(hipGetLastError());
        return AWCPML_OK;
    };
    auto fold_now = [&]() -> int {
        awcpml::fold_gradient<T><<<dim3((unsigned)nb_fold), dim3(256), 0, P->stream>>>(
            (T*)Bf.grad, (double*)Bf.grad64, P->field_elems, shape.ngroups_cp);
        HIP_TRY(hipGetLastError());
        return fold_rq((double*)Bf.grad64);
    };

    // trace cotangents and the distinct interior receiver nodes per shot
    const size_t ntr = (size_t)P->n_rec_total * nt;
    if ((st = dalloc(&Bf.tbar, ntr * sizeof(T)))) return st;
    if ((st = upload(P, Bf.tbar, tbar, ntr))) return st;
    std::vector<int64_t> node_lin;
    std::vector<int> node_start(1, 0), rec_start(1, 0), rec_idx;
    int gr = 0;
    for (int s = 0; s < P->nshots; ++s) {
        const auto& R = P->rec_lin[s];
        std::vector<std::pair<int64_t, int>> v;
        for (size_t r = 0; r < R.size(); ++r) {
            const int64_t lin = R[r];
            const int64_t i = lin % P->px, j = (lin / P->px) % P->n[1], k = lin / (P->px * P->n[1]);
            bool inside = i >= 1 && i <= P->n[0] - 2 && j >= 1 && j <= P->n[1] - 2;
            if (P->ndim == 3) inside = inside && k >= 1 && k <= P->n[2] - 2;
            if (inside) v.push_back({lin, gr + (int)r});
        }
        gr += (int)R.size();
        std::stable_sort(v.begin(), v.end(),
                         [](const std::pair<int64_t, int>& x, const std::pair<int64_t, int>& y) {
                             return x.first < y.first;
                         });
        for (size_t e = 0; e < v.size(); ++e) {
            if (e == 0 || v[e].first != v[e - 1].first) {
                if (e) rec_start.push_back((int)rec_idx.size());
                node_lin.push_back(v[e].first);
            }
            rec_idx.push_back(v[e].second);
        }
        if (!v.empty()) rec_start.push_back((int)rec_idx.size());
        node_start.push_back((int)node_lin.size());
    }
    if ((st = dalloc(&Bf.node_lin, node_lin.size() * 8))) return st;
    if ((st = dalloc(&Bf.node_start, node_start.size() * 4))) return st;
    if ((st = dalloc(&Bf.rec_start, rec_start.size() * 4))) return st;
    if ((st = dalloc(&Bf.rec_idx, rec_idx.size() * 4))) return st;
    if (!node_lin.empty()) {
        HIP_TRY(hipMemcpyAsync(Bf.node_lin, node_lin.data(), node_lin.size() * 8, hipMemcpyHostToDevice, P->stream));
        HIP_TRY(hipMemcpyAsync(Bf.rec_idx, rec_idx.data(), rec_idx.size() * 4, hipMemcpyHostToDevice, P->stream));
    }
    HIP_TRY(hipMemcpyAsync(Bf.node_start, node_start.data(), node_start.size() * 4, hipMemcpyHostToDevice, P->stream));
    HIP_TRY(hipMemcpyAsync(Bf.rec_start, rec_start.data(), rec_start.size() * 4, hipMemcpyHostToDevice, P->stream));
    // the receiver gradient terms: one accumulator per entry of node_lin,
    // folded per distinct node over all shots in shot order
    std::vector<int64_t> rf_lin;
    std::vector<int> rf_start(1, 0), rf_idx;
    {
        std::vector<std::pair<int64_t, int>> v;
        for (size_t u = 0; u < node_lin.size(); ++u) v.push_back({node_lin[u], (int)u});
        std::stable_sort(v.begin(), v.end(),
                         [](const std::pair<int64_t, int>& x, const std::pair<int64_t, int>& y) {
                             return x.first < y.first;
                         });
        for (size_t e = 0; e < v.size(); ++e) {
            if (e && v[e].first != v[e - 1].first) rf_start.push_back((int)rf_idx.size());
            if (e == 0 || v[e].first != v[e - 1].first) rf_lin.push_back(v[e].first);
            rf_idx.push_back(v[e].second);
        }
        if (!v.empty()) rf_start.push_back((int)rf_idx.size());
    }
    rf_count = (int)rf_lin.size();
    if (rf_count) {
        if ((st = dalloc(&Bf.rq, node_lin.size() * sizeof(T)))) return st;
        HIP_TRY(hipMemsetAsync(Bf.rq, 0, node_lin.size() * sizeof(T), P->stream));
        if ((st = dalloc(&Bf.rf_lin, rf_lin.size() * 8))) return st;
        if ((st = dalloc(&Bf.rf_start, rf_start.size() * 4))) return st;
        if ((st = dalloc(&Bf.rf_idx, rf_idx.size() * 4))) return st;
        HIP_TRY(hipMemcpyAsync(Bf.rf_lin, rf_lin.data(), rf_lin.size() * 8, hipMemcpyHostToDevice, P->stream));
        HIP_TRY(hipMemcpyAsync(Bf.rf_start, rf_start.data(), rf_start.size() * 4, hipMemcpyHostToDevice, P->stream));
        HIP_TRY(hipMemcpyAsync(Bf.rf_idx, rf_idx.data(), rf_idx.size() * 4, hipMemcpyHostToDevice, P->stream));
    }

    awcpml::AdjIoArgs<T> io;
    memset(&io, 0, sizeof io);
    io.fact = (const T*)P->fact;
    io.field_stride = P->field_elems;
    io.node_lin = (const int64_t*)Bf.node_lin;
    io.node_start = (const int*)Bf.node_start;
    io.rec_start = (const int*)Bf.rec_start;
    io.rec_idx = (const int*)Bf.rec_idx;
    io.tbar = (const T*)Bf.tbar;
    io.nt = nt;
    for (int a = 0; a < 3; ++a) io.n[a] = P->n[a];
    io.px = P->px;
    io.halo = P->halo;
    io.ndim = P->ndim;
    io.rq = (T*)Bf.rq;

    int lr = 0, par = 0;
    for (int64_t seg = P->ck_count - 1; seg >= 0; --seg) {
        const int64_t c0 = seg * K, c1 = std::min<int64_t>(c0 + K, nt);
        if ((st = ck_copy(P, seg, false))) return st;
        for (int64_t s = c0 + 1; s <= c1;) {  // replay, storing q^s
            P->q_cur = (char*)Bf.q + (size_t)(s - c0 - 1) * fb;
            P->q_next = s + 1 <= c1 ? (char*)Bf.q + (size_t)(s - c0) * fb : nullptr;
            const int64_t n = P->q_next && two_ok(P) ? 2 : 1;  // bit-identical q either way
            if (n == 1) P->q_next = nullptr;
            st = n == 2 ? do_step2(P, s - 1) : do_step(P, s - 1);
            P->q_cur = P->q_next = nullptr;
            if (st) return st;
            P->it += n;
            s += n;
        }
        for (int64_t s = c1; s > c0; --s) {
            const void* q = (const char*)Bf.q + (size_t)(s - c0 - 1) * fb;
            awcpml::AdjointArgs<T> A;
            for (int cp = 0; cp < 2; ++cp) {  // blocks without, then with CPML nodes
                A = make_adj_args<T>(P, Bf, shape, lr, par, q, cp != 0);
                if (A.logical_blocks == 0) continue;
                const dim3 grid(A.grid_blocks), block(awcpml::kTileX, kAdjBlockY);
                if (P->ndim == 3) {
                    if (cp) awcpml::adjoint_step<T, 3, true><<<grid, block, 0, P->stream>>>(A);
                    else awcpml::adjoint_step<T, 3, false><<<grid, block, 0, P->stream>>>(A);
                } else {
                    if (cp) awcpml::adjoint_step<T, 2, true><<<grid, block, 0, P->stream>>>(A);
                    else awcpml::adjoint_step<T, 2, false><<<grid, block, 0, P->stream>>>(A);
                }
                HIP_TRY(hipGetLastError());
            }
            if (!node_lin.empty()) {
                io.lam0 = A.lam0;
                io.q = (const T*)q;
                for (int a = 0; a < 3; ++a)
                    for (int sd = 0; sd < 2; ++sd) {
                        io.side[a][sd] = A.side[a][sd];
                        io.active[a][sd] = A.active[a][sd];
                        io.xi_stride[a] = A.xi_stride[a];
                    }
                awcpml::adjoint_inject<T><<<dim3(P->nshots), dim3(256), 0, P->stream>>>(io, s - 1);
                HIP_TRY(hipGetLastError());
            }
            lr = (lr + 2) % 3;
            par ^= 1;
            if (fold && (s == c0 + 1 || (c1 - s + 1) % fold_every == 0) && (st = fold_now())) return st;
        }
    }
    const int64_t nx = P->n[0], rows = P->n[1] * P->n[2];
    if (fold) {
        std::vector<double> host((size_t)P->field_elems);
        HIP_TRY(hipMemcpyAsync(host.data(), Bf.grad64, host.size() * 8, hipMemcpyDeviceToHost, P->stream));
        HIP_TRY(hipStreamSynchronize(P->stream));
        for (int64_t r = 0; r < rows; ++r)
            for (int64_t i = 0; i < nx; ++i) fbar[r * nx + i] = host[(size_t)(r * P->px + i)];
        return AWCPML_OK;
    }
    if (shape.ngroups_cp > 1) {
        const int64_t nb = (P->field_elems + 255) / 256;
        awcpml::sum_shots<T><<<dim3((unsigned)nb), dim3(256), 0, P->stream>>>((T*)Bf.grad, P->field_elems,
                                                                            shape.ngroups_cp);
        HIP_TRY(hipGetLastError());
    }
    if ((st = fold_rq((T*)Bf.grad))) return st;
    std::vector<T> host((size_t)P->field_elems);
    HIP_TRY(hipMemcpyAsync(host.data(), Bf.grad, host.size() * sizeof(T), hipMemcpyDeviceToHost, P->stream));
    HIP_TRY(hipStreamSynchronize(P->stream));
    for (int64_t r = 0; r < rows; ++r)
        for (int64_t i = 0; i < nx; ++i) fbar[r * nx + i] = (double)host[(size_t)(r * P->px + i)];
    return AWCPML_OK;
}

}  // namespace

extern "C" {

int awcpml_checkpoint_run(awcpml_plan* P, int64_t every) {
    g_err.clear();
    int st = check_plan(P);
    if (st) return st;
    if ((st = ready(P))) return st;
    if ((st = check_adjoint_shape(P))) return st;
    if (P->nt < 1) return fail(AWCPML_ERR_STATE, "sources not set (nt unknown)");
    if (every <= 0) every = std::max<int64_t>(1, (int64_t)llround(sqrt((double)P->nt)));
    every = std::min<int64_t>(every, P->nt);
    const int64_t count = (P->nt + every - 1) / every;
    const CkLayout L = ck_layout(P);
    const size_t need = L.total * (size_t)count;
    if (need > P->ck_alloc) {
        dfree(P->ck);
        P->ck_alloc = 0;
        if ((st = dalloc(&P->ck, need))) return st;
        P->ck_alloc = need;
    }
    P->ck_bytes = L.total;
    P->ck_every = every;
    P->ck_count = count;
    P->ck_rot.assign(count, 0);
    P->ck_par.assign(count, 0);
    P->ck_valid = false;
    if ((st = awcpml_reset(P))) return st;
    for (int64_t c = 0; c < count; ++c) {
        if ((st = ck_copy(P, c, true))) return st;
        const int64_t n = std::min<int64_t>(every, P->nt - P->it);
        // two steps per launch where the plan takes them (bit-identical)
        if ((st = enqueue_steps(P, P->it, n))) return st;
        P->it += n;
    }
    HIP_TRY(hipStreamSynchronize(P->stream));
    P->ck_valid = true;
    return AWCPML_OK;
}

int awcpml_set_gradient_fold(awcpml_plan* P, int64_t steps) {
    g_err.clear();
    int st = check_plan(P);
    if (st) return st;
    if (steps < -1) return fail(AWCPML_ERR_ARG, "steps must be >= -1 (got %lld)", (long long)steps);
    P->grad_fold = steps;
    return AWCPML_OK;
}

int awcpml_set_adjoint_group(awcpml_plan* P, int shots) {
    g_err.clear();
    int st = check_plan(P);
    if (st) return st;
    if (shots < 0 || shots > P->nshots)
        return fail(AWCPML_ERR_ARG, "shots per thread must be in 0..%d (got %d)", P->nshots, shots);
    P->adj_group = shots;
    return AWCPML_OK;
}

int awcpml_get_adjoint_group(const awcpml_plan* P, int* shots) {
    g_err.clear();
    int st = check_plan(P);
    if (st) return st;
    if (!shots) return fail(AWCPML_ERR_ARG, "NULL argument");
    *shots = adj_shape(P).group;
    return AWCPML_OK;
}

int awcpml_adjoint(awcpml_plan* P, const double* trace_bar, int64_t count, double* fact_bar) {
    g_err.clear();
    int st = check_plan(P);
    if (st) return st;
    if (!P->ck_valid) return fail(AWCPML_ERR_STATE, "no valid checkpoints: call awcpml_checkpoint_run first");
    if (!fact_bar || (!trace_bar && count)) return fail(AWCPML_ERR_ARG, "NULL argument");
    if (count != P->n_rec_total * P->nt)
        return fail(AWCPML_ERR_ARG, "trace cotangent has %lld values, expected %lld (receivers x nt)",
                    (long long)count, (long long)(P->n_rec_total * P->nt));
    st = P->dtype == AWCPML_F64 ? adjoint_t<double>(P, trace_bar, fact_bar)
                                : adjoint_t<float>(P, trace_bar, fact_bar);
    P->q_cur = P->q_next = nullptr;
    return st;
}

}  // extern "C"

// ---------------------------------------------------------------------
// Allocation-free one-shot solve (awcpml_run_noalloc).  The wavefields,
// psi/xi and the source/receiver tables live in `scratch`, a file-scope
// device array of AWCPML_NOALLOC_SCRATCH_BYTES; the model, coefficients,
// wavelets, traces and final field are the caller's device buffers.  The
// steps are the plan path's own single-step launches (do_step on a plan
// description that owns nothing), so the results are those of awcpml_run
// on a one-shot plan with one step per launch.
// ---------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[AWCPML_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxScratchDevices] = {};  // scratch's address on each device

// The current device's copy of `scratch`, looked up on the first call on
// that device (with it current).  In host code the array's name is only a
// stand-in for the symbol; every use goes through this device address.
int scratch_address(void** out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices)
        return fail(AWCPML_ERR_ARG, "device %d: the scratch table holds %d devices", dev, kMaxScratchDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* ptr = nullptr;
        HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(AWCPML_ERR_HIP, "hipGetSymbolAddress gave no address for the scratch array");
        g_scratch_dev[dev] = ptr;
    }
    *out = g_scratch_dev[dev];
    return AWCPML_OK;
}

// Zero n 16-byte words (a kernel rather than hipMemsetAsync, so that a
// caller may capture the call in a graph).
__global__ void zero_words(uint4* p, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride)
        p[q] = make_uint4(0u, 0u, 0u, 0u);
}

// Up to 256 values per launch, carried in the kernel argument (no host
// memory is referenced after the launch is enqueued).
struct Put64 {
    int64_t v[256];
};
__global__ void put_i64(int64_t* dst, int count, Put64 a) {
    if ((int)threadIdx.x < count) dst[threadIdx.x] = a.v[threadIdx.x];
}

// CSR bounds of the one shot: src_start = {0, nsrcs}, src_seq = 1 (the
// sources are added by one thread in source order, exact for any number of
// sources sharing a node), rec_start = {0, nrecs}.
__global__ void put_io_bounds(int* meta, int nsrcs, int nrecs) {
    if (threadIdx.x == 0) {
        meta[0] = 0;
        meta[1] = nsrcs;
        meta[2] = 1;
        meta[3] = 0;
        meta[4] = nrecs;
    }
}

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

// Linear element offsets of `count` 1-based positions into dst (device).
int upload_lin(const awcpml_plan& P, const int64_t* pos, int count, int64_t* dst, hipStream_t s) {
    for (int c0 = 0; c0 < count; c0 += 256) {
        Put64 a;
        const int cnt = std::min(256, count - c0);
        for (int q = 0; q < cnt; ++q) a.v[q] = lin_index(&P, pos, count, c0 + q);
        put_i64<<<dim3(1), dim3(256), 0, s>>>(dst + c0, cnt, a);
        HIP_TRY(hipGetLastError());
    }
    return AWCPML_OK;
}

}  // namespace

extern "C" {

int awcpml_run_noalloc(int ndim, int64_t nx, int64_t ny, int64_t nz, int halo, int dtype,
                       const void* fact, const double* inv_d, const double* inv_d2,
                       const void* const* coef, const int* active, int nsrcs,
                       const int64_t* srcpos, int64_t nt, const void* dt2srctf, int nrecs,
                       const int64_t* recpos, void* traces, void* pend, void* stream) {
    g_err.clear();
    if (ndim != 2 && ndim != 3) return fail(AWCPML_ERR_ARG, "ndim must be 2 or 3 (got %d)", ndim);
    if (dtype != AWCPML_F64 && dtype != AWCPML_F32) return fail(AWCPML_ERR_ARG, "unknown dtype %d", dtype);
    if (halo < 0) return fail(AWCPML_ERR_ARG, "halo must be >= 0");
    if (nsrcs < 0 || nrecs < 0 || nt < 0) return fail(AWCPML_ERR_ARG, "negative count");
    if (ndim == 2) nz = 1;
    const int64_t n[3] = {nx, ny, nz};
    for (int a = 0; a < ndim; ++a)
        if (n[a] < 3 || n[a] < (int64_t)halo + 2)
            return fail(AWCPML_ERR_ARG, "axis %d has %lld points (need >= 3 and >= halo+2 = %d)", a,
                        (long long)n[a], halo + 2);
    if (!fact || !inv_d || !inv_d2) return fail(AWCPML_ERR_ARG, "NULL model argument");
    if (halo > 0 && !coef) return fail(AWCPML_ERR_ARG, "coef is NULL");
    if (nsrcs > 0 && (!srcpos || (nt > 0 && !dt2srctf))) return fail(AWCPML_ERR_ARG, "NULL source argument");
    if (nrecs > 0 && (!recpos || (nt > 0 && !traces))) return fail(AWCPML_ERR_ARG, "NULL receiver argument");

    // a plan description that owns nothing: every pointer is the caller's
    // or inside scratch (its destructor frees nothing)
    awcpml_plan P;
    P.ndim = ndim;
    P.dtype = dtype;
    P.nshots = 1;
    P.halo = halo;
    for (int a = 0; a < 3; ++a) P.n[a] = n[a];
    P.esize = dtype == AWCPML_F64 ? 8 : 4;
    P.px = nx;  // the caller's fact is dense, so the fields are too
    P.field_elems = nx * ny * nz;
    P.stream = (hipStream_t)stream;
    int st;
    if ((st = check_positions(&P, srcpos, nsrcs, "source"))) return st;
    if ((st = check_positions(&P, recpos, nrecs, "receiver"))) return st;

    // scratch layout: io bounds, source and receiver offsets, 3 fields,
    // psi (2 copies) and xi per absorbing side, each 256-byte aligned
    const size_t fbytes = (size_t)P.field_elems * P.esize;
    size_t off = 0;
    const size_t o_meta = off;
    off += 256;
    const size_t o_src = off;
    off += align256((size_t)nsrcs * 8);
    const size_t o_rec = off;
    off += align256((size_t)nrecs * 8);
    const size_t o_state = off;
    size_t o_p[3];
    for (int b = 0; b < 3; ++b) {
        o_p[b] = off;
        off += align256(fbytes);
    }
    size_t o_psi[3][2][2] = {}, o_xi[3][2] = {};
    for (int a = 0; a < ndim; ++a) {
        P.psi_elems[a] = slab_elems(&P, a, halo + 1);
        P.xi_elems[a] = slab_elems(&P, a, halo);
        for (int s = 0; s < 2; ++s) {
            for (int c = 0; c < 2; ++c) {
                o_psi[a][s][c] = off;
                off += align256((size_t)P.psi_elems[a] * P.esize);
            }
            o_xi[a][s] = off;
            off += align256((size_t)P.xi_elems[a] * P.esize);
        }
    }
    if (off > (size_t)AWCPML_NOALLOC_SCRATCH_BYTES)
        return fail(AWCPML_ERR_ARG,
                    "this problem needs %zu bytes of scratch; the library's allocation-free array holds "
                    "%zu (use a plan, awcpml_plan_create, for larger grids)",
                    off, (size_t)AWCPML_NOALLOC_SCRATCH_BYTES);

    void* base = nullptr;
    if ((st = scratch_address(&base))) return st;
    char* sc = (char*)base;
    for (int b = 0; b < 3; ++b) P.p[b] = sc + o_p[b];
    P.fact = const_cast<void*>(fact);
    for (int a = 0; a < ndim; ++a) {
        P.inv_d[a] = inv_d[a];
        P.inv_d2[a] = inv_d2[a];
        for (int s = 0; s < 2; ++s) {
            P.psi[a][s][0] = sc + o_psi[a][s][0];
            P.psi[a][s][1] = sc + o_psi[a][s][1];
            P.xi[a][s] = sc + o_xi[a][s];
            P.active[a][s] = halo > 0 && (!active || active[2 * a + s]) ? 1 : 0;
        }
        for (int c = 0; c < 8; ++c) {
            P.coef[a][c] = halo > 0 ? const_cast<void*>(coef[8 * a + c]) : nullptr;
            if (halo > 0 && !P.coef[a][c])
                return fail(AWCPML_ERR_ARG, "coefficient vector %d of axis %d is NULL", c, a);
        }
    }
    int* meta = (int*)(sc + o_meta);
    P.d_src_lin = sc + o_src;
    P.d_src_wav = const_cast<void*>(dt2srctf);
    P.d_src_start = meta;
    P.d_src_seq = meta + 2;
    P.d_rec_lin = sc + o_rec;
    P.d_rec_start = meta + 3;
    P.d_traces = traces;
    P.nt = nt;
    P.n_src_total = nsrcs;
    P.n_rec_total = nrecs;
    P.has_io = nsrcs + nrecs > 0;
    P.io_dirty = false;
    if ((st = default_tiling(&P))) return st;

    hipStream_t s = P.stream;
    put_io_bounds<<<dim3(1), dim3(64), 0, s>>>(meta, nsrcs, nrecs);
    HIP_TRY(hipGetLastError());
    if ((st = upload_lin(P, srcpos, nsrcs, (int64_t*)P.d_src_lin, s))) return st;
    if ((st = upload_lin(P, recpos, nrecs, (int64_t*)P.d_rec_lin, s))) return st;
    // the previous call's state: fields, psi and xi start at zero
    zero_words<<<dim3(1024), dim3(256), 0, s>>>((uint4*)(sc + o_state), (off - o_state) / 16);
    HIP_TRY(hipGetLastError());
    for (int64_t it = 0; it < nt; ++it)
        if ((st = do_step(&P, it))) return st;
    if (pend)
        HIP_TRY(hipMemcpyAsync(pend, P.p[(P.rot + 1) % 3], fbytes, hipMemcpyDeviceToDevice, s));
    if (!stream) HIP_TRY(hipStreamSynchronize(nullptr));
    return AWCPML_OK;
}

}  // extern "C"
