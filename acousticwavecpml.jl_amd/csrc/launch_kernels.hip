// launch_kernels.hip -- definitions of the launchers declared in launch.h,
// instantiated for AWCPML_REAL (float or double), AWCPML_NDIM (2 or 3) and
// AWCPML_BATCH (0: one plan, 1: a batch of plans, 2: a batch of plans, two
// steps per launch), given on the compiler command line (see build.py).
#include "launch.h"

#if !defined(AWCPML_REAL) || !defined(AWCPML_NDIM) || !defined(AWCPML_BATCH)
#error "compile with -DAWCPML_REAL=<float|double> -DAWCPML_NDIM=<2|3> -DAWCPML_BATCH=<0|1|2>"
#endif

namespace awcpml {

#if AWCPML_BATCH == 0
template <typename T, typename Off, int P, int NDIM>
static void go_plan(const StencilArgs<T>& A, const LaunchShape& L) {
    if constexpr (NDIM == 3) stencil3d<T, Off, P><<<L.grid, L.block, L.lds, L.stream>>>(A);
    else stencil2d<T, Off, P><<<L.grid, L.block, L.lds, L.stream>>>(A);
}

template <typename T, int NDIM>
void launch_plan_stencil(const StencilArgs<T>& A, const LaunchShape& L, int pipe, bool off64) {
    if (off64) {  // fields of 4 GiB and more: one pipeline shape
        go_plan<T, uint64_t, 2, NDIM>(A, L);
        return;
    }
    switch (pipe) {
        case 1: go_plan<T, uint32_t, 1, NDIM>(A, L); break;
        case 3: go_plan<T, uint32_t, 3, NDIM>(A, L); break;
        case 4: go_plan<T, uint32_t, 4, NDIM>(A, L); break;
        case 5: go_plan<T, uint32_t, 5, NDIM>(A, L); break;
        default: go_plan<T, uint32_t, 2, NDIM>(A, L); break;
    }
}
template void launch_plan_stencil<AWCPML_REAL, AWCPML_NDIM>(const StencilArgs<AWCPML_REAL>&,
                                                           const LaunchShape&, int, bool);
template <typename T, int W>
static void go_two(const Stencil2Args<T>& B, const LaunchShape& L, int pf) {
#if AWCPML_NDIM == 3
    switch (pf) {
        case 1: stencil3d_two<T, uint32_t, W, 1><<<L.grid, L.block, L.lds, L.stream>>>(B); break;
        case 3: stencil3d_two<T, uint32_t, W, 3><<<L.grid, L.block, L.lds, L.stream>>>(B); break;
        default: stencil3d_two<T, uint32_t, W, 2><<<L.grid, L.block, L.lds, L.stream>>>(B); break;
    }
#else
    switch (pf) {
        case 1: stencil2d_two<T, uint32_t, W, 1><<<L.grid, L.block, L.lds, L.stream>>>(B); break;
        case 3: stencil2d_two<T, uint32_t, W, 3><<<L.grid, L.block, L.lds, L.stream>>>(B); break;
        default: stencil2d_two<T, uint32_t, W, 2><<<L.grid, L.block, L.lds, L.stream>>>(B); break;
    }
#endif
}

template <typename T, int NDIM>
void launch_plan_stencil2(const Stencil2Args<T>& B, const LaunchShape& L, int rows, int pf) {
#if AWCPML_NDIM == 3
    switch (rows) {
        case 4: go_two<T, 4>(B, L, pf); break;
        case 16: go_two<T, 16>(B, L, pf); break;
        default: go_two<T, 8>(B, L, pf); break;
    }
#else
    switch (rows) {
        case 8: go_two<T, 8>(B, L, pf); break;
        default: go_two<T, 4>(B, L, pf); break;
    }
#endif
}
template void launch_plan_stencil2<AWCPML_REAL, AWCPML_NDIM>(const Stencil2Args<AWCPML_REAL>&,
                                                            const LaunchShape&, int, int);
#elif AWCPML_BATCH == 1
template <typename T, typename Off, int P, int NDIM>
static void go_batch(const BatchArgs<T>& A, const LaunchShape& L) {
    stencil_batch<T, Off, P, NDIM><<<L.grid, L.block, L.lds, L.stream>>>(A);
}

template <typename T, int NDIM>
void launch_batch_stencil(const BatchArgs<T>& A, const LaunchShape& L, int pipe, bool off64) {
    if (off64) {
        go_batch<T, uint64_t, 2, NDIM>(A, L);
        return;
    }
    switch (pipe) {
        case 1: go_batch<T, uint32_t, 1, NDIM>(A, L); break;
        case 3: go_batch<T, uint32_t, 3, NDIM>(A, L); break;
        case 4: go_batch<T, uint32_t, 4, NDIM>(A, L); break;
        case 5: go_batch<T, uint32_t, 5, NDIM>(A, L); break;
        default: go_batch<T, uint32_t, 2, NDIM>(A, L); break;
    }
}
template void launch_batch_stencil<AWCPML_REAL, AWCPML_NDIM>(const BatchArgs<AWCPML_REAL>&,
                                                            const LaunchShape&, int, bool);
#else
template <typename T, int W>
static void go_two_batch(const Batch2Args<T>& B, const LaunchShape& L, int pf) {
#if AWCPML_NDIM == 3
    switch (pf) {
        case 1: stencil3d_two_batch<T, uint32_t, W, 1><<<L.grid, L.block, L.lds, L.stream>>>(B); break;
        case 3: stencil3d_two_batch<T, uint32_t, W, 3><<<L.grid, L.block, L.lds, L.stream>>>(B); break;
        default: stencil3d_two_batch<T, uint32_t, W, 2><<<L.grid, L.block, L.lds, L.stream>>>(B); break;
    }
#else
    switch (pf) {
        case 1: stencil2d_two_batch<T, uint32_t, W, 1><<<L.grid, L.block, L.lds, L.stream>>>(B); break;
        case 3: stencil2d_two_batch<T, uint32_t, W, 3><<<L.grid, L.block, L.lds, L.stream>>>(B); break;
        default: stencil2d_two_batch<T, uint32_t, W, 2><<<L.grid, L.block, L.lds, L.stream>>>(B); break;
    }
#endif
}

template <typename T, int NDIM>
void launch_batch_stencil2(const Batch2Args<T>& B, const LaunchShape& L, int rows, int pf) {
#if AWCPML_NDIM == 3
    switch (rows) {
        case 4: go_two_batch<T, 4>(B, L, pf); break;
        case 16: go_two_batch<T, 16>(B, L, pf); break;
        default: go_two_batch<T, 8>(B, L, pf); break;
    }
#else
    switch (rows) {
        case 8: go_two_batch<T, 8>(B, L, pf); break;
        default: go_two_batch<T, 4>(B, L, pf); break;
    }
#endif
}
template void launch_batch_stencil2<AWCPML_REAL, AWCPML_NDIM>(const Batch2Args<AWCPML_REAL>&,
                                                             const LaunchShape&, int, int);
#endif

}  // namespace awcpml
