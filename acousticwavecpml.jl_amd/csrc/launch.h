// launch.h -- host-side launchers of the stencil kernels.  Each (precision,
// dimension, plan / batch / two-step batch) combination is instantiated in
// its own object file (launch_kernels.hip, compiled once per combination by
// build.py), so the heavily templated kernels compile in parallel.
#pragma once

#include <hip/hip_runtime.h>

#include "stencil2_kernels.h"
#include "stencil_kernels.h"

namespace awcpml {

struct LaunchShape {
    dim3 grid, block;
    size_t lds;          // dynamic LDS bytes (march-axis CPML coefficients)
    hipStream_t stream;
};

// pipe: march pipeline (1..3 whole-node rings, 4..5 two-distance);
// off64: 64-bit byte offsets (fields of 4 GiB and more).
template <typename T, int NDIM>
void launch_plan_stencil(const StencilArgs<T>& A, const LaunchShape& L, int pipe, bool off64);

template <typename T, int NDIM>
void launch_batch_stencil(const BatchArgs<T>& A, const LaunchShape& L, int pipe, bool off64);

// Two time steps per launch; rows: waves per block (3D: 4, 8, 16 rows of a
// tile; 2D: 4 or 8 march chunks); pf: march positions of loads in flight
// per wave (1..3).
template <typename T, int NDIM>
void launch_plan_stencil2(const Stencil2Args<T>& B, const LaunchShape& L, int rows, int pf);

template <typename T, int NDIM>
void launch_batch_stencil2(const Batch2Args<T>& B, const LaunchShape& L, int rows, int pf);

}  // namespace awcpml
