// Launch interface of the HIP kernels (host side).
#pragma once

#include <hip/hip_runtime.h>

#include "plan_types.hpp"

namespace tblis_amd
{

struct GemmLaunch
{
    const TaskDesc* tasks;
    int ntasks;
    const SegDesc* segs;
    const int64_t* pool;
    const void* A;
    const void* B;
    void* C;
    double alpha, beta;
    int use_k;  // 0: ignore k segments (alpha == 0 or k == 0)
    int grid;
    hipStream_t stream;
};

int gemm_f64_num_tiles();
int gemm_f64_tile_m(int tile);
int gemm_f64_tile_k(int tile);
int gemm_f64_tile_n(int tile);
hipError_t launch_gemm_f64(const GemmLaunch& L, const KernelKey& k);
int gemm_f64_occupancy(const KernelKey& k);
hipError_t launch_mfma_probe(double* out, hipStream_t s);

}  // namespace tblis_amd
