// Launch interface of the HIP kernels (host side).
#pragma once

#include <hip/hip_runtime.h>

#include "plan_types.hpp"

namespace tblis_amd
{

// The chunk a pack workgroup takes: workgroup b of a launch runs on XCD
// b % 8 (round-robin dispatch), and XCD x takes the x-th contiguous eighth
// of the n chunks, so that neighbouring chunks, which share the cache lines
// at their edges, read them through one L2.
__device__ __forceinline__ int xcd_contiguous_chunk(int b, int n)
{
    constexpr int kX = 8;
    const int x = b % kX, i = b / kX, per = n / kX, rem = n % kX;
    return x * per + (x < rem ? x : rem) + i;
}

// Lane-layout probes of the two MFMA forms the GEMMs use (one wave each).
hipError_t launch_mfma_probe(double* out, hipStream_t s);
hipError_t launch_bf16_probe(const float* A, const float* B, float* D, hipStream_t s);

// Panel path (FP64 and split-BF16 families)
int panel_num_tiles(int family);
int panel_tile_m(int family, int tile);
int panel_tile_n(int family, int tile);
int panel_tile_waves(int family, int tile);
// FP64 tiles that stage both operands through LDS (every operand read from
// the tensors, single-block GEMMs only)
bool panel_tile_lds(int family, int tile);
struct PanelLaunch
{
    const PackChunk* chunks;
    int nchunks;
    const PanelTask* tasks;
    int ntasks;
    const UnpackJob* ujobs;
    const UnpackChunk* uchunks;
    int nuchunks;
    const int64_t* pool;
    const void* A;
    const void* B;
    void* C;
    void* ws;     // panels (double or BF16 values)
    double* cws;  // C panels
    double alpha, beta;
    int use_k;
    int dtype;   // 0: double, 1: float
    int family;  // FAMILY_F64 or a split-BF16 family
    int tile;
    bool direct_a, direct_b;  // FP64: operand read from the tensor (PanelTask)
    hipStream_t stream;
};
hipError_t launch_panel_pack(const PanelLaunch& L);
hipError_t launch_panel_gemm(const PanelLaunch& L);
hipError_t launch_panel_unpack(const PanelLaunch& L);
// n pseudo-random values in [-1, 1) (double, or BF16 when bf16) at p
hipError_t launch_fill_random(void* p, int64_t n, bool bf16, hipStream_t s);
// split-BF16 panels (gemm_panel_bf16.hip)
int panel_bf16_num_tiles();
int panel_bf16_tile_m(int tile);
int panel_bf16_tile_n(int tile);
int panel_bf16_tile_waves(int tile);
bool panel_bf16_tile_fast_only(int tile);
bool panel_bf16_tile_direct(int tile);  // has variants reading FP32 operands directly
// FP64 GEMM reading both operands from the tensors through LDS
// (gemm_lds.hip); its tiles follow the register-fed FP64 tiles.
int lds_num_tiles();
int lds_tile_m(int lt);
int lds_tile_n(int lt);
hipError_t launch_gemm_lds(const PanelLaunch& L, int lt);
// Workgroups of a persistent launch of `kernel` over ntasks tasks (kernels
// loop over tasks with a stride of the grid size).
int persistent_grid(const void* kernel, int block, int ntasks);
hipError_t launch_panel_pack_bf16(const PanelLaunch& L);
// Ozaki family (gemm_ozaki.hip): row maxima, digit planes and the int8 GEMM
// in one call; rmax holds nrows panel-row maxima in the workspace.
hipError_t launch_panel_ozaki(const PanelLaunch& L, int slices, unsigned long long* rmax, int64_t nrows);
int ozaki_slices();
hipError_t launch_panel_gemm_bf16(const PanelLaunch& L);

}  // namespace tblis_amd
