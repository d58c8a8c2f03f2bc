// Level-1 tensor operations (add, scale, set): host entry points and the
// kernel descriptor.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "common.hpp"

namespace tblis_amd
{

constexpr int LEVEL1_MAXDIM = 16;
constexpr int LEVEL1_ADD = 0, LEVEL1_SCALE = 1, LEVEL1_SET = 2;

// B's iteration dimensions (fastest first) with B's and A's strides, and
// the A-only dimensions summed per element.
struct Level1Desc
{
    int ndb = 0, nda = 0;
    int64_t nb = 1, na = 1;  // element counts (product of lengths)
    int64_t len_b[LEVEL1_MAXDIM], sb[LEVEL1_MAXDIM], sa_b[LEVEL1_MAXDIM];
    int64_t len_a[LEVEL1_MAXDIM], sa_a[LEVEL1_MAXDIM];
};

// Beta pass over the entries of an indexed tensor (same dense shape d,
// entry offsets and factors in device memory).
hipError_t launch_level1_entries(int dtype, const Level1Desc& d, const int64_t* offs, const double* fac,
                                 int64_t nent, double beta, void* B, hipStream_t s);
// C_e <- (f_e s) C_e over the entries of an interleaved complex indexed
// tensor of real type dtype (d and offs in complex elements; f_e from
// fre / fim, or 1 when fre is null; s = sr + i si).
hipError_t launch_level1_centries(int dtype, const Level1Desc& d, const int64_t* offs, const double* fre,
                                  const double* fim, int64_t nent, double sr, double si, void* B, hipStream_t s);
// C <- (br + i bi) C over an interleaved complex tensor of real type dtype
// (d in complex elements).
hipError_t launch_level1_cscale(int dtype, const Level1Desc& d, double br, double bi, void* B, hipStream_t s);
// One strided block copied between two layouts (up to 8 dimensions):
// dst[dst + sum_d i_d ds[d]] = src[src + sum_d i_d ss[d]].
constexpr int BLOCK_COPY_MAXDIM = 8;
struct BlockCopy
{
    int64_t src, dst;     // element offsets of the block in each layout
    int64_t e0, n;        // this piece: elements [e0, e0 + n) of the block
    int32_t nd, pad_;
    int64_t len[BLOCK_COPY_MAXDIM], ss[BLOCK_COPY_MAXDIM], ds[BLOCK_COPY_MAXDIM];
};
// dtype: 0 double, 1 float, 2 dcomplex, 3 scomplex (offsets in elements).
hipError_t launch_block_copy(int dtype, const BlockCopy* tasks, int64_t ntasks, const void* src, void* dst,
                             hipStream_t s);
// Device workspace a call needs (partial sums of a split A-only sum), bytes.
size_t level1_workspace_bytes(int mode, const Level1Desc& d);
hipError_t launch_level1(int mode, int dtype, const Level1Desc& d, const void* A, void* B,
                         double alpha, double beta, double value, void* ws, hipStream_t s);

// tblis_tensor_add / _scale / _set (src/iface/1t/{add,scale,set}.cxx).
void tensor_add(const tblis_config* cfg, const tblis_tensor* A, const label_type* idx_A,
                tblis_tensor* B, const label_type* idx_B);
void tensor_scale(const tblis_config* cfg, tblis_tensor* A, const label_type* idx_A);
void tensor_set(const tblis_config* cfg, const tblis_scalar* alpha, tblis_tensor* A,
                const label_type* idx_A);

}  // namespace tblis_amd
