// FP64-accurate contraction from int8 MFMA (Ozaki scheme, gfx950).
//
// Every panel row x (a row of A' or of B', i.e. one m or n index over all
// k) is scaled by 2^-e, e the exponent of its largest magnitude, so that
// |x 2^-e| < 1, and cut into S signed 7-bit digits,
//     x 2^-e = sum_{t=1..S} d_t 2^(-7 t) + r,   d_t in [-127, 127],
// |r| < 2^(-7 S) (each step multiplies the remainder by 128 and takes its
// integer part; every step is exact in FP64).  Then
//     A_ik B_kj ~= 2^(ea_i + eb_j) sum_{t + u <= S + 1} 2^(-7 (t + u)) dA_t,ik dB_u,kj,
// the products of digit pairs whose weight is above the truncation level
// (Ozaki, Ogita, Oishi and Rump, "Error-free transformations of matrix
// multiplication by using fast routines of matrix multiplication and its
// applications", Numer. Algorithms 59 (2012); the int8 form follows Ootomo,
// Ozaki and Yokota, "DGEMM on integer matrix multiplication unit", IJHPCA
// 2024).  Pairs with the same t + u share one int32 accumulator:
// v_mfma_i32_32x32x32_i8 sums them exactly (at most 8 pairs * 127^2 * 8160 k
// < 2^31 between flushes), and every 8160 k the accumulators are added to
// an FP64 sum with their weights 2^(-7 (t + u)).
//
// Kernels, in launch order:
//   zero    clears the row-maximum array;
//   rowmax  per pack chunk, the largest |x| of each row over the chunk's k,
//           combined per row with a 64-bit atomic max on the bit pattern
//           (non-negative doubles order like their bits);
//   pack    per pack chunk, the S digit planes in the MFMA operand order;
//           x is the job's scale (complex signs, indexed entry factors)
//           times the source value, in both passes;
//   gemm    one 64 x 64 C tile per workgroup, 32 x 32 per wave, S int32
//           accumulators (one per digit diagonal), five k groups of digit
//           planes in flight, FP64 epilogue with the row and column scales,
//           alpha and beta.
//
// Panels (plan_types.hpp jobs and tasks): 32-row by 32-k fragment groups
// of S planes of 1 KiB, group (rb, kg) plane t at
//     out + ((rb * pnkg + kg) * S + t) * 1024 bytes,
// lane l holding digit t of X[32 rb + l % 32][32 kg + 16 (l / 32) + j],
// j = 0..15, as one 16-byte vector -- the A and B operand layout of
// v_mfma_i32_32x32x32_i8.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace tblis_amd
{

namespace
{

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// Exponent e of a row whose largest magnitude has the bit pattern `bits`:
// the smallest e with max < 2^e (0 for an all-zero row).
__device__ __forceinline__ int oz_exponent(unsigned long long bits)
{
    if (bits == 0) return 0;
    const int biased = (int)(bits >> 52);  // sign bit is clear
    return biased == 0 ? -1022 : biased - 1022;
}

__device__ __forceinline__ double oz_abs(double x) { return __builtin_fabs(x); }

// Source offset of element (group row lr, group k kk) of group (ri, ki) of
// chunk j, or -1 when it lies outside the job.
__device__ __forceinline__ int64_t oz_offset(const PackChunk& j, const int64_t* __restrict__ pool, int ri,
                                             int ki, int lr, int kk)
{
    const int rb = j.rb0 + ri, kg = j.kg0 + ki;
    const int prow = rb * 32 + lr - j.r0;
    const int pk = kg * 32 + kk - j.k0;
    if (prow < 0 || prow >= j.nr || pk < 0 || pk >= j.nk) return -1;
    const int64_t rstr = j.rt[ri][1], kstr = j.kt[ki][1];
    const int64_t roff = rstr != SCATTERED ? j.rt[ri][0] + lr * rstr : pool[j.rowmap + prow];
    const int64_t koff = kstr != SCATTERED ? j.kt[ki][0] + kk * kstr : pool[j.kmap + pk];
    return j.src + roff + koff;
}

__global__ __launch_bounds__(256) void oz_zero_kernel(unsigned long long* __restrict__ p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = 0;
}

// Rows-fast source: one 32 x 32 group per wave and iteration, lane l reads
// row l % 32, k 16 (l / 32) .. + 15 (consecutive lanes, consecutive rows);
// the two halves of a row meet through one shuffle and lanes 0..31 publish
// the row maxima.  K-fast source: lanes along k instead.
template <typename T>
__global__ __launch_bounds__(256) void oz_rowmax_kernel(const PackChunk* __restrict__ chunks,
                                                        const int64_t* __restrict__ pool,
                                                        const T* __restrict__ A, const T* __restrict__ B,
                                                        unsigned long long* __restrict__ rmax)
{
    const PackChunk& j = chunks[xcd_contiguous_chunk(blockIdx.x, gridDim.x)];
    if (j.src == PANEL_MISSING) return;
    const T* __restrict__ X = j.operand ? B : A;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (j.kfast)
    {
        // k-fast source: one row per wave and iteration, lanes along k
        // (whole cache lines), the row maximum by a shuffle reduction
        const int nrows = j.nrb * 32, kspan = j.nkg * 32;
        for (int r = wave; r < nrows; r += 4)
        {
            const int ri = r >> 5, lr = r & 31;
            double m = 0.0;
            for (int kk = lane; kk < kspan; kk += 64)
            {
                const int64_t off = oz_offset(j, pool, ri, kk >> 5, lr, kk & 31);
                if (off >= 0) m = fmax(m, oz_abs(j.scale * (double)X[off]));
            }
#pragma unroll
            for (int w = 32; w >= 1; w >>= 1) m = fmax(m, __shfl_xor(m, w));
            const int prow = j.rb0 * 32 + r;  // panel row
            if (lane == 0 && prow >= j.r0 && prow < j.r0 + j.nr && m > 0.0)
                atomicMax(rmax + j.eoff + prow, (unsigned long long)__double_as_longlong(m));
        }
        return;
    }
    const int lr = lane & 31, lh = lane >> 5;
    const int ng = j.nrb * j.nkg;
    for (int g = wave; g < ng; g += 4)
    {
        const int ri = g / j.nkg, ki = g % j.nkg;
        double m = 0.0;
#pragma unroll
        for (int e = 0; e < 16; e++)
        {
            const int64_t off = oz_offset(j, pool, ri, ki, lr, 16 * lh + e);
            if (off >= 0) m = fmax(m, oz_abs(j.scale * (double)X[off]));
        }
        m = fmax(m, __shfl_xor(m, 32));
        const int prow = (j.rb0 + ri) * 32 + lr;  // panel row
        if (lh == 0 && prow >= j.r0 && prow < j.r0 + j.nr && m > 0.0)
            atomicMax(rmax + j.eoff + prow, (unsigned long long)__double_as_longlong(m));
    }
}

template <typename T, int S>
__global__ __launch_bounds__(256) void oz_pack_kernel(const PackChunk* __restrict__ chunks,
                                                      const int64_t* __restrict__ pool,
                                                      const T* __restrict__ A, const T* __restrict__ B,
                                                      const unsigned long long* __restrict__ rmax,
                                                      int8_t* __restrict__ ws)
{
    const PackChunk& j = chunks[xcd_contiguous_chunk(blockIdx.x, gridDim.x)];
    const T* __restrict__ X = j.operand ? B : A;
    const bool zero = j.src == PANEL_MISSING;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lr = lane & 31, lh = lane >> 5;
    const int ng = j.nrb * j.nkg;
    for (int g = wave; g < ng; g += 4)
    {
        const int ri = g / j.nkg, ki = g % j.nkg;
        const int rb = j.rb0 + ri, kg = j.kg0 + ki;
        const int prow = rb * 32 + lr;
        const bool rin = prow >= j.r0 && prow < j.r0 + j.nr;
        const int e = rin && !zero ? oz_exponent(rmax[j.eoff + prow]) : 0;
        uint32_t w[S][4];  // digit planes, four digits per word (byte q % 4 of word q / 4)
#pragma unroll
        for (int t = 0; t < S; t++)
#pragma unroll
            for (int c = 0; c < 4; c++) w[t][c] = 0;
        unsigned inmask = 0;
#pragma unroll
        for (int q = 0; q < 16; q++)
        {
            const int pk = kg * 32 + 16 * lh + q;
            const bool in = rin && pk >= j.k0 && pk < j.k0 + j.nk;
            inmask |= (unsigned)in << q;
            double v = 0.0;
            if (in && !zero)
            {
                const int64_t off = oz_offset(j, pool, ri, ki, lr, 16 * lh + q);
                v = ldexp(j.scale * (double)X[off], -e);  // |v| < 1
            }
#pragma unroll
            for (int t = 0; t < S; t++)
            {
                v *= 128.0;
                const double d = __builtin_trunc(v);
                v -= d;
                w[t][q >> 2] |= (uint32_t)(uint8_t)(int8_t)(int)d << (8 * (q & 3));
            }
        }
        int8_t* out = ws + j.out + ((int64_t)rb * j.pnkg + kg) * S * 1024 + lane * 16;
        if (inmask == 0xffffu)
        {
#pragma unroll
            for (int t = 0; t < S; t++)
            {
                v4i x;
#pragma unroll
                for (int c = 0; c < 4; c++) x[c] = (int)w[t][c];
                *(v4i*)(out + t * 1024) = x;
            }
        }
        else if (inmask)
        {
#pragma unroll
            for (int t = 0; t < S; t++)
#pragma unroll
                for (int q = 0; q < 16; q++)
                    if (inmask >> q & 1u) out[t * 1024 + q] = (int8_t)(w[t][q >> 2] >> (8 * (q & 3)));
        }
    }
}

// Digit pairs (t, u), t + u <= S + 1 (1-based), row by row.  Consecutive
// MFMAs into one diagonal's accumulator do not stall (an order that spaced
// them at least two apart measured the same).
template <int S>
__device__ __forceinline__ void oz_mma(v16i (&acc)[S], const v4i (&a)[S], const v4i (&b)[S])
{
#pragma unroll
    for (int t = 0; t < S; t++)
#pragma unroll
        for (int u = 0; u + t < S; u++) acc[t + u] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[t], b[u], acc[t + u], 0, 0, 0);
}

template <int S>
__device__ __forceinline__ void oz_load(v4i (&a)[S], v4i (&b)[S], const v4i* pa, const v4i* pb, int kg)
{
#pragma unroll
    for (int t = 0; t < S; t++)
    {
        a[t] = pa[((int64_t)kg * S + t) * 64];
        b[t] = pb[((int64_t)kg * S + t) * 64];
    }
}

template <typename TC, int S>
__global__ __launch_bounds__(256) void oz_gemm_kernel(const PanelTask* __restrict__ tasks,
                                                      const int64_t* __restrict__ pool,
                                                      const int8_t* __restrict__ ws,
                                                      const unsigned long long* __restrict__ rmax,
                                                      double* __restrict__ cws, TC* __restrict__ C,
                                                      double alpha, double beta, int use_k)
{
    constexpr int PF = 5;
    // k groups between FP64 flushes: a multiple of PF (a pass never straddles
    // a flush) and at most 256 (8 pairs * 127^2 * 8192 k < 2^31): 255 groups,
    // 8160 k
    constexpr int kFlush = 256 / PF * PF;
    const PanelTask t = tasks[blockIdx.x];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = wave & 1, wn = wave >> 1;
    if (wm * 32 >= t.mlen || wn * 32 >= t.nlen) return;  // no barriers below

    v16i acc[S];
    double tot[16];
#pragma unroll
    for (int d = 0; d < S; d++)
#pragma unroll
        for (int r = 0; r < 16; r++) acc[d][r] = 0;
#pragma unroll
    for (int r = 0; r < 16; r++) tot[r] = 0.0;

    const int nk = use_k ? t.nkg : 0;
    if (nk > 0)
    {
        const int64_t rs = (int64_t)t.nkg * S * 64;  // v4i per row block
        const v4i* pa = (const v4i*)(ws + t.a) + wm * rs + lane;
        const v4i* pb = (const v4i*)(ws + t.b) + wn * rs + lane;
        // PF k groups in flight: one group's MFMA work (S (S + 1) / 2
        // instructions) is shorter than a load's latency at one wave per SIMD.
        // Five fit the register file at S = 8 (492 of 512); measured on 96^4:
        // one group 38 ms, four 22.1 ms, five 21.8 ms.
        v4i a[PF][S], b[PF][S];
#pragma unroll
        for (int d = 0; d < PF; d++) oz_load<S>(a[d], b[d], pa, pb, d < nk ? d : nk - 1);
        for (int kg = 0; kg < nk; kg += PF)
        {
#pragma unroll
            for (int d = 0; d < PF; d++)
            {
                if (kg + d >= nk) break;  // the last pass may be partial
                oz_mma<S>(acc, a[d], b[d]);
                // past the end the final group is re-read (in bounds, unused)
                oz_load<S>(a[d], b[d], pa, pb, kg + d + PF < nk ? kg + d + PF : nk - 1);
            }
            if ((kg + PF) % kFlush == 0 || kg + PF >= nk)
            {
#pragma unroll
                for (int d = 0; d < S; d++)
                {
                    // diagonal t + u = d + 2 carries weight 2^(-7 (d + 2))
                    const double w = __builtin_ldexp(1.0, -7 * (d + 2));
#pragma unroll
                    for (int r = 0; r < 16; r++)
                    {
                        tot[r] += w * (double)acc[d][r];
                        acc[d][r] = 0;
                    }
                }
            }
        }
    }

    // D register r of lane l holds D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32]
    const int cl = lane & 31, rh = lane >> 5;
    const int col = wn * 32 + cl;
    const bool cin = col < t.nlen;
    const int eb = cin ? oz_exponent(rmax[t.b_rt + col]) : 0;
    const int64_t coff = t.c_ld == 0 && cin ? pool[t.c_colmap + col] : 0;
#pragma unroll
    for (int r = 0; r < 16; r++)
    {
        const int row = wm * 32 + 8 * (r / 4) + 4 * rh + r % 4;
        if (row >= t.mlen || !cin) continue;
        const double s = __builtin_ldexp(tot[r], oz_exponent(rmax[t.a_rt + row]) + eb);
        if (t.c_ld > 0)
        {
            cws[t.c_src + (int64_t)row * t.c_ld + col] = s;
            continue;
        }
        TC* p = C + t.c_src + pool[t.c_rowmap + row] + coff;
        double v = alpha * s;
        if (beta != 0.0) v += beta * (double)*p;
        *p = (TC)v;
    }
}

template <int S>
hipError_t oz_launch_s(const PanelLaunch& L, unsigned long long* rmax, int64_t nrows)
{
    if (nrows > 0)
    {
        const int64_t blocks = std::min<int64_t>((nrows + 255) / 256, 1024);
        hipLaunchKernelGGL(oz_zero_kernel, dim3((unsigned)blocks), dim3(256), 0, L.stream, rmax, nrows);
    }
    if (L.use_k && L.nchunks > 0)
    {
        const dim3 grid(L.nchunks);
        if (L.dtype == 0)
        {
            hipLaunchKernelGGL(oz_rowmax_kernel<double>, grid, dim3(256), 0, L.stream, L.chunks, L.pool,
                               (const double*)L.A, (const double*)L.B, rmax);
            hipLaunchKernelGGL((oz_pack_kernel<double, S>), grid, dim3(256), 0, L.stream, L.chunks, L.pool,
                               (const double*)L.A, (const double*)L.B, rmax, (int8_t*)L.ws);
        }
        else
        {
            hipLaunchKernelGGL(oz_rowmax_kernel<float>, grid, dim3(256), 0, L.stream, L.chunks, L.pool,
                               (const float*)L.A, (const float*)L.B, rmax);
            hipLaunchKernelGGL((oz_pack_kernel<float, S>), grid, dim3(256), 0, L.stream, L.chunks, L.pool,
                               (const float*)L.A, (const float*)L.B, rmax, (int8_t*)L.ws);
        }
    }
    if (L.ntasks > 0)
    {
        const dim3 grid(L.ntasks);
        if (L.dtype == 0)
            hipLaunchKernelGGL((oz_gemm_kernel<double, S>), grid, dim3(256), 0, L.stream, L.tasks, L.pool,
                               (const int8_t*)L.ws, rmax, L.cws, (double*)L.C, L.alpha, L.beta, L.use_k);
        else
            hipLaunchKernelGGL((oz_gemm_kernel<float, S>), grid, dim3(256), 0, L.stream, L.tasks, L.pool,
                               (const int8_t*)L.ws, rmax, L.cws, (float*)L.C, L.alpha, L.beta, L.use_k);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_panel_ozaki(const PanelLaunch& L, int slices, unsigned long long* rmax, int64_t nrows)
{
    switch (slices)
    {
        case 6: return oz_launch_s<6>(L, rmax, nrows);
        case 7: return oz_launch_s<7>(L, rmax, nrows);
        case 8: return oz_launch_s<8>(L, rmax, nrows);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace tblis_amd
