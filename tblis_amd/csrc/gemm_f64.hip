// FP64 block-scatter GEMM on gfx950 (v_mfma_f64_16x16x4_f64).
//
// One workgroup (4 waves, 2x2) computes BM x BN tiles of C taken from a task
// list (persistent: a workgroup walks tasks blockIdx.x, +gridDim.x, ...).
// For each k segment of a task it gathers BM x BK tiles of A and BK x BN
// tiles of B straight from the permuted tensors through the block-scatter
// vectors (element offset = base + row_offset + col_offset) into LDS,
// double-buffered with one barrier per k tile, and feeds the MFMA.  The
// epilogue applies alpha/beta and stores through the C scatter vectors;
// beta == 0 overwrites C without reading it.
//
// v_mfma_f64_16x16x4_f64 operand layouts (checked by the lane-layout probe
// test): A lane l holds A[l%16][l/16], B lane l holds B[l/16][l%16], and
// accumulator register r of lane l is D[l/16 + 4r][l%16].
//
// Thread maps for the loads are chosen per operand by the planner: a "fast"
// side with unit stride makes consecutive lanes touch consecutive elements
// (coalesced), the other side is a per-thread gather.
//
// Replaces the reference's pack_nb/pack_sb + micro-kernel pipeline
// (src/kernels/3m/packm.hpp:293-404, src/nodes/gemm_ukr.hpp:279-363,
// src/kernels/3m/gemm.hpp:30-67) for FP64.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace tblis_amd
{

typedef double dbl4 __attribute__((ext_vector_type(4)));

namespace
{

constexpr int NT = 256;

template <int BM, int BN, int BK, bool A_MFAST, bool B_NFAST, typename T, bool MFMA>
__global__ void __launch_bounds__(NT)
gemm_f64_kernel(const TaskDesc* __restrict__ tasks, int ntasks, const SegDesc* __restrict__ segs,
                const int64_t* __restrict__ pool, const T* __restrict__ A, const T* __restrict__ B,
                T* __restrict__ C, double alpha, double beta, int use_k)
{
    // offset of logical index i of one matrix side
    auto side = [pool](const SideRef& r, int64_t i) -> int64_t {
        return r.stride == SCATTERED ? pool[r.start + i] : (r.start + i) * r.stride;
    };

    constexpr int PAD = 1;
    constexpr int WM = BM / 2, WN = BN / 2;
    constexpr int FM = WM / 16, FN = WN / 16;
    constexpr int EA = BM * BK / NT, EB = BN * BK / NT;
    static_assert(EA >= 1 && EB >= 1, "tile too small");
    static_assert(NT % BM == 0 && NT % BN == 0 && NT % BK == 0, "thread map");

    __shared__ double As[2][BK][BM + PAD];
    __shared__ double Bs[2][BK][BN + PAD];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm0 = (wave & 1) * WM;
    const int wn0 = (wave >> 1) * WN;

    int parity = 0;

    for (int t = blockIdx.x; t < ntasks; t += gridDim.x)
    {
        const TaskDesc task = tasks[t];
        dbl4 acc[FM][FN];
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
            for (int j = 0; j < FN; j++) acc[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};

        if (use_k && task.nseg > 0)
        {
            SegDesc seg = segs[task.seg0];
            // Offsets fixed for a whole segment: A rows (m-fast: one row per
            // thread; k-fast: EA rows per thread), B cols likewise.
            int64_t a_fix[A_MFAST ? 1 : EA];
            int64_t b_fix[B_NFAST ? 1 : EB];
            auto load_fixed = [&]() {
                if (A_MFAST)
                {
                    a_fix[0] = side(seg.a_row, min(tid % BM, task.mlen - 1));
                }
                else
                {
#pragma unroll
                    for (int j = 0; j < EA; j++)
                        a_fix[j] = side(seg.a_row, min(tid / BK + j * (NT / BK), task.mlen - 1));
                }
                if (B_NFAST)
                {
                    b_fix[0] = side(seg.b_col, min(tid % BN, task.nlen - 1));
                }
                else
                {
#pragma unroll
                    for (int j = 0; j < EB; j++)
                        b_fix[j] = side(seg.b_col, min(tid / BK + j * (NT / BK), task.nlen - 1));
                }
            };

            double ra[EA], rb[EB];
            auto load_tile = [&](int64_t k0) {
                const int64_t klen = seg.klen;
                if (A_MFAST)
                {
#pragma unroll
                    for (int j = 0; j < EA; j++)
                    {
                        int64_t kg = k0 + tid / BM + j * (NT / BM);
                        bool ok = kg < klen;
                        int64_t co = side(seg.a_col, ok ? kg : klen - 1);
                        double v = (double)A[seg.a_base + a_fix[0] + co];
                        ra[j] = ok ? v : 0.0;
                    }
                }
                else
                {
                    int64_t kg = k0 + tid % BK;
                    bool ok = kg < klen;
                    int64_t co = side(seg.a_col, ok ? kg : klen - 1);
#pragma unroll
                    for (int j = 0; j < EA; j++)
                    {
                        double v = (double)A[seg.a_base + a_fix[j] + co];
                        ra[j] = ok ? v : 0.0;
                    }
                }
                if (B_NFAST)
                {
#pragma unroll
                    for (int j = 0; j < EB; j++)
                    {
                        int64_t kg = k0 + tid / BN + j * (NT / BN);
                        bool ok = kg < klen;
                        int64_t ro = side(seg.b_row, ok ? kg : klen - 1);
                        double v = (double)B[seg.b_base + ro + b_fix[0]];
                        rb[j] = ok ? v : 0.0;
                    }
                }
                else
                {
                    int64_t kg = k0 + tid % BK;
                    bool ok = kg < klen;
                    int64_t ro = side(seg.b_row, ok ? kg : klen - 1);
#pragma unroll
                    for (int j = 0; j < EB; j++)
                    {
                        double v = (double)B[seg.b_base + ro + b_fix[j]];
                        rb[j] = ok ? v : 0.0;
                    }
                }
            };

            load_fixed();
            load_tile(0);
            int s = 0;
            int64_t kt = 0;
            while (true)
            {
                // registers -> LDS[parity]
                if (A_MFAST)
                {
#pragma unroll
                    for (int j = 0; j < EA; j++) As[parity][tid / BM + j * (NT / BM)][tid % BM] = ra[j];
                }
                else
                {
#pragma unroll
                    for (int j = 0; j < EA; j++) As[parity][tid % BK][tid / BK + j * (NT / BK)] = ra[j];
                }
                if (B_NFAST)
                {
#pragma unroll
                    for (int j = 0; j < EB; j++) Bs[parity][tid / BN + j * (NT / BN)][tid % BN] = rb[j];
                }
                else
                {
#pragma unroll
                    for (int j = 0; j < EB; j++) Bs[parity][tid % BK][tid / BK + j * (NT / BK)] = rb[j];
                }
                __syncthreads();

                // prefetch the next tile (possibly of the next segment)
                int64_t nkt = (seg.klen + BK - 1) / BK;
                int ns = s;
                int64_t nk = kt + 1;
                if (nk >= nkt)
                {
                    ns = s + 1;
                    nk = 0;
                }
                const bool more = ns < task.nseg;
                if (more)
                {
                    if (ns != s)
                    {
                        seg = segs[task.seg0 + ns];
                        load_fixed();
                    }
                    load_tile(nk * BK);
                }

                // compute from LDS[parity]
                if (MFMA)
                {
#pragma unroll
                    for (int k4 = 0; k4 < BK / 4; k4++)
                    {
                        double af[FM], bf[FN];
                        const int kr = k4 * 4 + lane / 16;
#pragma unroll
                        for (int i = 0; i < FM; i++) af[i] = As[parity][kr][wm0 + i * 16 + lane % 16];
#pragma unroll
                        for (int j = 0; j < FN; j++) bf[j] = Bs[parity][kr][wn0 + j * 16 + lane % 16];
#pragma unroll
                        for (int i = 0; i < FM; i++)
#pragma unroll
                            for (int j = 0; j < FN; j++)
                                acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j],
                                                                                 0, 0, 0);
                    }
                }
                else
                {
#pragma unroll 4
                    for (int kk = 0; kk < BK; kk++)
                    {
#pragma unroll
                        for (int i = 0; i < FM; i++)
                        {
#pragma unroll
                            for (int j = 0; j < FN; j++)
                            {
                                double bv = Bs[parity][kk][wn0 + j * 16 + lane % 16];
#pragma unroll
                                for (int r = 0; r < 4; r++)
                                    acc[i][j][r] = fma(As[parity][kk][wm0 + i * 16 + lane / 16 + 4 * r],
                                                       bv, acc[i][j][r]);
                            }
                        }
                    }
                }
                parity ^= 1;
                if (!more) break;
                s = ns;
                kt = nk;
            }
        }

        // epilogue: C <- alpha*acc + beta*C through the C scatter vectors
        int64_t col_off[FN];
        bool col_ok[FN];
#pragma unroll
        for (int j = 0; j < FN; j++)
        {
            int col = wn0 + j * 16 + lane % 16;
            col_ok[j] = col < task.nlen;
            col_off[j] = side(task.c_col, col_ok[j] ? col : 0);
        }
#pragma unroll
        for (int i = 0; i < FM; i++)
        {
#pragma unroll
            for (int r = 0; r < 4; r++)
            {
                int row = wm0 + i * 16 + lane / 16 + 4 * r;
                if (row >= task.mlen) continue;
                int64_t ro = task.c_base + side(task.c_row, row);
#pragma unroll
                for (int j = 0; j < FN; j++)
                {
                    if (!col_ok[j]) continue;
                    T* p = C + ro + col_off[j];
                    double v = alpha * acc[i][j][r];
                    if (beta != 0.0) v = fma(beta, (double)*p, v);
                    *p = (T)v;
                }
            }
        }
    }
}

template <int BM, int BN, typename T, bool MFMA>
const void* select_layout(const KernelKey& k)
{
    constexpr int BK = 16;
    if (k.a_mfast && k.b_nfast) return (const void*)&gemm_f64_kernel<BM, BN, BK, true, true, T, MFMA>;
    if (k.a_mfast) return (const void*)&gemm_f64_kernel<BM, BN, BK, true, false, T, MFMA>;
    if (k.b_nfast) return (const void*)&gemm_f64_kernel<BM, BN, BK, false, true, T, MFMA>;
    return (const void*)&gemm_f64_kernel<BM, BN, BK, false, false, T, MFMA>;
}

template <typename T>
const void* select_typed(const KernelKey& k)
{
    if (!k.use_mfma) return select_layout<64, 64, T, false>(k);
    if (k.tile == 0) return select_layout<128, 128, T, true>(k);
    return select_layout<64, 64, T, true>(k);
}

const void* select_kernel(const KernelKey& k)
{
    return k.dtype == 0 ? select_typed<double>(k) : select_typed<float>(k);
}

}  // namespace

int gemm_f64_tile_m(int tile) { return tile == 0 ? 128 : 64; }
int gemm_f64_tile_n(int tile) { return tile == 0 ? 128 : 64; }

hipError_t launch_gemm_f64(const GemmLaunch& L, const KernelKey& k)
{
    if (L.ntasks == 0) return hipSuccess;
    const TaskDesc* tasks = L.tasks;
    int ntasks = L.ntasks;
    const SegDesc* segs = L.segs;
    const int64_t* pool = L.pool;
    const void* A = L.A;
    const void* B = L.B;
    void* C = L.C;
    double alpha = L.alpha, beta = L.beta;
    int use_k = L.use_k;
    void* args[] = {&tasks, &ntasks, &segs, &pool, &A, &B, &C, &alpha, &beta, &use_k};
    hipError_t e = hipLaunchKernel(select_kernel(k), dim3(L.grid), dim3(NT), args, 0, L.stream);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

int gemm_f64_occupancy(const KernelKey& k)
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, select_kernel(k), NT, 0) != hipSuccess)
    {
        (void)hipGetLastError();
        return 1;
    }
    return n;
}

// One 16x16x4 FP64 MFMA with A[i][k] = 1000*i + k + 1, B[k][j] = (k+1)*(j+1)
// loaded in the operand layout the GEMM assumes; each lane writes its four
// D values at the (row, col) positions the epilogue assumes
// (row l/16 + 4r, col l%16).  Lets a test
// check the lane layout against a host product.
__global__ void mfma_probe_kernel(double* out)
{
    const int lane = threadIdx.x;
    const int i = lane % 16, kq = lane / 16;
    double a = 1000.0 * i + kq + 1;
    double b = (kq + 1.0) * (i + 1.0);
    dbl4 d = {0.0, 0.0, 0.0, 0.0};
    d = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, d, 0, 0, 0);
    for (int r = 0; r < 4; r++) out[(lane / 16 + 4 * r) * 16 + lane % 16] = d[r];
}

hipError_t launch_mfma_probe(double* out, hipStream_t s)
{
    hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, s, out);
    return hipGetLastError();
}

}  // namespace tblis_amd
