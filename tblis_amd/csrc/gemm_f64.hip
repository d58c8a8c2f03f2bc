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

#ifndef TBLIS_SCHED_BARRIER
#define TBLIS_SCHED_BARRIER 1
#endif

namespace
{

// offset of logical index i of one matrix side
__device__ __forceinline__ int64_t side_off(const int64_t* __restrict__ pool, const SideRef& r, int64_t i)
{
    return r.stride == SCATTERED ? pool[r.start + i] : (r.start + i) * r.stride;
}

// (offset of the first k of tile kt, block stride or SCATTERED) of a k side;
// uniform across the workgroup.
__device__ __forceinline__ void k_block(const int64_t* __restrict__ pool, const SideRef& r,
                                        int64_t tbl, int64_t kt, int BK, int64_t& base, int64_t& stride)
{
    if (r.stride != SCATTERED)
    {
        base = (r.start + kt * BK) * r.stride;
        stride = r.stride;
    }
    else
    {
        base = pool[tbl + 2 * kt];
        stride = pool[tbl + 2 * kt + 1];
    }
}

template <int BM, int BN, int BK, int WGM, int WGN, bool A_MFAST, bool B_NFAST, typename T, bool MFMA>
__global__ void __launch_bounds__(64 * WGM * WGN)
gemm_f64_kernel(const TaskDesc* __restrict__ tasks, int ntasks, const SegDesc* __restrict__ segs,
                const int64_t* __restrict__ pool, const T* __restrict__ A, const T* __restrict__ B,
                T* __restrict__ C, double alpha, double beta, int use_k)
{
    constexpr int NT = 64 * WGM * WGN;
    constexpr int PAD = 1;
    constexpr int WM = BM / WGM, WN = BN / WGN;
    static_assert(WM % 16 == 0 && WN % 16 == 0, "wave tile");
    constexpr int FM = WM / 16, FN = WN / 16;
    constexpr int EA = BM * BK / NT, EB = BN * BK / NT;
    static_assert(EA >= 1 && EB >= 1, "tile too small");
    static_assert(NT % BM == 0 && NT % BN == 0 && NT % BK == 0, "thread map");

    __shared__ double As[2][BK][BM + PAD];
    __shared__ double Bs[2][BK][BN + PAD];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm0 = (wave % WGM) * WM;
    const int wn0 = (wave / WGM) * WN;

    // this thread's tile coordinates for the loads (m-fast / k-fast maps)
    const int a_i = A_MFAST ? tid % BM : tid / BK;   // first row
    const int a_k = A_MFAST ? tid / BM : tid % BK;   // first k
    const int b_j = B_NFAST ? tid % BN : tid / BK;   // first col
    const int b_k = B_NFAST ? tid / BN : tid % BK;   // first k
    constexpr int A_RSTEP = A_MFAST ? 0 : NT / BK, A_KSTEP = A_MFAST ? NT / BM : 0;
    constexpr int B_CSTEP = B_NFAST ? 0 : NT / BK, B_KSTEP = B_NFAST ? NT / BN : 0;

    int parity = 0;

    for (int t = blockIdx.x; t < ntasks; t += gridDim.x)
    {
        const TaskDesc task = tasks[t];
        dbl4 acc[FM][FN];
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
            for (int j = 0; j < FN; j++) acc[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};

        const int nseg = use_k ? task.nseg : 0;
        for (int s = 0; s < nseg; s++)
        {
            const SegDesc seg = segs[task.seg0 + s];
            const int64_t klen = seg.klen;
            const int64_t nkt = (klen + BK - 1) / BK;

            // per-thread row offsets of A and column offsets of B, fixed for
            // the segment (rows past the tile edge repeat the last valid one)
            int64_t a_fix[EA], b_fix[EB];
#pragma unroll
            for (int j = 0; j < EA; j++)
                a_fix[j] = seg.a_base + side_off(pool, seg.a_row, min(a_i + j * A_RSTEP, task.mlen - 1));
#pragma unroll
            for (int j = 0; j < EB; j++)
                b_fix[j] = seg.b_base + side_off(pool, seg.b_col, min(b_j + j * B_CSTEP, task.nlen - 1));

            double ra[EA], rb[EB];

            // global -> registers for k tile kt
            auto load = [&](int64_t kt) {
                int64_t akb, aks, bkb, bks;
                k_block(pool, seg.a_col, seg.a_kblk, kt, BK, akb, aks);
                k_block(pool, seg.b_row, seg.b_kblk, kt, BK, bkb, bks);
                const int64_t k0 = kt * BK;
                const bool full = k0 + BK <= klen;
#pragma unroll
                for (int j = 0; j < EA; j++)
                {
                    const int kk = a_k + j * A_KSTEP;
                    const bool ok = full || k0 + kk < klen;
                    const int64_t kc = ok ? kk : klen - 1 - k0;
                    const int64_t co = aks != SCATTERED ? akb + kc * aks : pool[seg.a_col.start + k0 + kc];
                    const double v = (double)A[a_fix[j] + co];
                    ra[j] = ok ? v : 0.0;
                }
#pragma unroll
                for (int j = 0; j < EB; j++)
                {
                    const int kk = b_k + j * B_KSTEP;
                    const bool ok = full || k0 + kk < klen;
                    const int64_t kc = ok ? kk : klen - 1 - k0;
                    const int64_t ro = bks != SCATTERED ? bkb + kc * bks : pool[seg.b_row.start + k0 + kc];
                    const double v = (double)B[b_fix[j] + ro];
                    rb[j] = ok ? v : 0.0;
                }
            };

            load(0);
            for (int64_t kt = 0; kt < nkt; kt++)
            {
#pragma unroll
                for (int j = 0; j < EA; j++)
                {
                    if (A_MFAST)
                        As[parity][a_k + j * A_KSTEP][a_i] = ra[j];
                    else
                        As[parity][a_k][a_i + j * A_RSTEP] = ra[j];
                }
#pragma unroll
                for (int j = 0; j < EB; j++)
                {
                    if (B_NFAST)
                        Bs[parity][b_k + j * B_KSTEP][b_j] = rb[j];
                    else
                        Bs[parity][b_k][b_j + j * B_CSTEP] = rb[j];
                }
                __syncthreads();
                if (kt + 1 < nkt) load(kt + 1);
#if TBLIS_SCHED_BARRIER
                // keep the next tile's global loads ahead of the MFMA block so
                // their latency hides behind it
                __builtin_amdgcn_sched_barrier(0);
#endif

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
            col_off[j] = side_off(pool, task.c_col, col_ok[j] ? col : 0);
        }
#pragma unroll
        for (int i = 0; i < FM; i++)
        {
#pragma unroll
            for (int r = 0; r < 4; r++)
            {
                int row = wm0 + i * 16 + lane / 16 + 4 * r;
                if (row >= task.mlen) continue;
                int64_t ro = task.c_base + side_off(pool, task.c_row, row);
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

// Tile table: (id, BM, BN, BK, waves along m, waves along n).
#define TBLIS_F64_TILES(X)      \
    X(0, 128, 128, 16, 2, 2)    \
    X(1, 64, 64, 16, 2, 2)      \
    X(2, 128, 64, 16, 2, 2)     \
    X(3, 64, 128, 16, 2, 2)     \
    X(4, 128, 128, 16, 2, 4)    \
    X(5, 32, 32, 16, 2, 2)      \
    X(6, 64, 64, 32, 2, 2)      \
    X(7, 128, 64, 16, 4, 2)

template <int BM, int BN, int BK, int WGM, int WGN, typename T, bool MFMA>
const void* select_layout(const KernelKey& k)
{
    if (k.a_mfast && k.b_nfast)
        return (const void*)&gemm_f64_kernel<BM, BN, BK, WGM, WGN, true, true, T, MFMA>;
    if (k.a_mfast) return (const void*)&gemm_f64_kernel<BM, BN, BK, WGM, WGN, true, false, T, MFMA>;
    if (k.b_nfast) return (const void*)&gemm_f64_kernel<BM, BN, BK, WGM, WGN, false, true, T, MFMA>;
    return (const void*)&gemm_f64_kernel<BM, BN, BK, WGM, WGN, false, false, T, MFMA>;
}

template <typename T>
const void* select_typed(const KernelKey& k)
{
    if (!k.use_mfma) return select_layout<64, 64, 16, 2, 2, T, false>(k);
    switch (k.tile)
    {
#define TBLIS_CASE(id, bm, bn, bk, wm, wn) \
    case id: return select_layout<bm, bn, bk, wm, wn, T, true>(k);
        TBLIS_F64_TILES(TBLIS_CASE)
#undef TBLIS_CASE
    }
    return nullptr;
}

const void* select_kernel(const KernelKey& k)
{
    return k.dtype == 0 ? select_typed<double>(k) : select_typed<float>(k);
}

struct TileShape
{
    int bm, bn, bk, threads;
};

TileShape tile_shape(int tile, int use_mfma)
{
    if (!use_mfma) return {64, 64, 16, 256};
    switch (tile)
    {
#define TBLIS_CASE(id, bm, bn, bk, wm, wn) \
    case id: return {bm, bn, bk, 64 * wm * wn};
        TBLIS_F64_TILES(TBLIS_CASE)
#undef TBLIS_CASE
    }
    return {64, 64, 16, 256};
}

}  // namespace

int gemm_f64_num_tiles() { return 8; }
int gemm_f64_tile_m(int tile) { return tile_shape(tile, 1).bm; }
int gemm_f64_tile_n(int tile) { return tile_shape(tile, 1).bn; }
int gemm_f64_tile_k(int tile) { return tile_shape(tile, 1).bk; }

hipError_t launch_gemm_f64(const GemmLaunch& L, const KernelKey& k)
{
    if (L.ntasks == 0) return hipSuccess;
    const void* f = select_kernel(k);
    if (!f) return hipErrorInvalidValue;
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
    hipError_t e = hipLaunchKernel(f, dim3(L.grid), dim3(tile_shape(k.tile, k.use_mfma).threads), args,
                                   0, L.stream);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

int gemm_f64_occupancy(const KernelKey& k)
{
    const void* f = select_kernel(k);
    int n = 0;
    if (!f ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, tile_shape(k.tile, k.use_mfma).threads, 0) !=
            hipSuccess)
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
