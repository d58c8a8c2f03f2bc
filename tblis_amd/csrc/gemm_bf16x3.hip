// FP32 block-scatter GEMM as 3-term split-BF16 products on gfx950
// (v_mfma_f32_32x32x16_bf16), FP32 accumulation.
//
// Every FP32 element x gathered from A or B is split once per workgroup into
// three BF16 terms, x = hi + mid + lo + e with |e| <= 2^-27 |x|
// (hi = RN(x), mid = RN(x - hi), lo = RN(x - hi - mid); the residuals are
// exact in FP32).  The product keeps the six terms of order >= 2^-18:
//     hi*hi + hi*mid + mid*hi + hi*lo + lo*hi + mid*mid,
// dropping mid*lo, lo*mid, lo*lo (order 2^-24 and below).  The three planes
// of each operand tile sit in LDS as [row][k] BF16 so a lane reads its
// 8 consecutive k with one ds_read_b128 per plane.
//
// Accumulation: FP32 MFMA accumulators restart every CHUNK k tiles (256 k)
// and the chunk sums are added into a second FP32 register set, which keeps
// the rounding error of long k loops (k = 9216) near 1e-7 relative instead
// of the ~1e-6 a single accumulator reaches.  The unchunked variant is the
// "bf16x3_fast" configuration.  The alpha/beta update of C runs in FP64.
//
// Task lists, scatter pool and k-block tables are the ones the FP64 kernel
// uses (plan_types.hpp); only the element type and the inner product differ.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace tblis_amd
{

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace
{

__device__ __forceinline__ int64_t side_off(const int64_t* __restrict__ pool, const SideRef& r, int64_t i)
{
    return r.stride == SCATTERED ? pool[r.start + i] : (r.start + i) * r.stride;
}

__device__ __forceinline__ void k_block(const int64_t* __restrict__ pool, const SideRef& r, int64_t tbl,
                                        int64_t kt, int BK, int64_t& base, int64_t& stride)
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

// x -> (hi, mid, lo), round to nearest even at each step
__device__ __forceinline__ void split3(float x, __bf16& h, __bf16& m, __bf16& l)
{
    h = (__bf16)x;
    const float r1 = x - (float)h;
    m = (__bf16)r1;
    const float r2 = r1 - (float)m;
    l = (__bf16)r2;
}

template <int BM, int BN, int WGM, int WGN, bool A_KFAST, bool B_KFAST, bool CHUNKED>
__global__ void __launch_bounds__(64 * WGM * WGN)
gemm_bf16x3_kernel(const TaskDesc* __restrict__ tasks, int ntasks, const SegDesc* __restrict__ segs,
                   const int64_t* __restrict__ pool, const float* __restrict__ A,
                   const float* __restrict__ B, float* __restrict__ C, double alpha, double beta,
                   int use_k)
{
    constexpr int BK = 16;
    constexpr int NT = 64 * WGM * WGN;
    constexpr int WM = BM / WGM, WN = BN / WGN;
    constexpr int FM = WM / 32, FN = WN / 32;
    constexpr int KP = BK + 8;  // padded LDS row (BF16 elements)
    constexpr int CHUNK = 16;   // k tiles per accumulation chunk
    constexpr int GA = 2 * BM, GB = 2 * BN;  // loader groups: one row x 8 k each
    constexpr int NG = (GA + GB + NT - 1) / NT;  // groups per thread (last may be idle)
    static_assert(GA % 64 == 0 && GB % 64 == 0, "loader map");
    static_assert(WM % 32 == 0 && WN % 32 == 0, "wave tile");

    __shared__ __bf16 As[2][3][BM][KP];
    __shared__ __bf16 Bs[2][3][BN][KP];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm0 = (wave % WGM) * WM;
    const int wn0 = (wave / WGM) * WN;

    int parity = 0;

    for (int t = blockIdx.x; t < ntasks; t += gridDim.x)
    {
        const TaskDesc task = tasks[t];
        f32x16 acc[FM][FN], sum[FM][FN];
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
            for (int j = 0; j < FN; j++)
                for (int r = 0; r < 16; r++)
                {
                    acc[i][j][r] = 0.0f;
                    sum[i][j][r] = 0.0f;
                }
        int chunk_left = CHUNK;

        const int nseg = use_k ? task.nseg : 0;
        for (int s = 0; s < nseg; s++)
        {
            const SegDesc seg = segs[task.seg0 + s];
            const int64_t klen = seg.klen;
            const int64_t nkt = (klen + BK - 1) / BK;

            // loader groups of this thread: (operand, row, k half) and the
            // row's element offset, fixed for the segment
            int g_row[NG], g_kh[NG];
            bool g_isA[NG], g_on[NG];
            int64_t g_fix[NG];
#pragma unroll
            for (int q = 0; q < NG; q++)
            {
                const int g = min(tid + q * NT, GA + GB - 1);
                g_on[q] = tid + q * NT < GA + GB;  // uniform per wave (64 | GA, GB)
                const bool isA = g < GA;
                const int gg = isA ? g : g - GA;
                const int R = isA ? BM : BN;
                const bool kfast = isA ? A_KFAST : B_KFAST;
                const int row = kfast ? gg / 2 : gg % R;
                const int kh = kfast ? gg % 2 : gg / R;
                g_isA[q] = isA;
                g_row[q] = row;
                g_kh[q] = kh;
                g_fix[q] = isA ? seg.a_base + side_off(pool, seg.a_row, min(row, task.mlen - 1))
                               : seg.b_base + side_off(pool, seg.b_col, min(row, task.nlen - 1));
            }

            float rv[NG][8];
            auto load = [&](int64_t kt) {
                int64_t akb, aks, bkb, bks;
                k_block(pool, seg.a_col, seg.a_kblk, kt, BK, akb, aks);
                k_block(pool, seg.b_row, seg.b_kblk, kt, BK, bkb, bks);
                const int64_t k0 = kt * BK;
                const bool full = k0 + BK <= klen;
#pragma unroll
                for (int q = 0; q < NG; q++)
                {
                    if (!g_on[q]) continue;
                    const float* base = g_isA[q] ? A : B;
                    const int64_t kb = g_isA[q] ? akb : bkb;
                    const int64_t ks = g_isA[q] ? aks : bks;
                    const int64_t kstart = g_isA[q] ? seg.a_col.start : seg.b_row.start;
#pragma unroll
                    for (int j = 0; j < 8; j++)
                    {
                        const int kk = 8 * g_kh[q] + j;
                        const bool ok = full || k0 + kk < klen;
                        const int64_t kc = ok ? kk : klen - 1 - k0;
                        const int64_t ko = ks != SCATTERED ? kb + kc * ks : pool[kstart + k0 + kc];
                        const float v = base[g_fix[q] + ko];
                        rv[q][j] = ok ? v : 0.0f;
                    }
                }
            };

            load(0);
            for (int64_t kt = 0; kt < nkt; kt++)
            {
                // split and store the three planes
#pragma unroll
                for (int q = 0; q < NG; q++)
                {
                    if (!g_on[q]) continue;
                    bf16x8 h, m, l;
#pragma unroll
                    for (int j = 0; j < 8; j++)
                    {
                        __bf16 a, b, c;
                        split3(rv[q][j], a, b, c);
                        h[j] = a;
                        m[j] = b;
                        l[j] = c;
                    }
                    __bf16* dst = g_isA[q] ? &As[parity][0][g_row[q]][8 * g_kh[q]]
                                           : &Bs[parity][0][g_row[q]][8 * g_kh[q]];
                    const int plane = g_isA[q] ? BM * KP : BN * KP;
                    *(bf16x8*)dst = h;
                    *(bf16x8*)(dst + plane) = m;
                    *(bf16x8*)(dst + 2 * plane) = l;
                }
                __syncthreads();
                if (kt + 1 < nkt) load(kt + 1);

                bf16x8 af[FM][3], bf[FN][3];
#pragma unroll
                for (int i = 0; i < FM; i++)
#pragma unroll
                    for (int p = 0; p < 3; p++)
                        af[i][p] = *(const bf16x8*)&As[parity][p][wm0 + i * 32 + lane % 32][8 * (lane / 32)];
#pragma unroll
                for (int j = 0; j < FN; j++)
#pragma unroll
                    for (int p = 0; p < 3; p++)
                        bf[j][p] = *(const bf16x8*)&Bs[parity][p][wn0 + j * 32 + lane % 32][8 * (lane / 32)];
#pragma unroll
                for (int i = 0; i < FM; i++)
#pragma unroll
                    for (int j = 0; j < FN; j++)
                    {
                        f32x16 c = acc[i][j];
                        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[i][1], bf[j][1], c, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[i][0], bf[j][2], c, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[i][2], bf[j][0], c, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[i][0], bf[j][1], c, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[i][1], bf[j][0], c, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[i][0], bf[j][0], c, 0, 0, 0);
                        acc[i][j] = c;
                    }
                parity ^= 1;
                if (CHUNKED && --chunk_left == 0)
                {
                    chunk_left = CHUNK;
#pragma unroll
                    for (int i = 0; i < FM; i++)
#pragma unroll
                        for (int j = 0; j < FN; j++)
                        {
                            sum[i][j] += acc[i][j];
                            for (int r = 0; r < 16; r++) acc[i][j][r] = 0.0f;
                        }
                }
            }
        }

        // epilogue: C <- alpha*(sum + acc) + beta*C, in FP64
        // accumulator register r of lane l is D[8*(r/4) + 4*(l/32) + r%4][l%32]
        int64_t col_off[FN];
        bool col_ok[FN];
#pragma unroll
        for (int j = 0; j < FN; j++)
        {
            const int col = wn0 + j * 32 + lane % 32;
            col_ok[j] = col < task.nlen;
            col_off[j] = side_off(pool, task.c_col, col_ok[j] ? col : 0);
        }
#pragma unroll
        for (int i = 0; i < FM; i++)
        {
#pragma unroll
            for (int r = 0; r < 16; r++)
            {
                const int row = wm0 + i * 32 + 8 * (r / 4) + 4 * (lane / 32) + r % 4;
                if (row >= task.mlen) continue;
                const int64_t ro = task.c_base + side_off(pool, task.c_row, row);
#pragma unroll
                for (int j = 0; j < FN; j++)
                {
                    if (!col_ok[j]) continue;
                    float* p = C + ro + col_off[j];
                    const float res = CHUNKED ? sum[i][j][r] + acc[i][j][r] : acc[i][j][r];
                    double v = alpha * (double)res;
                    if (beta != 0.0) v = fma(beta, (double)*p, v);
                    *p = (float)v;
                }
            }
        }
    }
}

// Tile table: (id, BM, BN, waves along m, waves along n).
#define TBLIS_BF16_TILES(X) \
    X(0, 128, 128, 2, 2)    \
    X(1, 64, 64, 2, 2)      \
    X(2, 128, 64, 2, 2)

template <int BM, int BN, int WGM, int WGN, bool CH>
const void* select_layout(const KernelKey& k)
{
    const bool ak = !k.a_mfast, bk = !k.b_nfast;
    if (ak && bk) return (const void*)&gemm_bf16x3_kernel<BM, BN, WGM, WGN, true, true, CH>;
    if (ak) return (const void*)&gemm_bf16x3_kernel<BM, BN, WGM, WGN, true, false, CH>;
    if (bk) return (const void*)&gemm_bf16x3_kernel<BM, BN, WGM, WGN, false, true, CH>;
    return (const void*)&gemm_bf16x3_kernel<BM, BN, WGM, WGN, false, false, CH>;
}

template <bool CH>
const void* select_chunked(const KernelKey& k)
{
    switch (k.tile)
    {
#define TBLIS_CASE(id, bm, bn, wm, wn) \
    case id: return select_layout<bm, bn, wm, wn, CH>(k);
        TBLIS_BF16_TILES(TBLIS_CASE)
#undef TBLIS_CASE
    }
    return nullptr;
}

const void* select_kernel(const KernelKey& k)
{
    return k.family == FAMILY_BF16X3 ? select_chunked<true>(k) : select_chunked<false>(k);
}

struct Shape
{
    int bm, bn, threads;
};

Shape shape(int tile)
{
    switch (tile)
    {
#define TBLIS_CASE(id, bm, bn, wm, wn) \
    case id: return {bm, bn, 64 * wm * wn};
        TBLIS_BF16_TILES(TBLIS_CASE)
#undef TBLIS_CASE
    }
    return {64, 64, 256};
}

// One 32x32x16 BF16 MFMA in the GEMM's operand layouts: lane l takes
// A[l%32][8*(l/32)+j] and B[8*(l/32)+j][l%32] (j = 0..7) from row-major
// A (32x16) and B (16x32) and writes D (32x32, row-major) with the output
// layout the epilogue assumes.
__global__ void bf16_probe_kernel(const float* A, const float* B, float* D)
{
    const int l = threadIdx.x;
    bf16x8 a, b;
    for (int j = 0; j < 8; j++)
    {
        a[j] = (__bf16)A[(l % 32) * 16 + 8 * (l / 32) + j];
        b[j] = (__bf16)B[(8 * (l / 32) + j) * 32 + l % 32];
    }
    f32x16 d = {};
    d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, d, 0, 0, 0);
    for (int r = 0; r < 16; r++) D[(8 * (r / 4) + 4 * (l / 32) + r % 4) * 32 + l % 32] = d[r];
}

}  // namespace

int gemm_bf16_num_tiles() { return 3; }
int gemm_bf16_tile_m(int tile) { return shape(tile).bm; }
int gemm_bf16_tile_n(int tile) { return shape(tile).bn; }

hipError_t launch_gemm_bf16x3(const GemmLaunch& L, const KernelKey& k)
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
    hipError_t e = hipLaunchKernel(f, dim3(L.grid), dim3(shape(k.tile).threads), args, 0, L.stream);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

int gemm_bf16x3_occupancy(const KernelKey& k)
{
    const void* f = select_kernel(k);
    int n = 0;
    if (!f || hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, shape(k.tile).threads, 0) != hipSuccess)
    {
        (void)hipGetLastError();
        return 1;
    }
    return n;
}

hipError_t launch_bf16_probe(const float* A, const float* B, float* D, hipStream_t s)
{
    hipLaunchKernelGGL(bf16_probe_kernel, dim3(1), dim3(64), 0, s, A, B, D);
    return hipGetLastError();
}

}  // namespace tblis_amd

// ---------------------------------------------------------------------------
// Packed path: split once into BF16 panels, then a GEMM over the panels.
// ---------------------------------------------------------------------------
namespace tblis_amd
{
namespace
{

constexpr int PR = 128;  // panel rows
constexpr int PK = 16;   // panel k per chunk
constexpr int PCH = 3 * PR * PK;  // BF16 elements per chunk

// One workgroup (256 threads) packs one chunk (p, kt): each thread gathers
// 8 consecutive k of one row (k-fast: lane pairs share a row; rows-fast:
// consecutive lanes take consecutive rows), splits, and writes 3 x 16 bytes.
template <bool KFAST>
__global__ void __launch_bounds__(256)
pack_bf16x3_kernel(PackDesc d, const int64_t* __restrict__ pool, const float* __restrict__ X,
                   __bf16* __restrict__ out)
{
    const int tid = threadIdx.x;
    const int p = blockIdx.x;
    const int64_t kt = blockIdx.y;
    const int row = KFAST ? tid / 2 : tid % PR;
    const int kh = KFAST ? tid % 2 : tid / PR;
    const int64_t r = (int64_t)p * PR + row;
    const bool rok = r < d.nrows;
    const int64_t ro = d.base + side_off(pool, d.rows, rok ? r : d.nrows - 1);
    int64_t kb, ks;
    k_block(pool, d.kside, d.ktbl, kt, PK, kb, ks);
    const int64_t k0 = kt * PK;
    bf16x8 h, m, l;
#pragma unroll
    for (int j = 0; j < 8; j++)
    {
        const int kk = 8 * kh + j;
        const bool ok = rok && k0 + kk < d.klen;
        const int64_t kc = k0 + kk < d.klen ? kk : d.klen - 1 - k0;
        const int64_t ko = ks != SCATTERED ? kb + kc * ks : pool[d.kside.start + k0 + kc];
        const float v = ok ? X[ro + ko] : 0.0f;
        __bf16 a, b, c;
        split3(v, a, b, c);
        h[j] = a;
        m[j] = b;
        l[j] = c;
    }
    __bf16* dst = out + d.out + ((int64_t)p * d.nkt + kt) * PCH + row * PK + 8 * kh;
    *(bf16x8*)dst = h;
    *(bf16x8*)(dst + PR * PK) = m;
    *(bf16x8*)(dst + 2 * PR * PK) = l;
}

// GEMM over packed panels: 128x128 C tile, 4 waves of 64x64 (2x2 blocks of
// 32x32x16 BF16 MFMA), 6 products per block, chunked FP32 accumulation.
// SegDesc.a_base / b_base hold the BF16 offsets of the task's A and B
// panels, klen the k extent.
template <bool CHUNKED>
__global__ void __launch_bounds__(256, 2)
gemm_bf16x3_packed_kernel(const TaskDesc* __restrict__ tasks, int ntasks,
                          const SegDesc* __restrict__ segs, const int64_t* __restrict__ pool,
                          const __bf16* __restrict__ Ap, const __bf16* __restrict__ Bp,
                          float* __restrict__ C, double alpha, double beta, int use_k)
{
    constexpr int KP = PK + 8;
    constexpr int CHUNK = 16;
    __shared__ __bf16 As[2][3][PR][KP];
    __shared__ __bf16 Bs[2][3][PR][KP];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm0 = (wave % 2) * 64;
    const int wn0 = (wave / 2) * 64;
    // copy map: piece (tid + 256 i) of a chunk = plane i, row tid/2, k half tid%2
    const int crow = tid / 2, ckh = tid % 2;

    int parity = 0;
    for (int t = blockIdx.x; t < ntasks; t += gridDim.x)
    {
        const TaskDesc task = tasks[t];
        f32x16 acc[2][2], sum[2][2];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
                for (int r = 0; r < 16; r++)
                {
                    acc[i][j][r] = 0.0f;
                    sum[i][j][r] = 0.0f;
                }
        int chunk_left = CHUNK;

        if (use_k && task.nseg > 0)
        {
            const SegDesc seg = segs[task.seg0];
            const int64_t nkt = (seg.klen + PK - 1) / PK;
            const uint4* ga = (const uint4*)(Ap + seg.a_base) + tid;
            const uint4* gb = (const uint4*)(Bp + seg.b_base) + tid;
            uint4 ra[3], rb[3];
#pragma unroll
            for (int i = 0; i < 3; i++)
            {
                ra[i] = ga[i * 256];
                rb[i] = gb[i * 256];
            }
            for (int64_t kt = 0; kt < nkt; kt++)
            {
#pragma unroll
                for (int i = 0; i < 3; i++)
                {
                    *(uint4*)&As[parity][i][crow][8 * ckh] = ra[i];
                    *(uint4*)&Bs[parity][i][crow][8 * ckh] = rb[i];
                }
                __syncthreads();
                if (kt + 1 < nkt)
                {
                    ga += PCH / 8;
                    gb += PCH / 8;
#pragma unroll
                    for (int i = 0; i < 3; i++)
                    {
                        ra[i] = ga[i * 256];
                        rb[i] = gb[i * 256];
                    }
                }
                bf16x8 af[2][3], bf[2][3];
#pragma unroll
                for (int i = 0; i < 2; i++)
#pragma unroll
                    for (int p = 0; p < 3; p++)
                    {
                        af[i][p] = *(const bf16x8*)&As[parity][p][wm0 + i * 32 + lane % 32][8 * (lane / 32)];
                        bf[i][p] = *(const bf16x8*)&Bs[parity][p][wn0 + i * 32 + lane % 32][8 * (lane / 32)];
                    }
#pragma unroll
                for (int i = 0; i < 2; i++)
#pragma unroll
                    for (int j = 0; j < 2; j++)
                    {
                        f32x16 c = acc[i][j];
                        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[i][1], bf[j][1], c, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[i][0], bf[j][2], c, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[i][2], bf[j][0], c, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[i][0], bf[j][1], c, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[i][1], bf[j][0], c, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[i][0], bf[j][0], c, 0, 0, 0);
                        acc[i][j] = c;
                    }
                parity ^= 1;
                if (CHUNKED && --chunk_left == 0)
                {
                    chunk_left = CHUNK;
#pragma unroll
                    for (int i = 0; i < 2; i++)
#pragma unroll
                        for (int j = 0; j < 2; j++)
                        {
                            sum[i][j] += acc[i][j];
                            for (int r = 0; r < 16; r++) acc[i][j][r] = 0.0f;
                        }
                }
            }
        }

        int64_t col_off[2];
        bool col_ok[2];
#pragma unroll
        for (int j = 0; j < 2; j++)
        {
            const int col = wn0 + j * 32 + lane % 32;
            col_ok[j] = col < task.nlen;
            col_off[j] = side_off(pool, task.c_col, col_ok[j] ? col : 0);
        }
#pragma unroll
        for (int i = 0; i < 2; i++)
        {
#pragma unroll
            for (int r = 0; r < 16; r++)
            {
                const int row = wm0 + i * 32 + 8 * (r / 4) + 4 * (lane / 32) + r % 4;
                if (row >= task.mlen) continue;
                const int64_t ro = task.c_base + side_off(pool, task.c_row, row);
#pragma unroll
                for (int j = 0; j < 2; j++)
                {
                    if (!col_ok[j]) continue;
                    float* p = C + ro + col_off[j];
                    const float res = CHUNKED ? sum[i][j][r] + acc[i][j][r] : acc[i][j][r];
                    double v = alpha * (double)res;
                    if (beta != 0.0) v = fma(beta, (double)*p, v);
                    *p = (float)v;
                }
            }
        }
    }
}

}  // namespace

int pack_bf16_rows() { return PR; }
int pack_bf16_k() { return PK; }

hipError_t launch_pack_bf16x3(const PackDesc& d, bool kfast, const int64_t* pool, const float* X,
                              __bf16* out, hipStream_t s)
{
    if (d.nrt == 0 || d.nkt == 0) return hipSuccess;
    dim3 grid(d.nrt, d.nkt);
    if (kfast)
        hipLaunchKernelGGL(pack_bf16x3_kernel<true>, grid, dim3(256), 0, s, d, pool, X, out);
    else
        hipLaunchKernelGGL(pack_bf16x3_kernel<false>, grid, dim3(256), 0, s, d, pool, X, out);
    return hipGetLastError();
}

hipError_t launch_gemm_bf16x3_packed(const GemmLaunch& L, bool chunked)
{
    if (L.ntasks == 0) return hipSuccess;
    const void* f = chunked ? (const void*)&gemm_bf16x3_packed_kernel<true>
                            : (const void*)&gemm_bf16x3_packed_kernel<false>;
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
    hipError_t e = hipLaunchKernel(f, dim3(L.grid), dim3(256), args, 0, L.stream);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

int gemm_bf16x3_packed_occupancy(bool chunked)
{
    const void* f = chunked ? (const void*)&gemm_bf16x3_packed_kernel<true>
                            : (const void*)&gemm_bf16x3_packed_kernel<false>;
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, 256, 0) != hipSuccess)
    {
        (void)hipGetLastError();
        return 1;
    }
    return n;
}

}  // namespace tblis_amd
