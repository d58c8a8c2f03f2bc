// Level-1 tensor operations around a contraction (gfx950):
//     add    B <- alpha * sum_{A-only} A + beta * B     (src/iface/1t/add.cxx)
//     scale  B <- beta * B                              (src/iface/1t/scale.cxx)
//     set    B <- value                                 (src/iface/1t/set.cxx)
// All three are HBM-bound.  B's dimensions arrive sorted by |B stride|
// (dimension 0 is the one consecutive lanes should store along) and the
// A-only dimensions by |A stride|.  Five kernels:
//   long rows  dimension 0 of at least 256: a block covers 2048 elements of
//              one row, eight per thread, all loads before the stores.
//   rows       a block covers R rows x W columns of dimension 0; the row
//              offsets (all other B dimensions) are decomposed once per row
//              into LDS, so the element loop has no 64-bit division.  A-only
//              sums run per element (accumulated in FP64), the inner loop
//              along the A-only dimension of smallest stride.
//   transpose  add with A's smallest stride on another dimension than B's:
//              a 64x64 tile is read along A's fast dimension into LDS and
//              written along B's, so both sides are coalesced.
//   reduce     add whose A-only sum is long and contiguous: one wave per B
//              element, lanes along the A-only dimension, shuffle reduction.
//   split      add with few B elements and a long, strided A-only sum: the
//              sum is cut into chunks (over the outer A-only dimensions, then
//              over the first), partial sums go to a workspace, and the reduce
//              kernel adds each element's chunks (a deterministic two-pass sum).
// beta == 0 writes B without reading it (the reference's rule,
// src/internal/1t/dense/add.cxx:65-90).
#include <hip/hip_runtime.h>

#include "level1.hpp"

namespace tblis_amd
{

namespace
{

constexpr int kThreads = 256;
constexpr int kRowElems = 1024;  // elements per block of the rows kernel
constexpr int kTile = 64;        // transpose tile edge
constexpr int64_t kMaxBlocks = 1 << 16;

// sum_{j < na} A[off + offset of j over the A-only dimensions]
template <typename TA>
__device__ inline double a_only_sum(const Level1Desc& d, const TA* __restrict__ A, int64_t off)
{
    const int64_t l0 = d.len_a[0], s0 = d.sa_a[0];
    const int64_t nouter = d.na / l0;
    double sum = 0.0;
    for (int64_t o = 0; o < nouter; o++)
    {
        int64_t base = off, r = o;
        for (int k = 1; k < d.nda; k++)
        {
            const int64_t q = r % d.len_a[k];
            r /= d.len_a[k];
            base += q * d.sa_a[k];
        }
        for (int64_t i = 0; i < l0; i++) sum += (double)A[base + i * s0];
    }
    return sum;
}

template <typename TA, typename T, int MODE>
__device__ inline void apply(const Level1Desc& d, const TA* __restrict__ A, T* __restrict__ B,
                             int64_t ob, int64_t oa, double alpha, double beta, double value)
{
    T* p = B + ob;
    if (MODE == LEVEL1_SET)
        *p = (T)value;
    else if (MODE == LEVEL1_SCALE)
        *p = (T)(beta * (double)*p);
    else
    {
        const double a = d.nda == 0 ? (double)A[oa] : a_only_sum(d, A, oa);
        double v = alpha * a;
        if (beta != 0.0) v += beta * (double)*p;
        *p = (T)v;
    }
}

// Offsets of row `row` (a multi-index over B dimensions 1..ndb-1, or over
// every dimension but p and q when skip >= 0).
__device__ inline void row_offsets(const Level1Desc& d, int64_t row, int first, int skip,
                                   int64_t& ob, int64_t& oa)
{
    ob = 0;
    oa = 0;
    for (int k = first; k < d.ndb; k++)
    {
        if (k == skip) continue;
        const int64_t q = row % d.len_b[k];
        row /= d.len_b[k];
        ob += q * d.sb[k];
        oa += q * d.sa_b[k];
    }
}

template <typename TA, typename T, int MODE>
__global__ __launch_bounds__(kThreads) void level1_rows_kernel(Level1Desc d, const TA* __restrict__ A,
                                                               T* __restrict__ B, double alpha,
                                                               double beta, double value)
{
    __shared__ int64_t s_ob[kRowElems], s_oa[kRowElems];
    const int64_t l0 = d.ndb > 0 ? d.len_b[0] : 1;
    const int64_t sb0 = d.ndb > 0 ? d.sb[0] : 0, sa0 = d.ndb > 0 ? d.sa_b[0] : 0;
    const int64_t nrows = d.nb / l0;
    const int W = (int)(l0 < kRowElems ? l0 : kRowElems);
    const int R = kRowElems / W;
    const int64_t nseg = (l0 + W - 1) / W;
    const int64_t nblocks = ((nrows + R - 1) / R) * nseg;
    for (int64_t blk = blockIdx.x; blk < nblocks; blk += gridDim.x)
    {
        const int64_t rb = blk / nseg, seg = blk - rb * nseg;
        const int64_t r0 = rb * R, c0 = seg * W;
        __syncthreads();  // the previous block's readers are done with s_ob / s_oa
        for (int t = threadIdx.x; t < R; t += kThreads)
            if (r0 + t < nrows) row_offsets(d, r0 + t, 1, -1, s_ob[t], s_oa[t]);
        __syncthreads();
        for (int e = threadIdx.x; e < R * W; e += kThreads)
        {
            const int r = e / W;
            const int64_t col = c0 + (e - r * W);
            if (r0 + r >= nrows || col >= l0) continue;
            apply<TA, T, MODE>(d, A, B, s_ob[r] + col * sb0, s_oa[r] + col * sa0, alpha, beta, value);
        }
    }
}

// Rows of at least kLongRow elements: a block covers kSeg consecutive
// elements of one row, kSeg / kThreads per thread with every load of a
// thread issued before its first store; no LDS, no barrier.
constexpr int kLongRow = 256;
constexpr int kPerThread = 8;
constexpr int kSeg = kThreads * kPerThread;

template <typename T, int MODE>
__global__ __launch_bounds__(kThreads) void level1_long_rows_kernel(Level1Desc d, const T* __restrict__ A,
                                                                    T* __restrict__ B, double alpha,
                                                                    double beta, double value)
{
    const int64_t l0 = d.len_b[0], sb0 = d.sb[0], sa0 = d.sa_b[0];
    const int64_t nrows = d.nb / l0, nseg = (l0 + kSeg - 1) / kSeg;
    const int64_t nblocks = nrows * nseg;
    for (int64_t blk = blockIdx.x; blk < nblocks; blk += gridDim.x)
    {
        const int64_t row = blk / nseg, seg = blk - row * nseg;
        int64_t ob, oa;
        row_offsets(d, row, 1, -1, ob, oa);
        const int64_t c0 = seg * kSeg + threadIdx.x;
        if (seg * kSeg + kSeg <= l0)
        {
            double av[kPerThread], bv[kPerThread];
#pragma unroll
            for (int j = 0; j < kPerThread; j++)
            {
                const int64_t c = c0 + j * kThreads;
                av[j] = MODE == LEVEL1_ADD ? (double)A[oa + c * sa0] : 0.0;
                bv[j] = (MODE == LEVEL1_SCALE || (MODE == LEVEL1_ADD && beta != 0.0)) ? (double)B[ob + c * sb0]
                                                                                        : 0.0;
            }
#pragma unroll
            for (int j = 0; j < kPerThread; j++)
            {
                const int64_t c = c0 + j * kThreads;
                double v;
                if (MODE == LEVEL1_SET)
                    v = value;
                else if (MODE == LEVEL1_SCALE)
                    v = beta * bv[j];
                else
                    v = alpha * av[j] + (beta != 0.0 ? beta * bv[j] : 0.0);
                B[ob + c * sb0] = (T)v;
            }
        }
        else
        {
            for (int j = 0; j < kPerThread; j++)
            {
                const int64_t c = c0 + j * kThreads;
                if (c < l0) apply<T, T, MODE>(d, A, B, ob + c * sb0, oa + c * sa0, alpha, beta, value);
            }
        }
    }
}

// B dims p = 0 (B's fast dimension) and q (A's fast dimension), tile 64x64.
template <typename T>
__global__ __launch_bounds__(kThreads) void level1_transpose_kernel(Level1Desc d, int q,
                                                                    const T* __restrict__ A,
                                                                    T* __restrict__ B, double alpha,
                                                                    double beta)
{
    __shared__ T tile[kTile][kTile + 1];
    const int64_t lp = d.len_b[0], lq = d.len_b[q];
    const int64_t ntp = (lp + kTile - 1) / kTile, ntq = (lq + kTile - 1) / kTile;
    const int64_t nouter = d.nb / (lp * lq);
    const int64_t nblocks = ntp * ntq * nouter;
    const int lane = threadIdx.x % kTile, row = threadIdx.x / kTile;  // 4 rows of 64
    for (int64_t blk = blockIdx.x; blk < nblocks; blk += gridDim.x)
    {
        const int64_t o = blk / (ntp * ntq), tpq = blk - o * (ntp * ntq);
        const int64_t tq = tpq / ntp, tp = tpq - tq * ntp;
        int64_t ob, oa;
        row_offsets(d, o, 1, q, ob, oa);
        const bool full = (tp + 1) * kTile <= lp && (tq + 1) * kTile <= lq;
        constexpr int kRows = kThreads / kTile;  // tile rows per pass
        __syncthreads();  // the previous tile has been written out
        // read: lanes along q (A's smallest stride)
        const int64_t qa = tq * kTile + lane;
        const T* a = A + oa + qa * d.sa_b[q];
        if (full)
        {
            T v[kTile / kRows];
#pragma unroll
            for (int i = 0; i < kTile / kRows; i++) v[i] = a[(tp * kTile + row + i * kRows) * d.sa_b[0]];
#pragma unroll
            for (int i = 0; i < kTile / kRows; i++) tile[row + i * kRows][lane] = v[i];
        }
        else
        {
            for (int i = row; i < kTile; i += kRows)
            {
                const int64_t pa = tp * kTile + i;
                if (pa < lp && qa < lq) tile[i][lane] = a[pa * d.sa_b[0]];
            }
        }
        __syncthreads();
        // write: lanes along p (B's smallest stride)
        const int64_t pb = tp * kTile + lane;
        T* b = B + ob + pb * d.sb[0];
        if (full)
        {
            double bv[kTile / kRows];
#pragma unroll
            for (int i = 0; i < kTile / kRows; i++)
                bv[i] = beta != 0.0 ? (double)b[(tq * kTile + row + i * kRows) * d.sb[q]] : 0.0;
#pragma unroll
            for (int i = 0; i < kTile / kRows; i++)
            {
                const double v = alpha * (double)tile[lane][row + i * kRows] + (beta != 0.0 ? beta * bv[i] : 0.0);
                b[(tq * kTile + row + i * kRows) * d.sb[q]] = (T)v;
            }
        }
        else
        {
            for (int i = row; i < kTile; i += kRows)
            {
                const int64_t qb = tq * kTile + i;
                if (pb >= lp || qb >= lq) continue;
                T* p = b + qb * d.sb[q];
                double v = alpha * (double)tile[lane][i];
                if (beta != 0.0) v += beta * (double)*p;
                *p = (T)v;
            }
        }
    }
}

// One wave per B element; lanes along A-only dimension 0, four independent
// accumulators per lane.
template <typename TA, typename T>
__global__ __launch_bounds__(kThreads) void level1_reduce_kernel(Level1Desc d, const TA* __restrict__ A,
                                                                 T* __restrict__ B, double alpha,
                                                                 double beta)
{
    const int lane = threadIdx.x % 64;
    const int64_t waves = (int64_t)gridDim.x * (kThreads / 64);
    const int64_t l0 = d.len_a[0], s0 = d.sa_a[0], nouter = d.na / l0;
    for (int64_t w = (int64_t)blockIdx.x * (kThreads / 64) + threadIdx.x / 64; w < d.nb; w += waves)
    {
        int64_t ob, oa;
        row_offsets(d, w, 0, -1, ob, oa);
        double sum = 0.0;
        for (int64_t o = 0; o < nouter; o++)
        {
            int64_t base = oa, r = o;
            for (int k = 1; k < d.nda; k++)
            {
                const int64_t qk = r % d.len_a[k];
                r /= d.len_a[k];
                base += qk * d.sa_a[k];
            }
            double s1 = 0.0, s2 = 0.0, s3 = 0.0;
            int64_t i = lane;
            for (; i + 192 < l0; i += 256)
            {
                sum += (double)A[base + i * s0];
                s1 += (double)A[base + (i + 64) * s0];
                s2 += (double)A[base + (i + 128) * s0];
                s3 += (double)A[base + (i + 192) * s0];
            }
            for (; i < l0; i += 64) sum += (double)A[base + i * s0];
            sum += (s1 + s2) + s3;
        }
        for (int m = 32; m > 0; m >>= 1) sum += __shfl_xor(sum, m, 64);
        if (lane == 0)
        {
            T* p = B + ob;
            double v = alpha * sum;
            if (beta != 0.0) v += beta * (double)*p;
            *p = (T)v;
        }
    }
}

// Chunks of the A-only sum: `so` ranges over the outer A-only dimensions
// (1..nda-1), each cut into `si` ranges of dimension 0.
struct Split
{
    int64_t so = 0, si = 0, po = 0, pi = 0;
    __host__ __device__ int64_t chunks() const { return so * si; }
};

// Partial sums: part[c * nb + i] = sum of B element i's A-only terms in chunk c
// (i in B's iteration order, dimension 0 fastest).
template <typename T>
__global__ __launch_bounds__(kThreads) void level1_split_kernel(Level1Desc d, Split sp,
                                                                const T* __restrict__ A,
                                                                double* __restrict__ part)
{
    const int64_t total = sp.chunks() * d.nb;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t l0 = d.len_a[0], s0 = d.sa_a[0], nouter = d.na / l0;
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < total; x += stride)
    {
        const int64_t c = x / d.nb, i = x - c * d.nb;
        const int64_t co = c / sp.si, ci = c - co * sp.si;
        int64_t ob, oa;
        row_offsets(d, i, 0, -1, ob, oa);
        const int64_t o_end = (co + 1) * sp.po < nouter ? (co + 1) * sp.po : nouter;
        const int64_t i_end = (ci + 1) * sp.pi < l0 ? (ci + 1) * sp.pi : l0;
        double sum = 0.0;
        for (int64_t o = co * sp.po; o < o_end; o++)
        {
            int64_t base = oa, r = o;
            for (int k = 1; k < d.nda; k++)
            {
                const int64_t q = r % d.len_a[k];
                r /= d.len_a[k];
                base += q * d.sa_a[k];
            }
            for (int64_t j = ci * sp.pi; j < i_end; j++) sum += (double)A[base + j * s0];
        }
        part[x] = sum;
    }
}

// Beta pass of an indexed contraction over C entries (same dense shape,
// entry offsets offs[e], factors fac[e]): beta == 0 or a zero factor writes
// 0 without reading; otherwise data *= fac[e] * beta
// (src/internal/1t/indexed/scale.cxx:12-30, set.cxx:12-20).  A block covers
// kSeg elements of one row (dimension 0) of one entry, like the long-rows
// kernel.
template <typename T>
__global__ __launch_bounds__(kThreads) void level1_entries_kernel(Level1Desc d, const int64_t* __restrict__ offs,
                                                                  const double* __restrict__ fac, int64_t nent,
                                                                  double beta, T* __restrict__ B)
{
    const int64_t l0 = d.ndb > 0 ? d.len_b[0] : 1, sb0 = d.ndb > 0 ? d.sb[0] : 0;
    const int64_t nrows = d.nb / l0, nseg = (l0 + kSeg - 1) / kSeg;
    const int64_t per = nrows * nseg, nblocks = nent * per;
    for (int64_t blk = blockIdx.x; blk < nblocks; blk += gridDim.x)
    {
        const int64_t e = blk / per, r = blk - e * per;
        const int64_t row = r / nseg, seg = r - row * nseg;
        int64_t ob, oa;
        row_offsets(d, row, 1, -1, ob, oa);
        T* base = B + offs[e] + ob;
        const double f = fac[e];
        const bool zero = beta == 0.0 || f == 0.0;
        const double c = f * beta;
#pragma unroll
        for (int j = 0; j < kPerThread; j++)
        {
            const int64_t col = seg * kSeg + threadIdx.x + j * kThreads;
            if (col >= l0) break;
            T* p = base + col * sb0;
            *p = zero ? (T)0 : (T)(c * (double)*p);
        }
    }
}

// C <- beta C for an interleaved complex tensor (d in complex elements):
// beta == 0 writes zeros without reading.
template <typename T>
__global__ __launch_bounds__(kThreads) void level1_cscale_kernel(Level1Desc d, double br, double bi,
                                                                 T* __restrict__ B)
{
    const int64_t l0 = d.ndb > 0 ? d.len_b[0] : 1, sb0 = d.ndb > 0 ? d.sb[0] : 0;
    const int64_t nrows = d.nb / l0, nseg = (l0 + kSeg - 1) / kSeg;
    const bool zero = br == 0.0 && bi == 0.0;
    for (int64_t blk = blockIdx.x; blk < nrows * nseg; blk += gridDim.x)
    {
        const int64_t row = blk / nseg, seg = blk - row * nseg;
        int64_t ob, oa;
        row_offsets(d, row, 1, -1, ob, oa);
#pragma unroll
        for (int j = 0; j < kPerThread; j++)
        {
            const int64_t col = seg * kSeg + threadIdx.x + j * kThreads;
            if (col >= l0) break;
            T* p = B + 2 * (ob + col * sb0);
            if (zero)
            {
                p[0] = (T)0;
                p[1] = (T)0;
            }
            else
            {
                const double re = (double)p[0], im = (double)p[1];
                p[0] = (T)(br * re - bi * im);
                p[1] = (T)(br * im + bi * re);
            }
        }
    }
}

// C_e <- (f_e s) C_e over the entries of an interleaved complex indexed
// tensor (s = sr + i si; f_e = fre[e] + i fim[e], or 1 when fre is null);
// a zero product writes zeros without reading.
template <typename T>
__global__ __launch_bounds__(kThreads) void level1_centries_kernel(Level1Desc d, const int64_t* __restrict__ offs,
                                                                   const double* __restrict__ fre,
                                                                   const double* __restrict__ fim, int64_t nent,
                                                                   double sr, double si, T* __restrict__ B)
{
    const int64_t l0 = d.ndb > 0 ? d.len_b[0] : 1, sb0 = d.ndb > 0 ? d.sb[0] : 0;
    const int64_t nrows = d.nb / l0, nseg = (l0 + kSeg - 1) / kSeg;
    const int64_t per = nrows * nseg, nblocks = nent * per;
    for (int64_t blk = blockIdx.x; blk < nblocks; blk += gridDim.x)
    {
        const int64_t e = blk / per, r = blk - e * per;
        const int64_t row = r / nseg, seg = r - row * nseg;
        int64_t ob, oa;
        row_offsets(d, row, 1, -1, ob, oa);
        T* base = B + 2 * (offs[e] + ob);
        const double fr = fre ? fre[e] : 1.0, fi = fre ? fim[e] : 0.0;
        const double cr = fr * sr - fi * si, ci = fr * si + fi * sr;
        const bool zero = cr == 0.0 && ci == 0.0;
#pragma unroll
        for (int j = 0; j < kPerThread; j++)
        {
            const int64_t col = seg * kSeg + threadIdx.x + j * kThreads;
            if (col >= l0) break;
            T* p = base + 2 * col * sb0;
            if (zero)
            {
                p[0] = (T)0;
                p[1] = (T)0;
            }
            else
            {
                const double re = (double)p[0], im = (double)p[1];
                p[0] = (T)(cr * re - ci * im);
                p[1] = (T)(cr * im + ci * re);
            }
        }
    }
}

// Block copies between layouts (indexed DPD expansion and write-back): one
// workgroup per piece of a block, the elements decomposed over the block's
// dimensions.
template <typename T>
__global__ __launch_bounds__(kThreads) void block_copy_kernel(const BlockCopy* __restrict__ tasks, int64_t ntasks,
                                                              const T* __restrict__ src, T* __restrict__ dst)
{
    for (int64_t t = blockIdx.x; t < ntasks; t += gridDim.x)
    {
        const BlockCopy& b = tasks[t];
        const int nd = b.nd;
        for (int64_t e = b.e0 + threadIdx.x; e < b.e0 + b.n; e += kThreads)
        {
            int64_t r = e, so = b.src, dof = b.dst;
            for (int d = 0; d < nd; d++)
            {
                const int64_t q = r % b.len[d];
                r /= b.len[d];
                so += q * b.ss[d];
                dof += q * b.ds[d];
            }
            dst[dof] = src[so];
        }
    }
}

unsigned grid_for(int64_t work)
{
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(work, kMaxBlocks));
}

int64_t abs64(int64_t x) { return x < 0 ? -x : x; }

constexpr int64_t kSplitMaxB = 1 << 18;   // B elements below which a long sum is split
constexpr int64_t kSplitTarget = 1 << 21;  // partial sums aimed for
constexpr int64_t kSplitMaxChunks = 512;   // the combining pass sums at most this many

// Whether (and how) an add's A-only sum is split; so == 0 when it is not.
Split plan_split(int mode, const Level1Desc& d)
{
    Split sp;
    if (mode != LEVEL1_ADD || d.nda == 0 || d.nb >= kSplitMaxB || d.na < 256) return sp;
    const bool rows_coalesced = d.ndb > 0 && abs64(d.sa_b[0]) == 1 && d.nb >= 65536;
    if (d.len_a[0] >= 32 && abs64(d.sa_a[0]) <= 2 && !rows_coalesced) return sp;  // reduce kernel
    const int64_t want = std::min(kSplitMaxChunks, std::max<int64_t>(2, kSplitTarget / d.nb));
    const int64_t l0 = d.len_a[0], nouter = d.na / l0;
    sp.so = std::min(nouter, want);
    sp.si = std::min(l0, std::max<int64_t>(1, want / sp.so));
    sp.po = (nouter + sp.so - 1) / sp.so;
    sp.pi = (l0 + sp.si - 1) / sp.si;
    sp.so = (nouter + sp.po - 1) / sp.po;  // no empty chunks
    sp.si = (l0 + sp.pi - 1) / sp.pi;
    return sp;
}

template <typename T>
hipError_t launch_t(int mode, const Level1Desc& d, const void* A, void* B, double alpha, double beta,
                    double value, void* ws, hipStream_t s)
{
    if (d.nb == 0) return hipSuccess;
    const T* a = (const T*)A;
    T* b = (T*)B;
    const int64_t l0 = d.ndb > 0 ? d.len_b[0] : 1;
    const int64_t W = std::min<int64_t>(l0, kRowElems), R = kRowElems / W;
    const unsigned rows_grid = grid_for(((d.nb / l0 + R - 1) / R) * ((l0 + W - 1) / W));
    const bool long_rows = d.ndb > 0 && l0 >= kLongRow;
    const unsigned long_grid = grid_for((d.nb / l0) * ((l0 + kSeg - 1) / kSeg));
    if (mode == LEVEL1_SET)
    {
        if (long_rows)
            hipLaunchKernelGGL((level1_long_rows_kernel<T, LEVEL1_SET>), dim3(long_grid), dim3(kThreads), 0,
                               s, d, a, b, alpha, beta, value);
        else
            hipLaunchKernelGGL((level1_rows_kernel<T, T, LEVEL1_SET>), dim3(rows_grid), dim3(kThreads), 0,
                               s, d, a, b, alpha, beta, value);
        return hipGetLastError();
    }
    if (mode == LEVEL1_SCALE)
    {
        if (long_rows)
            hipLaunchKernelGGL((level1_long_rows_kernel<T, LEVEL1_SCALE>), dim3(long_grid), dim3(kThreads),
                               0, s, d, a, b, alpha, beta, value);
        else
            hipLaunchKernelGGL((level1_rows_kernel<T, T, LEVEL1_SCALE>), dim3(rows_grid), dim3(kThreads),
                               0, s, d, a, b, alpha, beta, value);
        return hipGetLastError();
    }
    if (d.nda > 0)
    {
        // a long A-only sum with a contiguous fast dimension: wave per element,
        // unless consecutive B elements already read consecutive A elements
        const bool rows_coalesced = d.ndb > 0 && abs64(d.sa_b[0]) == 1 && d.nb >= 65536;
        if (d.len_a[0] >= 32 && abs64(d.sa_a[0]) <= 2 && !rows_coalesced)
        {
            hipLaunchKernelGGL((level1_reduce_kernel<T, T>), dim3(grid_for((d.nb + 3) / 4)), dim3(kThreads),
                               0, s, d, a, b, alpha, beta);
            return hipGetLastError();
        }
        const Split sp = plan_split(mode, d);
        if (sp.so > 0)
        {
            if (!ws) return hipErrorInvalidValue;
            double* part = (double*)ws;
            hipLaunchKernelGGL((level1_split_kernel<T>), dim3(grid_for((sp.chunks() * d.nb + 255) / 256)),
                               dim3(kThreads), 0, s, d, sp, a, part);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
            // B <- alpha * sum_c part[c][i] + beta * B: the partials as an A
            // whose only A-only dimension is the chunk, one wave per element
            Level1Desc c = d;
            int64_t acc = 1;
            for (int k = 0; k < c.ndb; k++)
            {
                c.sa_b[k] = acc;
                acc *= c.len_b[k];
            }
            c.nda = 1;
            c.len_a[0] = sp.chunks();
            c.sa_a[0] = d.nb;
            c.na = sp.chunks();
            hipLaunchKernelGGL((level1_reduce_kernel<double, T>), dim3(grid_for((d.nb + 3) / 4)),
                               dim3(kThreads), 0, s, c, (const double*)part, b, alpha, beta);
            return hipGetLastError();
        }
    }
    else if (d.ndb >= 2 && abs64(d.sa_b[0]) > 1)
    {
        // A's smallest stride on another dimension: transpose through LDS
        int q = -1;
        for (int k = 1; k < d.ndb; k++)
            if (d.sa_b[k] != 0 && (q < 0 || abs64(d.sa_b[k]) < abs64(d.sa_b[q]))) q = k;
        if (q > 0 && abs64(d.sa_b[q]) < abs64(d.sa_b[0]) && d.len_b[0] >= 16 && d.len_b[q] >= 16)
        {
            const int64_t tiles = ((d.len_b[0] + kTile - 1) / kTile) * ((d.len_b[q] + kTile - 1) / kTile) *
                                  (d.nb / (d.len_b[0] * d.len_b[q]));
            hipLaunchKernelGGL((level1_transpose_kernel<T>), dim3(grid_for(tiles)), dim3(kThreads), 0, s,
                               d, q, a, b, alpha, beta);
            return hipGetLastError();
        }
    }
    if (long_rows && d.nda == 0)
        hipLaunchKernelGGL((level1_long_rows_kernel<T, LEVEL1_ADD>), dim3(long_grid), dim3(kThreads), 0, s,
                           d, a, b, alpha, beta, value);
    else
        hipLaunchKernelGGL((level1_rows_kernel<T, T, LEVEL1_ADD>), dim3(rows_grid), dim3(kThreads), 0, s,
                           d, a, b, alpha, beta, value);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_level1_entries(int dtype, const Level1Desc& d, const int64_t* offs, const double* fac,
                                int64_t nent, double beta, void* B, hipStream_t s)
{
    if (nent == 0 || d.nb == 0) return hipSuccess;
    const int64_t l0 = d.ndb > 0 ? d.len_b[0] : 1;
    const unsigned grid = grid_for(nent * (d.nb / l0) * ((l0 + kSeg - 1) / kSeg));
    if (dtype == 0)
        hipLaunchKernelGGL((level1_entries_kernel<double>), dim3(grid), dim3(kThreads), 0, s, d, offs, fac, nent,
                           beta, (double*)B);
    else
        hipLaunchKernelGGL((level1_entries_kernel<float>), dim3(grid), dim3(kThreads), 0, s, d, offs, fac, nent,
                           beta, (float*)B);
    return hipGetLastError();
}

hipError_t launch_level1_centries(int dtype, const Level1Desc& d, const int64_t* offs, const double* fre,
                                  const double* fim, int64_t nent, double sr, double si, void* B, hipStream_t s)
{
    if (nent == 0 || d.nb == 0) return hipSuccess;
    const int64_t l0 = d.ndb > 0 ? d.len_b[0] : 1;
    const unsigned grid = grid_for(nent * (d.nb / l0) * ((l0 + kSeg - 1) / kSeg));
    if (dtype == 0)
        hipLaunchKernelGGL((level1_centries_kernel<double>), dim3(grid), dim3(kThreads), 0, s, d, offs, fre, fim,
                           nent, sr, si, (double*)B);
    else
        hipLaunchKernelGGL((level1_centries_kernel<float>), dim3(grid), dim3(kThreads), 0, s, d, offs, fre, fim,
                           nent, sr, si, (float*)B);
    return hipGetLastError();
}

hipError_t launch_level1_cscale(int dtype, const Level1Desc& d, double br, double bi, void* B, hipStream_t s)
{
    if (d.nb == 0) return hipSuccess;
    const int64_t l0 = d.ndb > 0 ? d.len_b[0] : 1;
    const unsigned grid = grid_for((d.nb / l0) * ((l0 + kSeg - 1) / kSeg));
    if (dtype == 0)
        hipLaunchKernelGGL((level1_cscale_kernel<double>), dim3(grid), dim3(kThreads), 0, s, d, br, bi, (double*)B);
    else
        hipLaunchKernelGGL((level1_cscale_kernel<float>), dim3(grid), dim3(kThreads), 0, s, d, br, bi, (float*)B);
    return hipGetLastError();
}

hipError_t launch_block_copy(int dtype, const BlockCopy* tasks, int64_t ntasks, const void* src, void* dst,
                             hipStream_t s)
{
    if (ntasks == 0) return hipSuccess;
    const unsigned grid = grid_for(ntasks);
    if (dtype == 0)
        hipLaunchKernelGGL((block_copy_kernel<double>), dim3(grid), dim3(kThreads), 0, s, tasks, ntasks,
                           (const double*)src, (double*)dst);
    else if (dtype == 1)
        hipLaunchKernelGGL((block_copy_kernel<float>), dim3(grid), dim3(kThreads), 0, s, tasks, ntasks,
                           (const float*)src, (float*)dst);
    else if (dtype == 2)  // dcomplex: (re, im) pairs
        hipLaunchKernelGGL((block_copy_kernel<double2>), dim3(grid), dim3(kThreads), 0, s, tasks, ntasks,
                           (const double2*)src, (double2*)dst);
    else
        hipLaunchKernelGGL((block_copy_kernel<float2>), dim3(grid), dim3(kThreads), 0, s, tasks, ntasks,
                           (const float2*)src, (float2*)dst);
    return hipGetLastError();
}

size_t level1_workspace_bytes(int mode, const Level1Desc& d)
{
    const Split sp = plan_split(mode, d);
    return sp.so > 0 ? (size_t)(sp.chunks() * d.nb) * sizeof(double) : 0;
}

hipError_t launch_level1(int mode, int dtype, const Level1Desc& d, const void* A, void* B,
                         double alpha, double beta, double value, void* ws, hipStream_t s)
{
    if (dtype == 0) return launch_t<double>(mode, d, A, B, alpha, beta, value, ws, s);
    return launch_t<float>(mode, d, A, B, alpha, beta, value, ws, s);
}

}  // namespace tblis_amd
