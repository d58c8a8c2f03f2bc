// Split-BF16 MFMA over fragment-ordered panels (gfx950).
//
// FP32 operands are split once, at packing time, into three BF16 terms
// x = hi + mid + lo (hi = RN(x), mid = RN(x - hi), lo = RN(x - hi - mid)),
// and every C element accumulates the six products that carry FP32
// accuracy, smallest first: mid*mid, hi*lo, lo*hi, hi*mid, mid*hi, hi*hi,
// with v_mfma_f32_32x32x16_bf16 (FP32 accumulation).  With CHUNKED the
// products of every 256 k are summed in a fresh accumulator that is then
// added to a second one, which keeps the FP32 rounding of long k sums near
// that of a 256-term sum.
//
// Panels use the FP64 path's jobs and tasks (plan_types.hpp) with a larger
// fragment group: 32 rows x 16 k, three terms, group (rb, ks) at
//     out + ((rb * pnkg + ks) * 3 + term) * 512,
// lane l holding X[32 rb + l % 32][16 ks + 8 (l / 32) + j], j = 0..7, as one
// 16-byte vector -- the A and B operand of v_mfma_f32_32x32x16_bf16.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace tblis_amd
{

namespace
{

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ void split3(float x, __bf16& h, __bf16& m, __bf16& l)
{
    h = (__bf16)x;
    const float r1 = x - (float)h;
    m = (__bf16)r1;
    const float r2 = r1 - (float)m;
    l = (__bf16)r2;
}

// ------------------------------------------------------------------ pack

// As the FP64 pack: every address of a chunk follows from its PackChunk record.
// Loads per batch: one pair of k groups (k-fast sources) or two groups
// (others) -- 66 registers, 7 waves per SIMD; twice that batch took 96
// registers (5 waves per SIMD) and was 22% slower on the 4096-instance batch
// and on 96^4.
__global__ __launch_bounds__(256) void pack_bf16_panel_kernel(const PackChunk* __restrict__ chunks,
                                                              const int64_t* __restrict__ pool,
                                                              const float* __restrict__ A,
                                                              const float* __restrict__ B,
                                                              __bf16* __restrict__ ws)
{
    const PackChunk& j = chunks[xcd_contiguous_chunk(blockIdx.x, gridDim.x)];
    const float* __restrict__ X = j.operand ? B : A;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lr = lane & 31, lh = lane >> 5;
    const bool zero = j.src == PANEL_MISSING;
    const int ng = j.nrb * j.nkg;
    const float scale = (float)j.scale;
    if (j.kfast && !zero)
    {
        // k-fast source: per pair of k groups, 16 loads of two rows x 32
        // consecutive k (whole 128-byte lines), transposed through a per-wave
        // LDS tile into the fragment order.  The k tables are unit-stride.
        // The loads of kBatch pairs are issued before the first transpose.
        constexpr int kBatch = 1;
        __shared__ float tile[4][32][33];
        const int npair = (j.nkg + 1) / 2;
        const int niter = j.nrb * npair;
        const int kk = lane & 31;
        for (int g0 = wave; g0 < niter; g0 += 4 * kBatch)
        {
            float v[kBatch][16];
#pragma unroll
            for (int u = 0; u < kBatch; u++)
            {
                const int g = g0 + 4 * u;
#pragma unroll
                for (int jj = 0; jj < 16; jj++) v[u][jj] = 0.0f;
                if (g >= niter) continue;
                const int ri = g / npair, ki = 2 * (g % npair);  // groups within the chunk
                const int rb = j.rb0 + ri;
                const int ks = j.kg0 + ki;
                const bool two = ki + 1 < j.nkg;
                const int64_t rbase = j.rt[ri][0];
                const int64_t rstr = j.rt[ri][1];
                const int h = kk >> 4;
                const int64_t kb = j.kt[ki + (two ? h : 0)][0];
                const int pk = ks * 16 + kk - j.k0;  // job-relative k
                const bool kin = pk >= 0 && pk < j.nk && (h == 0 || two);
                const int64_t koff = kb + (kk & 15);
#pragma unroll
                for (int jj = 0; jj < 16; jj++)
                {
                    const int r = (lane >> 5) + 2 * jj;
                    const int prow = rb * 32 + r - j.r0;
                    if (kin && prow >= 0 && prow < j.nr)
                    {
                        const int64_t roff = rstr != SCATTERED ? rbase + r * rstr : pool[j.rowmap + prow];
                        v[u][jj] = X[j.src + roff + koff] * scale;
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < kBatch; u++)
            {
                const int g = g0 + 4 * u;
                if (g >= niter) break;
                const int rb = j.rb0 + g / npair;
                const int ks = j.kg0 + 2 * (g % npair);
                const bool two = ks + 1 < j.kg0 + j.nkg;
#pragma unroll
                for (int jj = 0; jj < 16; jj++) tile[wave][(lane >> 5) + 2 * jj][kk] = v[u][jj];
                __builtin_amdgcn_wave_barrier();
                const int row = rb * 32 + lr - j.r0;
                const bool rin = row >= 0 && row < j.nr;
                for (int hh = 0; hh < (two ? 2 : 1); hh++)
                {
                    const int kb0 = (ks + hh) * 16 + lh * 8 - j.k0;
                    bf16x8 hv, mv, lv;
                    unsigned inmask = 0;
#pragma unroll
                    for (int e = 0; e < 8; e++)
                    {
                        const int k = kb0 + e;
                        const bool in = rin && k >= 0 && k < j.nk;
                        __bf16 a, b, c;
                        split3(tile[wave][lr][16 * hh + 8 * lh + e], a, b, c);
                        hv[e] = a;
                        mv[e] = b;
                        lv[e] = c;
                        inmask |= (unsigned)in << e;
                    }
                    __bf16* out = ws + j.out + ((int64_t)rb * j.pnkg + ks + hh) * 1536 + lane * 8;
                    if (inmask == 0xffu)
                    {
                        *(bf16x8*)out = hv;
                        *(bf16x8*)(out + 512) = mv;
                        *(bf16x8*)(out + 1024) = lv;
                    }
                    else if (inmask)
                    {
#pragma unroll
                        for (int e = 0; e < 8; e++)
                            if (inmask >> e & 1u)
                            {
                                out[e] = hv[e];
                                out[512 + e] = mv[e];
                                out[1024 + e] = lv[e];
                            }
                    }
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
        return;
    }
    // general source: one fragment group per iteration, the loads of
    // kBatch groups issued before their stores
    constexpr int kBatch = 2;
    for (int g0 = wave; g0 < ng; g0 += 4 * kBatch)
    {
        float v[kBatch][8];
#pragma unroll
        for (int u = 0; u < kBatch; u++)
        {
            const int g = g0 + 4 * u;
#pragma unroll
            for (int e = 0; e < 8; e++) v[u][e] = 0.0f;
            if (g >= ng || zero) continue;
            const int ri = g / j.nkg, ki = g % j.nkg;  // groups within the chunk
            const int rb = j.rb0 + ri;
            const int ks = j.kg0 + ki;
            const int row = rb * 32 + lr - j.r0;
            const bool rin = row >= 0 && row < j.nr;
            const int kb = ks * 16 + lh * 8 - j.k0;
            // group tables: wave-uniform (base, stride) of this row / k group
            const int64_t rbase = j.rt[ri][0];
            const int64_t rstr = j.rt[ri][1];
            const int64_t kbase = j.kt[ki][0];
            const int64_t kstr = j.kt[ki][1];
            if (!rin) continue;
            const int64_t roff = j.src + (rstr != SCATTERED ? rbase + lr * rstr : pool[j.rowmap + row]);
#pragma unroll
            for (int e = 0; e < 8; e++)
            {
                const int k = kb + e;
                if (k >= 0 && k < j.nk)
                    v[u][e] = X[roff + (kstr != SCATTERED ? kbase + (lh * 8 + e) * kstr : pool[j.kmap + k])] * scale;
            }
        }
#pragma unroll
        for (int u = 0; u < kBatch; u++)
        {
            const int g = g0 + 4 * u;
            if (g >= ng) break;
            const int rb = j.rb0 + g / j.nkg;
            const int ks = j.kg0 + g % j.nkg;
            const int row = rb * 32 + lr - j.r0;
            const bool rin = row >= 0 && row < j.nr;
            const int kb = ks * 16 + lh * 8 - j.k0;
            bf16x8 h, m, l;
            unsigned inmask = 0;
#pragma unroll
            for (int e = 0; e < 8; e++)
            {
                const int k = kb + e;
                const bool in = rin && k >= 0 && k < j.nk;
                __bf16 a, b, c;
                split3(v[u][e], a, b, c);
                h[e] = a;
                m[e] = b;
                l[e] = c;
                inmask |= (unsigned)in << e;
            }
            __bf16* out = ws + j.out + ((int64_t)rb * j.pnkg + ks) * 1536 + lane * 8;
            if (inmask == 0xffu)
            {
                *(bf16x8*)out = h;
                *(bf16x8*)(out + 512) = m;
                *(bf16x8*)(out + 1024) = l;
            }
            else if (inmask)
            {
#pragma unroll
                for (int e = 0; e < 8; e++)
                    if (inmask >> e & 1u)
                    {
                        out[e] = h[e];
                        out[512 + e] = m[e];
                        out[1024 + e] = l[e];
                    }
            }
        }
    }
}

// ------------------------------------------------------------------ gemm

template <int MB, int NB>
__device__ __forceinline__ void load_frags(bf16x8 (&a)[MB][3], bf16x8 (&b)[NB][3], const uint4* pa,
                                           const uint4* pb, int64_t rs, int ks)
{
#pragma unroll
    for (int i = 0; i < MB; i++)
#pragma unroll
        for (int t = 0; t < 3; t++)
        {
            uint4 v = pa[i * rs + ks * 192 + t * 64];
            a[i][t] = __builtin_bit_cast(bf16x8, v);
        }
#pragma unroll
    for (int jb = 0; jb < NB; jb++)
#pragma unroll
        for (int t = 0; t < 3; t++)
        {
            uint4 v = pb[jb * rs + ks * 192 + t * 64];
            b[jb][t] = __builtin_bit_cast(bf16x8, v);
        }
}

// Six products per block pair, interleaved across blocks so that
// consecutive MFMAs on one accumulator are MB*NB instructions apart.
template <int MB, int NB>
__device__ __forceinline__ void mma_frags(f32x16 (&acc)[MB][NB], const bf16x8 (&a)[MB][3],
                                          const bf16x8 (&b)[NB][3])
{
    constexpr int PA[6] = {1, 0, 2, 0, 1, 0};
    constexpr int PB[6] = {1, 2, 0, 1, 0, 0};
#pragma unroll
    for (int p = 0; p < 6; p++)
#pragma unroll
        for (int i = 0; i < MB; i++)
#pragma unroll
            for (int jb = 0; jb < NB; jb++)
                acc[i][jb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][PA[p]], b[jb][PB[p]],
                                                                    acc[i][jb], 0, 0, 0);
}

// An FP32 operand read straight from its tensor (PanelTask "direct"): the
// lane's row offset per 32-row block (taken once per task) and the k
// tables, (base, stride) per 16-k group or per-element offsets for groups
// that are not affine.  Each 16-k step a lane loads its 8 consecutive k of
// every block and splits them into the three BF16 terms in registers, so
// no pack pass runs.
template <int NBLK>
struct DirectF32
{
    __amdgpu_buffer_rsrc_t rsrc;
    const int64_t* kt;
    const int64_t* km;
    uint32_t roff[NBLK];
    bool rvalid[NBLK];
};

template <int NBLK>
__device__ __forceinline__ void direct_f32_init(DirectF32<NBLK>& d, const float* X, int64_t src,
                                                uint32_t bytes, const int64_t* __restrict__ pool,
                                                int64_t rows, int64_t kt, int64_t km, int rows0, int len,
                                                int lr)
{
    d.rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)(X + src), (short)0, (int)bytes, 0x00020000);
    d.kt = pool + kt;
    d.km = pool + km;
#pragma unroll
    for (int i = 0; i < NBLK; i++)
    {
        const int r = rows0 + i * 32 + lr;
        d.rvalid[i] = r < len;
        d.roff[i] = d.rvalid[i] ? (uint32_t)pool[rows + r] : 0u;
    }
}

// Raw values of one 16-k step: lane (lr, lh) gets row lr of each block at
// k = 16 ks + 8 lh + e, e = 0..7 (zero past the valid rows or k).
template <int NBLK>
__device__ __forceinline__ void direct_f32_load(float (&v)[NBLK][8], const DirectF32<NBLK>& d, int ks, int lh,
                                                int klen)
{
    const int64_t kb = d.kt[2 * ks], kst = d.kt[2 * ks + 1];
    const int k0 = ks * 16 + lh * 8;
    uint32_t o[8];
    if (kst != SCATTERED)  // wave-uniform
    {
#pragma unroll
        for (int e = 0; e < 8; e++) o[e] = (uint32_t)(kb + (lh * 8 + e) * kst);
    }
    else
    {
#pragma unroll
        for (int e = 0; e < 8; e++) o[e] = k0 + e < klen ? (uint32_t)d.km[k0 + e] : 0u;
    }
#pragma unroll
    for (int i = 0; i < NBLK; i++)
#pragma unroll
        for (int e = 0; e < 8; e++)
        {
            const uint32_t off = d.rvalid[i] && k0 + e < klen ? (d.roff[i] + o[e]) * 4u : 0xffffffffu;
            v[i][e] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(d.rsrc, off, 0, 0));
        }
}

template <int NBLK>
__device__ __forceinline__ void split_frags(bf16x8 (&x)[NBLK][3], const float (&v)[NBLK][8])
{
#pragma unroll
    for (int i = 0; i < NBLK; i++)
#pragma unroll
        for (int e = 0; e < 8; e++)
        {
            __bf16 h, m, l;
            split3(v[i][e], h, m, l);
            x[i][0][e] = h;
            x[i][1][e] = m;
            x[i][2][e] = l;
        }
}

// One operand's registers for one 16-k step: split fragments loaded from a
// panel, or raw FP32 values loaded from the tensor.
template <int NBLK, bool D>
struct Stage
{
    bf16x8 f[NBLK][3];
};
template <int NBLK>
struct Stage<NBLK, true>
{
    float v[NBLK][8];
};

template <int NBLK, bool D>
__device__ __forceinline__ void stage_load(Stage<NBLK, D>& st, const uint4* p, int64_t rs, const DirectF32<NBLK>& d,
                                           int ks, int lh, int klen)
{
    if constexpr (D)
        direct_f32_load<NBLK>(st.v, d, ks, lh, klen);
    else
#pragma unroll
        for (int i = 0; i < NBLK; i++)
#pragma unroll
            for (int t = 0; t < 3; t++) st.f[i][t] = __builtin_bit_cast(bf16x8, p[i * rs + ks * 192 + t * 64]);
}

template <int NBLK, bool D>
__device__ __forceinline__ void stage_frags(bf16x8 (&x)[NBLK][3], const Stage<NBLK, D>& st)
{
    if constexpr (D)
        split_frags<NBLK>(x, st.v);
    else
#pragma unroll
        for (int i = 0; i < NBLK; i++)
#pragma unroll
            for (int t = 0; t < 3; t++) x[i][t] = st.f[i][t];
}

// One C tile.  KS > 1 splits k inside the workgroup: KS sets of WGM x WGN
// waves take interleaved pairs of k groups (pair p goes to set p % KS) and
// their sums meet in LDS before set 0 runs the epilogue; every wave reaches
// that barrier.
template <int BM, int BN, int WGM, int WGN, bool CHUNKED, bool DA, bool DB, int KS>
__device__ __forceinline__ void bf16_task(const PanelTask& t, const int64_t* __restrict__ pool,
                                          const __bf16* __restrict__ ws, double* __restrict__ cws,
                                          const float* A, const float* B, float* __restrict__ C, double alpha,
                                          double beta, int use_k)
{
    constexpr int WM = BM / WGM, WN = BN / WGN, MB = WM / 32, NB = WN / 32;
    constexpr int CHUNK = 16;  // k16 steps per accumulation chunk (256 k)
    const int lane = threadIdx.x & 63;
    const int wave_all = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int kw = wave_all / (WGM * WGN), wave = wave_all % (WGM * WGN);
    const int wm = wave % WGM, wn = wave / WGM;
    const bool active = !(wm * WM >= t.mlen || wn * WN >= t.nlen);
    if (KS == 1 && !active) return;  // no barriers below when KS == 1

    f32x16 acc[MB][NB], tot[CHUNKED ? MB : 1][CHUNKED ? NB : 1];
#pragma unroll
    for (int i = 0; i < MB; i++)
#pragma unroll
        for (int jb = 0; jb < NB; jb++)
            for (int r = 0; r < 16; r++) acc[i][jb][r] = 0.0f;
    if (CHUNKED)
#pragma unroll
        for (int i = 0; i < (CHUNKED ? MB : 1); i++)
#pragma unroll
            for (int jb = 0; jb < (CHUNKED ? NB : 1); jb++)
                for (int r = 0; r < 16; r++) tot[i][jb][r] = 0.0f;

    // this set's pairs of k groups: kw, kw + KS, ... (t.nkg is even)
    const int np = use_k && active ? (t.nkg / 2 - kw + KS - 1) / KS : 0;
    if (np > 0)
    {
        const int lr = lane & 31, lh = lane >> 5;
        const int64_t rs = (int64_t)t.nkg * 192;  // uint4 per row block
        const uint4* pa = (const uint4*)(ws + t.a) + (int64_t)wm * MB * rs + lane;
        const uint4* pb = (const uint4*)(ws + t.b) + (int64_t)wn * NB * rs + lane;
        DirectF32<MB> da;
        DirectF32<NB> db;
        if constexpr (DA)
            direct_f32_init<MB>(da, A, t.a_src, t.a_bytes, pool, t.a_rt, t.a_kt, t.a_km, wm * WM, t.mlen, lr);
        if constexpr (DB)
            direct_f32_init<NB>(db, B, t.b_src, t.b_bytes, pool, t.b_rt, t.b_kt, t.b_km, wn * WN, t.nlen, lr);
        Stage<MB, DA> a0, a1;
        Stage<NB, DB> b0, b1;
        stage_load<MB, DA>(a0, pa, rs, da, 2 * kw, lh, t.klen);
        stage_load<NB, DB>(b0, pb, rs, db, 2 * kw, lh, t.klen);
        stage_load<MB, DA>(a1, pa, rs, da, 2 * kw + 1, lh, t.klen);
        stage_load<NB, DB>(b1, pb, rs, db, 2 * kw + 1, lh, t.klen);
        for (int q = 0; q < np; q++)
        {
            // the last iteration re-reads its own groups (in bounds, unused)
            const int nx = 2 * (kw + (q + 1 < np ? q + 1 : q) * KS);
            bf16x8 fa[MB][3], fb[NB][3];
            stage_frags<MB, DA>(fa, a0);
            stage_frags<NB, DB>(fb, b0);
            mma_frags<MB, NB>(acc, fa, fb);
            stage_load<MB, DA>(a0, pa, rs, da, nx, lh, t.klen);
            stage_load<NB, DB>(b0, pb, rs, db, nx, lh, t.klen);
            stage_frags<MB, DA>(fa, a1);
            stage_frags<NB, DB>(fb, b1);
            mma_frags<MB, NB>(acc, fa, fb);
            stage_load<MB, DA>(a1, pa, rs, da, nx + 1, lh, t.klen);
            stage_load<NB, DB>(b1, pb, rs, db, nx + 1, lh, t.klen);
            if (CHUNKED && ((2 * q + 2) % CHUNK == 0 || q + 1 >= np))
            {
#pragma unroll
                for (int i = 0; i < MB; i++)
#pragma unroll
                    for (int jb = 0; jb < NB; jb++)
                    {
                        tot[i % (CHUNKED ? MB : 1)][jb % (CHUNKED ? NB : 1)] += acc[i][jb];
                        for (int r = 0; r < 16; r++) acc[i][jb][r] = 0.0f;
                    }
            }
        }
    }
    if constexpr (KS > 1)
    {
        // the other sets' sums (the chunk totals when CHUNKED) through LDS
        __shared__ f32x16 red[KS - 1][WGM * WGN][MB * NB][64];
        if (kw > 0 && active)
#pragma unroll
            for (int i = 0; i < MB; i++)
#pragma unroll
                for (int jb = 0; jb < NB; jb++)
                    red[kw - 1][wave][i * NB + jb][lane] =
                        CHUNKED ? tot[i % (CHUNKED ? MB : 1)][jb % (CHUNKED ? NB : 1)] : acc[i][jb];
        __syncthreads();
        if (kw > 0 || !active) return;
#pragma unroll
        for (int s = 0; s < KS - 1; s++)
#pragma unroll
            for (int i = 0; i < MB; i++)
#pragma unroll
                for (int jb = 0; jb < NB; jb++)
                {
                    if constexpr (CHUNKED)
                        tot[i % (CHUNKED ? MB : 1)][jb % (CHUNKED ? NB : 1)] += red[s][wave][i * NB + jb][lane];
                    else
                        acc[i][jb] += red[s][wave][i * NB + jb][lane];
                }
    }

    // D register r of lane l holds D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32]
    const int cl = lane & 31, rh = lane >> 5;
    int64_t coff[NB], cs[NB];
#pragma unroll
    for (int jb = 0; jb < NB; jb++)
    {
        const int col = wn * WN + jb * 32 + cl;
        coff[jb] = 0;
        cs[jb] = 0;
        if (t.c_ld <= 0 && col < t.nlen)
        {
            const int64_t v = pool[t.c_colmap + col];
            if (t.c_ld < 0)
            {
                cs[jb] = v >> 32;  // column patch; index in coff
                coff[jb] = v & 0xffffffff;
            }
            else
            {
                coff[jb] = v;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < MB; i++)
#pragma unroll
        for (int r = 0; r < 16; r++)
        {
            const int row = wm * WM + i * 32 + 8 * (r / 4) + 4 * rh + r % 4;
            if (row >= t.mlen) continue;
            int64_t roff = 0, rp = 0;
            if (t.c_ld == 0)
                roff = t.c_src + pool[t.c_rowmap + row];
            else if (t.c_ld < 0)
            {
                const int64_t code = pool[t.c_rowmap + row];
                rp = code >> 32;
                roff = code & 0xffffffff;  // index in the row patch
            }
#pragma unroll
            for (int jb = 0; jb < NB; jb++)
            {
                const int col = wn * WN + jb * 32 + cl;
                if (col >= t.nlen) continue;
                const float s = CHUNKED ? tot[i % (CHUNKED ? MB : 1)][jb % (CHUNKED ? NB : 1)][r]
                                        : acc[i][jb][r];
                if (t.c_ld > 0)
                {
                    cws[t.c_src + (int64_t)row * t.c_ld + col] = (double)s;
                    continue;
                }
                float* p;
                if (t.c_ld < 0)
                {
                    const int64_t* blk = pool + t.c_src + 3 * (rp * (-t.c_ld) + cs[jb]);
                    p = C + blk[0] + pool[blk[1] + roff] + pool[blk[2] + coff[jb]];
                }
                else
                {
                    p = C + roff + coff[jb];
                }
                double v = alpha * (double)s;
                if (beta != 0.0) v += beta * (double)*p;
                *p = (float)v;
            }
        }
}

// Persistent over the task list, as the FP64 panel GEMM.
template <int BM, int BN, int WGM, int WGN, bool CHUNKED, bool DA, bool DB, int KS>
__global__ __launch_bounds__(64 * WGM * WGN * KS) void gemm_bf16_panel_kernel(
    const PanelTask* __restrict__ tasks, int ntasks, const int64_t* __restrict__ pool,
    const __bf16* __restrict__ ws, double* __restrict__ cws, const float* A, const float* B,
    float* __restrict__ C, double alpha, double beta, int use_k)
{
    for (int ti = blockIdx.x; ti < ntasks; ti += gridDim.x)
    {
        const PanelTask t = tasks[ti];
        bf16_task<BM, BN, WGM, WGN, CHUNKED, DA, DB, KS>(t, pool, ws, cws, A, B, C, alpha, beta, use_k);
        // the next task's partial sums reuse the LDS that set 0 just read
        if constexpr (KS > 1) __syncthreads();
    }
}

// ------------------------------------------------------------------ table

struct Bf16Tile
{
    int bm, bn, wgm, wgn;
    bool fast_only;  // the chunked accumulator does not fit the register file
    bool direct;     // variants that read FP32 operands from the tensors
    int ks;          // k sets inside the workgroup
};

// id, BM, BN, waves (WGM x WGN), fast only, direct variants.  Registers
// per wave (gfx950): 128x64 and 64x128 wave tiles 228 (single accumulator)
// / 410 (chunked), 64x64 174 / 230.  (A 128x128 wave tile needs all 512
// registers for the single accumulator and spills; it is not offered.)
// Direct operands hold two 16-k steps of raw FP32 per block and split them
// in registers, so they are offered on the 64x64 and 32x32 wave tiles.
// Tiles 4 and 5 are tile 3 with k split between two and four sets of waves
// (short k: DPD irrep GEMMs), reading packed panels only.
#define TBLIS_PANEL_BF16_TILES(X)        \
    X(0, 256, 128, 2, 2, false, false, 1) \
    X(1, 128, 256, 2, 2, false, false, 1) \
    X(2, 128, 128, 2, 2, false, true, 1)  \
    X(3, 64, 64, 2, 2, false, true, 1)    \
    X(4, 64, 64, 2, 2, false, false, 2)   \
    X(5, 64, 64, 2, 2, false, false, 4)

constexpr Bf16Tile kBTiles[] = {
#define X(id, bm, bn, wgm, wgn, fo, dir, ks) {bm, bn, wgm, wgn, fo, dir, ks},
    TBLIS_PANEL_BF16_TILES(X)
#undef X
};
constexpr int kNumBTiles = sizeof(kBTiles) / sizeof(kBTiles[0]);

template <int BM, int BN, int WGM, int WGN, bool CHUNKED, bool DA, bool DB, int KS>
hipError_t launch_bf16_t(const PanelLaunch& L)
{
    auto* k = &gemm_bf16_panel_kernel<BM, BN, WGM, WGN, CHUNKED, DA, DB, KS>;
    const int threads = 64 * WGM * WGN * KS;
    const dim3 grid(persistent_grid((const void*)k, threads, L.ntasks));
    hipLaunchKernelGGL(k, grid, dim3(threads), 0, L.stream, L.tasks, L.ntasks, L.pool, (const __bf16*)L.ws,
                       L.cws, (const float*)L.A, (const float*)L.B, (float*)L.C, L.alpha, L.beta, L.use_k);
    return hipGetLastError();
}

template <int BM, int BN, int WGM, int WGN, bool CHUNKED, bool DIR, int KS>
hipError_t launch_bf16_d(const PanelLaunch& L)
{
    if constexpr (DIR)
    {
        if (L.direct_a && L.direct_b) return launch_bf16_t<BM, BN, WGM, WGN, CHUNKED, true, true, KS>(L);
        if (L.direct_a) return launch_bf16_t<BM, BN, WGM, WGN, CHUNKED, true, false, KS>(L);
        if (L.direct_b) return launch_bf16_t<BM, BN, WGM, WGN, CHUNKED, false, true, KS>(L);
    }
    else if (L.direct_a || L.direct_b)
        return hipErrorInvalidValue;
    return launch_bf16_t<BM, BN, WGM, WGN, CHUNKED, false, false, KS>(L);
}

}  // namespace

int panel_bf16_num_tiles() { return kNumBTiles; }
int panel_bf16_tile_m(int t) { return kBTiles[t].bm; }
int panel_bf16_tile_n(int t) { return kBTiles[t].bn; }
int panel_bf16_tile_waves(int t) { return kBTiles[t].wgm * kBTiles[t].wgn * kBTiles[t].ks; }
bool panel_bf16_tile_fast_only(int t) { return kBTiles[t].fast_only; }
bool panel_bf16_tile_direct(int t) { return t >= 0 && t < kNumBTiles && kBTiles[t].direct; }

hipError_t launch_panel_pack_bf16(const PanelLaunch& L)
{
    if (L.nchunks == 0) return hipSuccess;
    if (L.dtype != 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(pack_bf16_panel_kernel, dim3(L.nchunks), dim3(256), 0, L.stream, L.chunks, L.pool,
                       (const float*)L.A, (const float*)L.B, (__bf16*)L.ws);
    return hipGetLastError();
}

hipError_t launch_panel_gemm_bf16(const PanelLaunch& L)
{
    if (L.ntasks == 0) return hipSuccess;
    if (L.tile < 0 || L.tile >= kNumBTiles || L.dtype != 1) return hipErrorInvalidValue;
    const bool chunked = L.family == FAMILY_BF16X3;
    if (chunked && kBTiles[L.tile].fast_only) return hipErrorInvalidValue;
    switch (L.tile * 2 + (chunked ? 1 : 0))
    {
#define X(id, bm, bn, wgm, wgn, fo, dir, ks)                       \
    case 2 * id:                                                   \
        return launch_bf16_d<bm, bn, wgm, wgn, false, dir, ks>(L); \
    case 2 * id + 1:                                               \
        if constexpr (!fo) return launch_bf16_d<bm, bn, wgm, wgn, true, dir, ks>(L); \
        return hipErrorInvalidValue;
        TBLIS_PANEL_BF16_TILES(X)
#undef X
        default:
            return hipErrorInvalidValue;
    }
}

// The v_mfma_f32_32x32x16_bf16 lane layout the GEMM relies on: lane l takes
// A[l%32][8*(l/32)+j] and B[8*(l/32)+j][l%32] (j = 0..7) from row-major A
// (32x16) and B (16x32), and register r of lane l,
// D[8(r/4) + 4(l/32) + r%4][l%32], is stored row-major into D (32x32).
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

hipError_t launch_bf16_probe(const float* A, const float* B, float* D, hipStream_t s)
{
    hipLaunchKernelGGL(bf16_probe_kernel, dim3(1), dim3(64), 0, s, A, B, D);
    return hipGetLastError();
}

}  // namespace tblis_amd
