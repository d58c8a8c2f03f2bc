// FP64 MFMA over fragment-ordered panels (gfx950).
//
// Three kernels per contraction:
//   pack    gathers every operand block through its row and k scatter
//           vectors (the reference's block-scatter packing,
//           src/matrix/block_scatter_matrix.hpp:353-431 and
//           src/kernels/3m/packm.hpp:293-404) into panels whose 1 KiB
//           fragment groups are the exact lane order of v_mfma_f64_16x16x4,
//           zero-filling padding;
//   gemm    one workgroup per C tile; each wave streams its row blocks of
//           A' and B' with one coalesced 16-byte load per lane and per
//           block, straight into registers (no LDS, no barriers), two
//           k groups in flight, and accumulates with v_mfma_f64_16x16x4;
//           the epilogue applies alpha/beta through C's scatter vectors
//           (beta == 0 does not read C, src/kernels/3m/gemm.hpp:54-58), or
//           stores a dense C panel when the GEMM spans several C blocks;
//   unpack  (DPD) scatters C panels into the C blocks with alpha/beta.
#include <algorithm>
#include <cstdlib>
#include <map>
#include <mutex>

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace tblis_amd
{

namespace
{

typedef double v4d __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------ pack

// Every address of a chunk follows from its PackChunk record (job fields
// and the chunk's group tables), so a wave goes from the record straight
// to its data loads.
template <typename T>
__global__ __launch_bounds__(256) void pack_panel_kernel(const PackChunk* __restrict__ chunks,
                                                         const int64_t* __restrict__ pool,
                                                         const T* __restrict__ A,
                                                         const T* __restrict__ B,
                                                         double* __restrict__ ws)
{
    const PackChunk& j = chunks[xcd_contiguous_chunk(blockIdx.x, gridDim.x)];
    const T* __restrict__ X = j.operand ? B : A;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lr = lane & 15, lk = lane >> 4;
    const bool zero = j.src == PANEL_MISSING;
    const int ng = j.nrb * j.nkg;
    if (j.kfast && !zero)
    {
        // k-fast source: per pair of k groups, four loads of four rows x 16
        // consecutive k (whole 128-byte lines), transposed through a per-wave
        // LDS tile into the fragment order.  The k tables are unit-stride.
        // The loads of kBatch pairs are issued before the first transpose,
        // so a wave waits on memory once per batch rather than per pair.
        constexpr int kBatch = 4;
        __shared__ double tile[4][16][17];
        const int npair = (j.nkg + 1) / 2;
        const int niter = j.nrb * npair;
        const int kk = lane & 15;
        for (int g0 = wave; g0 < niter; g0 += 4 * kBatch)
        {
            double v[kBatch][4];
#pragma unroll
            for (int u = 0; u < kBatch; u++)
            {
                const int g = g0 + 4 * u;
#pragma unroll
                for (int jj = 0; jj < 4; jj++) v[u][jj] = 0.0;
                if (g >= niter) continue;
                const int ri = g / npair, ki = 2 * (g % npair);  // groups within the chunk
                const int rb = j.rb0 + ri;
                const int kg = j.kg0 + ki;
                const bool two = ki + 1 < j.nkg;
                const int64_t rbase = j.rt[ri][0];
                const int64_t rstr = j.rt[ri][1];
                const int h = kk >> 3;
                const int64_t kb = j.kt[ki + (two ? h : 0)][0];
                const int pk = kg * 8 + kk - j.k0;  // job-relative k
                const bool kin = pk >= 0 && pk < j.nk && (h == 0 || two);
                const int64_t koff = kb + (kk & 7);
#pragma unroll
                for (int jj = 0; jj < 4; jj++)
                {
                    const int r = (lane >> 4) + 4 * jj;
                    const int prow = rb * 16 + r - j.r0;
                    if (kin && prow >= 0 && prow < j.nr)
                    {
                        const int64_t roff = rstr != SCATTERED ? rbase + r * rstr : pool[j.rowmap + prow];
                        v[u][jj] = j.scale * (double)X[j.src + roff + koff];
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < kBatch; u++)
            {
                const int g = g0 + 4 * u;
                if (g >= niter) break;
                const int rb = j.rb0 + g / npair;
                const int kg = j.kg0 + 2 * (g % npair);
                const bool two = kg + 1 < j.kg0 + j.nkg;
#pragma unroll
                for (int jj = 0; jj < 4; jj++) tile[wave][(lane >> 4) + 4 * jj][kk] = v[u][jj];
                __builtin_amdgcn_wave_barrier();
                const int row = rb * 16 + lr - j.r0;
                const bool rin = row >= 0 && row < j.nr;
                for (int hh = 0; hh < (two ? 2 : 1); hh++)
                {
                    const int k0 = (kg + hh) * 8 + lk - j.k0, k1 = k0 + 4;
                    const bool in0 = rin && k0 >= 0 && k0 < j.nk;
                    const bool in1 = rin && k1 >= 0 && k1 < j.nk;
                    const double v0 = tile[wave][lr][8 * hh + lk], v1 = tile[wave][lr][8 * hh + 4 + lk];
                    double* out = ws + j.out + ((int64_t)rb * j.pnkg + kg + hh) * 128 + lane * 2;
                    if (in0 && in1)
                    {
                        double2 w;
                        w.x = v0;
                        w.y = v1;
                        *(double2*)out = w;
                    }
                    else
                    {
                        if (in0) out[0] = v0;
                        if (in1) out[1] = v1;
                    }
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
        return;
    }
    // general source: one fragment group per iteration, the loads of
    // kBatch groups issued before their stores
    constexpr int kBatch = 8;
    for (int g0 = wave; g0 < ng; g0 += 4 * kBatch)
    {
        double v0[kBatch], v1[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; u++)
        {
            const int g = g0 + 4 * u;
            v0[u] = 0.0;
            v1[u] = 0.0;
            if (g >= ng || zero) continue;
            const int ri = g / j.nkg, ki = g % j.nkg;  // groups within the chunk
            const int rb = j.rb0 + ri;
            const int kg = j.kg0 + ki;
            const int row = rb * 16 + lr - j.r0;
            const bool rin = row >= 0 && row < j.nr;
            const int k0 = kg * 8 + lk - j.k0, k1 = k0 + 4;
            const bool in0 = rin && k0 >= 0 && k0 < j.nk;
            const bool in1 = rin && k1 >= 0 && k1 < j.nk;
            // group tables: wave-uniform (base, stride) of this row / k group
            const int64_t rbase = j.rt[ri][0];
            const int64_t rstr = j.rt[ri][1];
            const int64_t kbase = j.kt[ki][0];
            const int64_t kstr = j.kt[ki][1];
            if (rin)
            {
                const int64_t roff = j.src + (rstr != SCATTERED ? rbase + lr * rstr : pool[j.rowmap + row]);
                if (in0)
                    v0[u] = j.scale * (double)X[roff + (kstr != SCATTERED ? kbase + lk * kstr : pool[j.kmap + k0])];
                if (in1)
                    v1[u] = j.scale *
                            (double)X[roff + (kstr != SCATTERED ? kbase + (lk + 4) * kstr : pool[j.kmap + k1])];
            }
        }
#pragma unroll
        for (int u = 0; u < kBatch; u++)
        {
            const int g = g0 + 4 * u;
            if (g >= ng) break;
            const int rb = j.rb0 + g / j.nkg;
            const int kg = j.kg0 + g % j.nkg;
            const int row = rb * 16 + lr - j.r0;
            const bool rin = row >= 0 && row < j.nr;
            const int k0 = kg * 8 + lk - j.k0, k1 = k0 + 4;
            const bool in0 = rin && k0 >= 0 && k0 < j.nk;
            const bool in1 = rin && k1 >= 0 && k1 < j.nk;
            double* out = ws + j.out + ((int64_t)rb * j.pnkg + kg) * 128 + lane * 2;
            if (in0 && in1)
            {
                double2 w;
                w.x = v0[u];
                w.y = v1[u];
                *(double2*)out = w;
            }
            else
            {
                if (in0) out[0] = v0[u];
                if (in1) out[1] = v1[u];
            }
        }
    }
}

// ------------------------------------------------------------------ gemm

template <int MB, int NB>
__device__ __forceinline__ void load_groups(double2 (&a)[MB], double2 (&b)[NB], const double2* pa,
                                            const double2* pb, int64_t rs, int kg)
{
#pragma unroll
    for (int i = 0; i < MB; i++) a[i] = pa[i * rs + kg * 64];
#pragma unroll
    for (int j = 0; j < NB; j++) b[j] = pb[j * rs + kg * 64];
}

template <int MB, int NB>
__device__ __forceinline__ void mma_group(v4d (&acc)[MB][NB], const double2 (&a)[MB],
                                          const double2 (&b)[NB])
{
#pragma unroll
    for (int i = 0; i < MB; i++)
#pragma unroll
        for (int j = 0; j < NB; j++)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].x, b[j].x, acc[i][j], 0, 0, 0);
#pragma unroll
    for (int i = 0; i < MB; i++)
#pragma unroll
        for (int j = 0; j < NB; j++)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].y, b[j].y, acc[i][j], 0, 0, 0);
}

// One operand read straight from its tensor (PanelTask "direct"): the
// lane's row offsets of its row blocks (taken once per task), and the k
// tables: (base, stride) per 8-k group, per-element offsets for groups
// that are not affine.
template <typename TC, int NBLK>
struct DirectSide
{
    __amdgpu_buffer_rsrc_t rsrc;
    const int64_t* kt;
    const int64_t* km;
    uint32_t roff[NBLK];
    bool rvalid[NBLK];
};

template <typename TC>
__device__ __forceinline__ double direct_load(__amdgpu_buffer_rsrc_t r, uint32_t byte_off)
{
    if constexpr (sizeof(TC) == 8)
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, byte_off, 0, 0));
    else
        return (double)__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, byte_off, 0, 0));
}

// rows: pool index of the offset of the task's first row; this wave's
// rows start at rows0 (of len valid rows).
template <typename TC, int NBLK>
__device__ __forceinline__ void direct_init(DirectSide<TC, NBLK>& d, const void* X, int64_t src,
                                            uint32_t bytes, const int64_t* __restrict__ pool,
                                            int64_t rows, int64_t kt, int64_t km, int rows0, int len,
                                            int lr)
{
    d.rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)((const TC*)X + src), (short)0, (int)bytes,
                                               0x00020000);
    d.kt = pool + kt;
    d.km = pool + km;
#pragma unroll
    for (int i = 0; i < NBLK; i++)
    {
        const int r = rows0 + i * 16 + lr;
        d.rvalid[i] = r < len;
        d.roff[i] = d.rvalid[i] ? (uint32_t)pool[rows + r] : 0u;
    }
}

// Two k steps of NBLK row blocks of a direct operand as the double2 a
// panel would hold: rows lr, k = 8 kg + lk and 8 kg + 4 + lk.
template <typename TC, int NBLK>
__device__ __forceinline__ void direct_groups(double2 (&x)[NBLK], const DirectSide<TC, NBLK>& d,
                                              int kg, int lk, int klen)
{
    const int64_t kb = d.kt[2 * kg], ks = d.kt[2 * kg + 1];
    const int k0 = kg * 8 + lk;
    uint32_t o0, o1;
    if (ks != SCATTERED)  // wave-uniform
    {
        o0 = (uint32_t)(kb + lk * ks);
        o1 = (uint32_t)(kb + (lk + 4) * ks);
    }
    else
    {
        o0 = k0 < klen ? (uint32_t)d.km[k0] : 0u;
        o1 = k0 + 4 < klen ? (uint32_t)d.km[k0 + 4] : 0u;
    }
#pragma unroll
    for (int i = 0; i < NBLK; i++)
    {
        const uint32_t b0 = d.rvalid[i] && k0 < klen ? (d.roff[i] + o0) * (uint32_t)sizeof(TC) : 0xffffffffu;
        const uint32_t b1 =
            d.rvalid[i] && k0 + 4 < klen ? (d.roff[i] + o1) * (uint32_t)sizeof(TC) : 0xffffffffu;
        x[i].x = direct_load<TC>(d.rsrc, b0);
        x[i].y = direct_load<TC>(d.rsrc, b1);
    }
}

// One C tile.  KS > 1 splits k inside the workgroup: KS sets of WGM x WGN
// waves take interleaved k groups (group kg goes to set kg % KS) and the
// partial sums meet in LDS before set 0 runs the epilogue.  For short k
// (DPD irrep GEMMs, small batch instances) this doubles the waves resident
// per SIMD.  Every wave reaches the LDS barrier when KS > 1.
template <int BM, int BN, int WGM, int WGN, typename TC, bool DA, bool DB, int KS>
__device__ __forceinline__ void panel_task(const PanelTask& t, const int64_t* __restrict__ pool,
                                           const double* __restrict__ ws, double* __restrict__ cws,
                                           const void* A, const void* B, TC* __restrict__ C, double alpha,
                                           double beta, int use_k)
{
    constexpr int WM = BM / WGM, WN = BN / WGN, MB = WM / 16, NB = WN / 16;
    const int lane = threadIdx.x & 63;
    const int wave_all = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int kw = wave_all / (WGM * WGN), wave = wave_all % (WGM * WGN);
    const int wm = wave % WGM, wn = wave / WGM;
    const bool active = !(wm * WM >= t.mlen || wn * WN >= t.nlen);
    if (KS == 1 && !active) return;  // no barriers below when KS == 1

    v4d acc[MB][NB];
#pragma unroll
    for (int i = 0; i < MB; i++)
#pragma unroll
        for (int j = 0; j < NB; j++) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

    // this wave's k groups: kw, kw + KS, ... (nk is even)
    const int nk = use_k && active ? (t.nkg - kw + KS - 1) / KS : 0;
    if (nk > 0)
    {
        const int lr = lane & 15, lk = lane >> 4;
        const int64_t rs = (int64_t)t.nkg * 64;  // double2 per row block
        const double2* pa = (const double2*)(ws + t.a) + (int64_t)wm * MB * rs + lane;
        const double2* pb = (const double2*)(ws + t.b) + (int64_t)wn * NB * rs + lane;
        DirectSide<TC, MB> da;
        DirectSide<TC, NB> db;
        if constexpr (DA)
            direct_init(da, A, t.a_src, t.a_bytes, pool, t.a_rt, t.a_kt, t.a_km, wm * WM, t.mlen, lr);
        if constexpr (DB)
            direct_init(db, B, t.b_src, t.b_bytes, pool, t.b_rt, t.b_kt, t.b_km, wn * WN, t.nlen, lr);
        auto load = [&](double2 (&a)[MB], double2 (&b)[NB], int li) {
            const int kg = kw + li * KS;
            if constexpr (DA)
                direct_groups<TC, MB>(a, da, kg, lk, t.klen);
            else
#pragma unroll
                for (int i = 0; i < MB; i++) a[i] = pa[i * rs + kg * 64];
            if constexpr (DB)
                direct_groups<TC, NB>(b, db, kg, lk, t.klen);
            else
#pragma unroll
                for (int j = 0; j < NB; j++) b[j] = pb[j * rs + kg * 64];
        };
        // k groups in flight: five for one-wave 32x32 tiles with one or two
        // k sets (their MFMA work per group is short; 160 registers, three
        // waves per SIMD; three in flight ties on the permuted workload),
        // three otherwise: 32x32 waves in four-wave tiles or with k split
        // four ways then hold 128 registers and four waves per SIMD, which
        // gain more from resident waves (permuted workload on the 2x2-wave
        // 64x64 tile 49.4-49.9 -> 53.3-53.7 TFLOP/s, 8-irrep DPD on the
        // four-set tile 0.0630 -> 0.0609 ms per call), and 64x64 waves hold
        // as many as fit next to their accumulators at two waves per SIMD
        // (1-18% faster than two)
        constexpr int PF = MB * NB <= 4 && WGM * WGN == 1 && KS < 4 ? 5 : 3;
        double2 a[PF][MB], b[PF][NB];
#pragma unroll
        for (int d = 0; d < PF; d++) load(a[d], b[d], d < nk ? d : nk - 1);
        // whole passes of PF groups with no exit inside the pass, so that
        // the loop header is reached with the same loads in flight on every
        // path and each group waits only for its own loads (an exit inside
        // the pass made the compiler drain every load at the header)
        const int nmain = nk - nk % PF;
        for (int kg = 0; kg < nmain; kg += PF)
        {
#pragma unroll
            for (int d = 0; d < PF; d++)
            {
                mma_group<MB, NB>(acc, a[d], b[d]);
                // past the end: re-read the final group (in bounds, unused)
                const int nx = kg + d + PF < nk ? kg + d + PF : nk - 1;
                load(a[d], b[d], nx);
            }
        }
        // the last nk % PF groups sit in a[0 ..], b[0 ..]
#pragma unroll
        for (int d = 0; d < PF; d++)
            if (d < nk - nmain) mma_group<MB, NB>(acc, a[d], b[d]);
    }
    if constexpr (KS > 1)
    {
        // partial sums of the other k sets through LDS, then set 0 finishes
        __shared__ v4d red[KS - 1][WGM * WGN][MB * NB][64];
        if (kw > 0 && active)
#pragma unroll
            for (int i = 0; i < MB; i++)
#pragma unroll
                for (int j = 0; j < NB; j++) red[kw - 1][wave][i * NB + j][lane] = acc[i][j];
        __syncthreads();
        if (kw > 0 || !active) return;
#pragma unroll
        for (int q = 0; q < KS - 1; q++)
#pragma unroll
            for (int i = 0; i < MB; i++)
#pragma unroll
                for (int j = 0; j < NB; j++) acc[i][j] += red[q][wave][i * NB + j][lane];
    }

    // D register r of lane l holds D[l / 16 + 4 r][l % 16] of each block
    const int cl = lane & 15, rl = lane >> 4;
    if (t.c_ld < 0)
    {
        // C over several blocks: each element finds its block
        const int ns = -t.c_ld;
        int64_t cs[NB], cc[NB];
#pragma unroll
        for (int j = 0; j < NB; j++)
        {
            const int col = wn * WN + j * 16 + cl;
            const int64_t code = col < t.nlen ? pool[t.c_colmap + col] : 0;
            cs[j] = code >> 32;
            cc[j] = code & 0xffffffff;
        }
#pragma unroll
        for (int i = 0; i < MB; i++)
#pragma unroll
            for (int r = 0; r < 4; r++)
            {
                const int row = wm * WM + i * 16 + rl + 4 * r;
                if (row >= t.mlen) continue;
                const int64_t code = pool[t.c_rowmap + row];
                const int64_t rp = code >> 32, ri = code & 0xffffffff;
#pragma unroll
                for (int j = 0; j < NB; j++)
                {
                    const int col = wn * WN + j * 16 + cl;
                    if (col >= t.nlen) continue;
                    const int64_t* blk = pool + t.c_src + 3 * (rp * ns + cs[j]);
                    TC* p = C + blk[0] + pool[blk[1] + ri] + pool[blk[2] + cc[j]];
                    double v = alpha * acc[i][j][r];
                    if (beta != 0.0) v += beta * (double)*p;
                    *p = (TC)v;
                }
            }
        return;
    }
    if (t.c_ld)
    {
#pragma unroll
        for (int i = 0; i < MB; i++)
#pragma unroll
            for (int r = 0; r < 4; r++)
            {
                const int row = wm * WM + i * 16 + rl + 4 * r;
                if (row >= t.mlen) continue;
#pragma unroll
                for (int j = 0; j < NB; j++)
                {
                    const int col = wn * WN + j * 16 + cl;
                    if (col < t.nlen) cws[t.c_src + (int64_t)row * t.c_ld + col] = acc[i][j][r];
                }
            }
        return;
    }
    int64_t coff[NB];
#pragma unroll
    for (int j = 0; j < NB; j++)
    {
        const int col = wn * WN + j * 16 + cl;
        coff[j] = col < t.nlen ? pool[t.c_colmap + col] : 0;
    }
#pragma unroll
    for (int i = 0; i < MB; i++)
#pragma unroll
        for (int r = 0; r < 4; r++)
        {
            const int row = wm * WM + i * 16 + rl + 4 * r;
            if (row >= t.mlen) continue;
            const int64_t roff = t.c_src + pool[t.c_rowmap + row];
#pragma unroll
            for (int j = 0; j < NB; j++)
            {
                const int col = wn * WN + j * 16 + cl;
                if (col >= t.nlen) continue;
                TC* p = C + roff + coff[j];
                double v = alpha * acc[i][j][r];
                if (beta != 0.0) v += beta * (double)*p;
                *p = (TC)v;
            }
        }
}

// Persistent over the task list: workgroup b runs tasks b, b + G, b + 2G,
// ... (G = gridDim.x, a multiple of 8, so every task stays on the XCD the
// task order planned it for: task i on XCD i % 8).  With G = ntasks it is
// one task per workgroup.
template <int BM, int BN, int WGM, int WGN, typename TC, bool DA, bool DB, int KS = 1>
__global__ __launch_bounds__(64 * WGM * WGN * KS) void gemm_panel_kernel(
    const PanelTask* __restrict__ tasks, int ntasks, const int64_t* __restrict__ pool,
    const double* __restrict__ ws, double* __restrict__ cws, const void* A, const void* B,
    TC* __restrict__ C, double alpha, double beta, int use_k)
{
    for (int ti = blockIdx.x; ti < ntasks; ti += gridDim.x)
    {
        const PanelTask t = tasks[ti];
        panel_task<BM, BN, WGM, WGN, TC, DA, DB, KS>(t, pool, ws, cws, A, B, C, alpha, beta, use_k);
        // the next task's partial sums reuse the LDS that set 0 just read
        if constexpr (KS > 1) __syncthreads();
    }
}

// ------------------------------------------------------------------ unpack

template <typename TC>
__global__ __launch_bounds__(256) void unpack_panel_kernel(const UnpackJob* __restrict__ jobs,
                                                           const UnpackChunk* __restrict__ chunks,
                                                           const int64_t* __restrict__ pool,
                                                           const double* __restrict__ ws,
                                                           TC* __restrict__ C, double alpha,
                                                           double beta)
{
    const UnpackChunk ch = chunks[blockIdx.x];
    const UnpackJob j = jobs[ch.job];
    for (int e = threadIdx.x; e < 64 * 64; e += 256)
    {
        const int i = ch.r0 + (j.rows_fast ? (e & 63) : (e >> 6));
        const int c = ch.c0 + (j.rows_fast ? (e >> 6) : (e & 63));
        if (i >= j.nr || c >= j.nc) continue;
        TC* p = C + j.dst + pool[j.rowmap + i] + pool[j.colmap + c];
        double v = alpha * ws[j.panel + (int64_t)i * j.ld + c];
        if (beta != 0.0) v += beta * (double)*p;
        *p = (TC)v;
    }
}

// ------------------------------------------------------------------ fill

// Pseudo-random values in [-1, 1) (a 64-bit mix of the index) so that
// timing runs see data with the switching activity of real operands.
__global__ __launch_bounds__(256) void fill_random_kernel(void* p, int64_t n, int bf16)
{
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
    {
        uint64_t x = (uint64_t)i * 0x9E3779B97F4A7C15ULL;
        x ^= x >> 31;
        x *= 0xBF58476D1CE4E5B9ULL;
        x ^= x >> 29;
        const double v = (double)(x >> 11) * (2.0 / 9007199254740992.0) - 1.0;
        if (bf16)
            ((__bf16*)p)[i] = (__bf16)(float)v;
        else
            ((double*)p)[i] = v;
    }
}

// ------------------------------------------------------------------ table

struct PanelTile
{
    int bm, bn, wgm, wgn, ks;
};

// id, BM, BN, waves (WGM x WGN), k sets inside the workgroup
#define TBLIS_PANEL_TILES(X)  \
    X(0, 128, 128, 2, 2, 1)   \
    X(1, 64, 64, 2, 2, 1)     \
    X(2, 64, 64, 1, 1, 1)     \
    X(3, 128, 64, 2, 1, 1)    \
    X(4, 64, 128, 1, 2, 1)    \
    X(5, 32, 32, 1, 1, 1)     \
    X(6, 64, 64, 2, 2, 2)     \
    X(7, 32, 32, 1, 1, 2)     \
    X(8, 32, 32, 1, 1, 4)

constexpr PanelTile kTiles[] = {
#define X(id, bm, bn, wgm, wgn, ks) {bm, bn, wgm, wgn, ks},
    TBLIS_PANEL_TILES(X)
#undef X
};
constexpr int kNumTiles = sizeof(kTiles) / sizeof(kTiles[0]);

template <typename TC, bool DA, bool DB>
hipError_t launch_gemm_t(const PanelLaunch& L)
{
    const int threads = 64 * kTiles[L.tile].wgm * kTiles[L.tile].wgn * kTiles[L.tile].ks;
    const dim3 block(threads);
    switch (L.tile)
    {
#define X(id, bm, bn, wgm, wgn, ks)                                                                        \
    case id:                                                                                               \
    {                                                                                                      \
        auto* k = &gemm_panel_kernel<bm, bn, wgm, wgn, TC, DA, DB, ks>;                                   \
        const dim3 grid(persistent_grid((const void*)k, threads, L.ntasks));                              \
        hipLaunchKernelGGL(k, grid, block, 0, L.stream, L.tasks, L.ntasks, L.pool, (const double*)L.ws,   \
                           L.cws, L.A, L.B, (TC*)L.C, L.alpha, L.beta, L.use_k);                          \
        break;                                                                                             \
    }
        TBLIS_PANEL_TILES(X)
#undef X
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename TC>
hipError_t launch_gemm_d(const PanelLaunch& L)
{
    if (L.direct_a && L.direct_b) return launch_gemm_t<TC, true, true>(L);
    if (L.direct_a) return launch_gemm_t<TC, true, false>(L);
    if (L.direct_b) return launch_gemm_t<TC, false, true>(L);
    return launch_gemm_t<TC, false, false>(L);
}

// The v_mfma_f64_16x16x4 lane layout the GEMM relies on (the lane-layout
// test): lane l supplies A[l%16][l/16] and B[l/16][l%16] of
// A[i][k] = 1000 i + k + 1 and B[k][j] = (k + 1)(j + 1), and register r of
// lane l, D[l/16 + 4r][l%16], is stored row-major into out (16 x 16).
__global__ void mfma_probe_kernel(double* out)
{
    const int lane = threadIdx.x;
    const int i = lane % 16, kq = lane / 16;
    const double a = 1000.0 * i + kq + 1;
    const double b = (kq + 1.0) * (i + 1.0);
    v4d d = {0.0, 0.0, 0.0, 0.0};
    d = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, d, 0, 0, 0);
    for (int r = 0; r < 4; r++) out[(lane / 16 + 4 * r) * 16 + lane % 16] = d[r];
}

}  // namespace

hipError_t launch_mfma_probe(double* out, hipStream_t s)
{
    hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, s, out);
    return hipGetLastError();
}

// Workgroups of a GEMM launch.  By default one per task: the hardware
// dispatcher then hands tasks to CUs as they free up, which balances the
// uneven tasks of batches (4096-instance batch, FP64: 38.1 TFLOP/s against
// 31.8 with a static stride over resident workgroups).  TBLIS_HIP_PERSISTENT=1
// launches as many workgroups as are resident together (a multiple of the
// 8 XCDs), each striding over the task list.
int persistent_grid(const void* kernel, int block, int ntasks)
{
    static const bool on = [] {
        const char* e = std::getenv("TBLIS_HIP_PERSISTENT");
        return e && e[0] == '1';
    }();
    if (!on) return ntasks;
    static std::mutex mu;
    static std::map<const void*, int> cap;
    int slots = 0;
    {
        std::lock_guard<std::mutex> g(mu);
        auto it = cap.find(kernel);
        if (it != cap.end())
            slots = it->second;
        else
        {
            int per_cu = 0, dev = 0, cus = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess ||
                hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            {
                (void)hipGetLastError();
                per_cu = 0;
            }
            slots = per_cu > 0 && cus > 0 ? per_cu * cus / 8 * 8 : 0;
            cap[kernel] = slots;
        }
    }
    return slots >= 8 && ntasks > slots ? slots : ntasks;
}

// The Ozaki family has one tile: 64 x 64 per workgroup, 2 x 2 waves.  The
// FP64 family's tiles are the register-fed ones, then the LDS-staged ones
// (gemm_lds.hip, four waves each).
int panel_num_tiles(int family)
{
    if (family == FAMILY_OZAKI) return 1;
    return family == FAMILY_F64 ? kNumTiles + lds_num_tiles() : panel_bf16_num_tiles();
}
bool panel_tile_lds(int family, int t) { return family == FAMILY_F64 && t >= kNumTiles; }
int panel_tile_m(int family, int t)
{
    if (family == FAMILY_OZAKI) return 64;
    if (panel_tile_lds(family, t)) return lds_tile_m(t - kNumTiles);
    return family == FAMILY_F64 ? kTiles[t].bm : panel_bf16_tile_m(t);
}
int panel_tile_n(int family, int t)
{
    if (family == FAMILY_OZAKI) return 64;
    if (panel_tile_lds(family, t)) return lds_tile_n(t - kNumTiles);
    return family == FAMILY_F64 ? kTiles[t].bn : panel_bf16_tile_n(t);
}
int panel_tile_waves(int family, int t)
{
    if (family == FAMILY_OZAKI) return 4;
    if (panel_tile_lds(family, t)) return 4;
    return family == FAMILY_F64 ? kTiles[t].wgm * kTiles[t].wgn * kTiles[t].ks : panel_bf16_tile_waves(t);
}

hipError_t launch_panel_pack(const PanelLaunch& L)
{
    if (L.nchunks == 0) return hipSuccess;
    if (L.family != FAMILY_F64) return launch_panel_pack_bf16(L);
    if (L.dtype == 0)
        hipLaunchKernelGGL(pack_panel_kernel<double>, dim3(L.nchunks), dim3(256), 0, L.stream, L.chunks,
                           L.pool, (const double*)L.A, (const double*)L.B, (double*)L.ws);
    else
        hipLaunchKernelGGL(pack_panel_kernel<float>, dim3(L.nchunks), dim3(256), 0, L.stream, L.chunks,
                           L.pool, (const float*)L.A, (const float*)L.B, (double*)L.ws);
    return hipGetLastError();
}

hipError_t launch_panel_gemm(const PanelLaunch& L)
{
    if (L.ntasks == 0) return hipSuccess;
    if (L.family != FAMILY_F64) return launch_panel_gemm_bf16(L);
    if (L.tile >= kNumTiles) return launch_gemm_lds(L, L.tile - kNumTiles);
    if (L.tile < 0) return hipErrorInvalidValue;
    return L.dtype == 0 ? launch_gemm_d<double>(L) : launch_gemm_d<float>(L);
}

hipError_t launch_fill_random(void* p, int64_t n, bool bf16, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_random_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n, bf16 ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_panel_unpack(const PanelLaunch& L)
{
    if (L.nuchunks == 0) return hipSuccess;
    if (L.dtype == 0)
        hipLaunchKernelGGL(unpack_panel_kernel<double>, dim3(L.nuchunks), dim3(256), 0, L.stream,
                           L.ujobs, L.uchunks, L.pool, (const double*)L.cws, (double*)L.C, L.alpha,
                           L.beta);
    else
        hipLaunchKernelGGL(unpack_panel_kernel<float>, dim3(L.nuchunks), dim3(256), 0, L.stream,
                           L.ujobs, L.uchunks, L.pool, (const double*)L.cws, (float*)L.C, L.alpha,
                           L.beta);
    return hipGetLastError();
}

}  // namespace tblis_amd
