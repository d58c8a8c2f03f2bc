// FP64 MFMA GEMM that reads both operands straight from the tensors through
// LDS (gfx950).
//
// For GEMMs whose A, B and C are single blocks (dense contractions, merged
// DPD irrep GEMMs, batch instances): no pack pass.  Per 32-k stage the
// workgroup's 256 threads load the tile's A' (BM x 32) and B' (BN x 32)
// elements with lanes along whichever side of the operand is unit-stride
// (rows, or k), so that every wave reads whole cache lines, and store them
// in LDS as [k][row] with a one-element pad per k row (one stage per
// operand in LDS).  The next stage's loads are in flight, into registers,
// while the waves run this stage's MFMAs
// (v_mfma_f64_16x16x4, operands read from LDS with one 8-byte load per lane
// and block), two barriers per stage (before the next stage's stores and
// after them).  Element (r, k) of an operand is at
// tensor + src + rows[r] + koff(k), koff from the (base, stride) pair of
// k's 8-k group or, where the group is not affine, the per-element k
// offsets (PanelTask).  Loads go through buffer resources bounded by the
// operand's extent; lanes past the valid rows or k read zero.  The epilogue
// applies alpha / beta through C's offsets (beta == 0 does not read C,
// src/kernels/3m/gemm.hpp:54-58), or, for a task with c_ld > 0, stores the
// unscaled sums into a dense C panel in cws.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace tblis_amd
{

namespace
{

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kBK = 32;   // k per stage
constexpr int kNT = 256;  // threads per workgroup

template <typename TC>
__device__ __forceinline__ double lds_src_load(__amdgpu_buffer_rsrc_t r, uint32_t byte_off)
{
    if constexpr (sizeof(TC) == 8)
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, byte_off, 0, 0));
    else
        return (double)__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, byte_off, 0, 0));
}

struct LdsSide
{
    __amdgpu_buffer_rsrc_t rsrc;
    const int64_t* kt;  // (base, stride) per 8-k group
    const int64_t* km;  // per-element k offsets
    int rows;           // valid rows of the tile
    int kfast;          // lanes along k (else along rows)
};

__device__ __forceinline__ uint32_t k_offset(const LdsSide& s, int k)
{
    const int64_t kb = s.kt[2 * (k >> 3)], ks = s.kt[2 * (k >> 3) + 1];
    return ks != SCATTERED ? (uint32_t)(kb + (k & 7) * ks) : (uint32_t)s.km[k];
}

// This thread's elements of one stage (k from k0):
//   rows along lanes: r = tid % R, k = tid / R + (kNT / R) j  (k wave-uniform)
//   k along lanes:    r = tid / kBK + (kNT / kBK) j, k = tid % kBK
template <int R, typename TC>
__device__ __forceinline__ void stage_load(double (&v)[R * kBK / kNT], const LdsSide& s, const uint32_t* srow,
                                           int k0, int klen, int tid)
{
    constexpr int NE = R * kBK / kNT;
    if (!s.kfast)
    {
        // k = k0 + tid / R + KT j: element j lies in the stage's k group
        // (KT j) / 8 (tid / R < KT), so the four groups' (base, stride)
        // pairs are loaded once per stage (wave-uniform addresses); only
        // groups that are not affine read per-element k offsets
        constexpr int KT = kNT / R;
        static_assert(KT <= 8 && NE * KT == kBK, "stage shape");
        const int r = tid % R;
        const bool rin = r < s.rows;
        const uint32_t ro = srow[r];
        const int g0 = k0 >> 3;
        int64_t kb[kBK / 8], ks[kBK / 8];
#pragma unroll
        for (int q = 0; q < kBK / 8; q++)
        {
            const bool gin = k0 + 8 * q < klen;
            kb[q] = gin ? s.kt[2 * (g0 + q)] : 0;
            ks[q] = gin ? s.kt[2 * (g0 + q) + 1] : 0;
        }
#pragma unroll
        for (int j = 0; j < NE; j++)
        {
            const int q = (KT * j) >> 3;
            const int k = k0 + tid / R + KT * j;
            const bool in = rin && k < klen;
            uint32_t ko = 0u;
            if (k < klen) ko = ks[q] != SCATTERED ? (uint32_t)(kb[q] + (k & 7) * ks[q]) : (uint32_t)s.km[k];
            v[j] = lds_src_load<TC>(s.rsrc, in ? (ro + ko) * (uint32_t)sizeof(TC) : 0xffffffffu);
        }
    }
    else
    {
        const int k = k0 + tid % kBK;
        const bool kin = k < klen;
        const uint32_t ko = kin ? k_offset(s, k) : 0u;
#pragma unroll
        for (int j = 0; j < NE; j++)
        {
            const int r = tid / kBK + (kNT / kBK) * j;
            const bool in = kin && r < s.rows;
            v[j] = lds_src_load<TC>(s.rsrc, in ? (srow[r] + ko) * (uint32_t)sizeof(TC) : 0xffffffffu);
        }
    }
}

// The stage into LDS as [k][row], rows padded to R + 1.
template <int R>
__device__ __forceinline__ void stage_store(double* __restrict__ lds, const double (&v)[R * kBK / kNT], int kfast,
                                            int tid)
{
    constexpr int NE = R * kBK / kNT;
    if (!kfast)
    {
        const int r = tid % R;
#pragma unroll
        for (int j = 0; j < NE; j++) lds[(tid / R + (kNT / R) * j) * (R + 1) + r] = v[j];
    }
    else
    {
        const int k = tid % kBK;
#pragma unroll
        for (int j = 0; j < NE; j++) lds[k * (R + 1) + tid / kBK + (kNT / kBK) * j] = v[j];
    }
}

template <int BM, int BN, int WGM, int WGN, typename TC>
__global__ __launch_bounds__(kNT) void gemm_lds_kernel(const PanelTask* __restrict__ tasks, int ntasks,
                                                       const int64_t* __restrict__ pool, const void* A,
                                                       const void* B, TC* __restrict__ C,
                                                       double* __restrict__ cws, double alpha, double beta,
                                                       int use_k)
{
    static_assert(WGM * WGN * 64 == kNT, "four waves");
    constexpr int WM = BM / WGM, WN = BN / WGN, MB = WM / 16, NB = WN / 16;
    constexpr int NEA = BM * kBK / kNT, NEB = BN * kBK / kNT;
    // one stage of each operand: the next stage waits in registers while
    // the waves compute from LDS (half the LDS of double buffering, twice
    // the resident workgroups)
    __shared__ double As[kBK * (BM + 1)];
    __shared__ double Bs[kBK * (BN + 1)];
    __shared__ uint32_t srow_a[BM], srow_b[BN];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave % WGM, wn = wave / WGM;
    const int lr = lane & 15, lk = lane >> 4;

    for (int ti = blockIdx.x; ti < ntasks; ti += gridDim.x)
    {
        const PanelTask t = tasks[ti];
        for (int r = tid; r < BM; r += kNT) srow_a[r] = r < t.mlen ? (uint32_t)pool[t.a_rt + r] : 0u;
        for (int r = tid; r < BN; r += kNT) srow_b[r] = r < t.nlen ? (uint32_t)pool[t.b_rt + r] : 0u;
        LdsSide sa, sb;
        sa.rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)((const TC*)A + t.a_src), (short)0, (int)t.a_bytes,
                                                    0x00020000);
        sb.rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)((const TC*)B + t.b_src), (short)0, (int)t.b_bytes,
                                                    0x00020000);
        sa.kt = pool + t.a_kt;
        sa.km = pool + t.a_km;
        sb.kt = pool + t.b_kt;
        sb.km = pool + t.b_km;
        sa.rows = t.mlen;
        sb.rows = t.nlen;
        sa.kfast = t.pad_ & 1;
        sb.kfast = (t.pad_ >> 1) & 1;

        v4d acc[MB][NB];
#pragma unroll
        for (int i = 0; i < MB; i++)
#pragma unroll
            for (int j = 0; j < NB; j++) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

        const int nst = use_k ? (t.klen + kBK - 1) / kBK : 0;
        double va[NEA], vb[NEB];
        __syncthreads();  // row offsets in LDS
        if (nst > 0)
        {
            stage_load<BM, TC>(va, sa, srow_a, 0, t.klen, tid);
            stage_load<BN, TC>(vb, sb, srow_b, 0, t.klen, tid);
            stage_store<BM>(As, va, sa.kfast, tid);
            stage_store<BN>(Bs, vb, sb.kfast, tid);
        }
        __syncthreads();
        for (int s = 0; s < nst; s++)
        {
            if (s + 1 < nst)
            {
                stage_load<BM, TC>(va, sa, srow_a, (s + 1) * kBK, t.klen, tid);
                stage_load<BN, TC>(vb, sb, srow_b, (s + 1) * kBK, t.klen, tid);
            }
            const double* __restrict__ a_s = As + wm * WM + lr;
            const double* __restrict__ b_s = Bs + wn * WN + lr;
            double fa[2][MB], fb[2][NB];
#pragma unroll
            for (int i = 0; i < MB; i++) fa[0][i] = a_s[lk * (BM + 1) + i * 16];
#pragma unroll
            for (int j = 0; j < NB; j++) fb[0][j] = b_s[lk * (BN + 1) + j * 16];
#pragma unroll
            for (int q = 0; q < kBK / 4; q++)
            {
                const int c = q & 1;
                if (q + 1 < kBK / 4)
                {
                    const int kr = (q + 1) * 4 + lk;
#pragma unroll
                    for (int i = 0; i < MB; i++) fa[c ^ 1][i] = a_s[kr * (BM + 1) + i * 16];
#pragma unroll
                    for (int j = 0; j < NB; j++) fb[c ^ 1][j] = b_s[kr * (BN + 1) + j * 16];
                }
#pragma unroll
                for (int i = 0; i < MB; i++)
#pragma unroll
                    for (int j = 0; j < NB; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[c][i], fb[c][j], acc[i][j], 0, 0, 0);
            }
            if (s + 1 < nst)
            {
                __syncthreads();  // every wave has read this stage
                stage_store<BM>(As, va, sa.kfast, tid);
                stage_store<BN>(Bs, vb, sb.kfast, tid);
            }
            __syncthreads();
        }

        // D register r of lane l holds D[l / 16 + 4 r][l % 16] of each block
        const int cl = lane & 15, rl = lane >> 4;
        if (t.c_ld > 0)
        {
            // a C panel (leading dimension c_ld at c_src in cws) takes the
            // unscaled sums: the tile timing runs write there
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
            continue;
        }
        // c_ld == 0: C through its row and column offsets (the planner gives
        // these tiles single-block C only)
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
}

struct LdsTile
{
    int bm, bn, wgm, wgn;
};

// id (after the register-fed tiles), BM, BN, waves WGM x WGN (four waves)
#define TBLIS_LDS_TILES(X) \
    X(0, 128, 128, 2, 2)   \
    X(1, 64, 64, 2, 2)     \
    X(2, 128, 64, 2, 2)

constexpr LdsTile kLTiles[] = {
#define X(id, bm, bn, wgm, wgn) {bm, bn, wgm, wgn},
    TBLIS_LDS_TILES(X)
#undef X
};
constexpr int kNumLTiles = sizeof(kLTiles) / sizeof(kLTiles[0]);

template <typename TC>
hipError_t launch_lds_t(const PanelLaunch& L, int lt)
{
    switch (lt)
    {
#define X(id, bm, bn, wgm, wgn)                                                                          \
    case id:                                                                                             \
    {                                                                                                    \
        auto* k = &gemm_lds_kernel<bm, bn, wgm, wgn, TC>;                                                \
        const dim3 grid(persistent_grid((const void*)k, kNT, L.ntasks));                                \
        hipLaunchKernelGGL(k, grid, dim3(kNT), 0, L.stream, L.tasks, L.ntasks, L.pool, L.A, L.B, (TC*)L.C, \
                           L.cws, L.alpha, L.beta, L.use_k);                                             \
        return hipGetLastError();                                                                        \
    }
        TBLIS_LDS_TILES(X)
#undef X
        default:
            return hipErrorInvalidValue;
    }
}

}  // namespace

int lds_num_tiles() { return kNumLTiles; }
int lds_tile_m(int lt) { return kLTiles[lt].bm; }
int lds_tile_n(int lt) { return kLTiles[lt].bn; }

hipError_t launch_gemm_lds(const PanelLaunch& L, int lt)
{
    if (L.ntasks == 0) return hipSuccess;
    if (lt < 0 || lt >= kNumLTiles || !L.direct_a || !L.direct_b) return hipErrorInvalidValue;
    return L.dtype == 0 ? launch_lds_t<double>(L, lt) : launch_lds_t<float>(L, lt);
}

}  // namespace tblis_amd
