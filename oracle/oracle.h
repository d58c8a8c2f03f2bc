/*
 * oracle.h -- CPU restatement of TBLIS 1.2.0's tensor-contraction path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP library is compared
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it; nothing under tblis_amd/ includes, links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the TBLIS source tree).  How each function is pinned (golden vectors, closed
 * forms, brute force, invariants) is listed in oracle.c's header and checked
 * by tests/test_oracle.py.
 */
#ifndef TBLIS_ORACLE_H
#define TBLIS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXDIM 16

/* ------------------------------------------------------------------ */
/* Canonical form of one dense contraction (src/iface/3t/mult.cxx:24-116) */
/* ------------------------------------------------------------------ */
typedef struct orc_groups
{
    int valid;           /* 0: A-, B- or C-only labels, or length mismatch */
    int n_ab, n_ac, n_bc, n_abc;
    char idx_ab[ORC_MAXDIM], idx_ac[ORC_MAXDIM], idx_bc[ORC_MAXDIM], idx_abc[ORC_MAXDIM];
    int64_t len_ab[ORC_MAXDIM], len_ac[ORC_MAXDIM], len_bc[ORC_MAXDIM], len_abc[ORC_MAXDIM];
    int64_t sA_ab[ORC_MAXDIM], sB_ab[ORC_MAXDIM];
    int64_t sA_ac[ORC_MAXDIM], sC_ac[ORC_MAXDIM];
    int64_t sB_bc[ORC_MAXDIM], sC_bc[ORC_MAXDIM];
    int64_t sA_abc[ORC_MAXDIM], sB_abc[ORC_MAXDIM], sC_abc[ORC_MAXDIM];
} orc_groups;

/* The three logical matrices handed to TensorGEMM by mult_blis
 * (src/internal/3t/dense/mult.cxx:197-266 and :460-571). */
typedef struct orc_tmat
{
    int nm, nn, nk, nl;
    int64_t len_m[ORC_MAXDIM], len_n[ORC_MAXDIM], len_k[ORC_MAXDIM], len_l[ORC_MAXDIM];
    int64_t sA_m[ORC_MAXDIM], sA_k[ORC_MAXDIM];
    int64_t sB_k[ORC_MAXDIM], sB_n[ORC_MAXDIM];
    int64_t sC_m[ORC_MAXDIM], sC_n[ORC_MAXDIM];
    int64_t sA_l[ORC_MAXDIM], sB_l[ORC_MAXDIM], sC_l[ORC_MAXDIM];
    int pack_m_3d, pack_n_3d, pack_k_3d;
} orc_tmat;

int  orc_diagonal(int ndim, const int64_t* len, const int64_t* stride, const char* idx,
                  int64_t* len_out, int64_t* stride_out, char* idx_out);
void orc_sort_by_stride(int n, int ntensor, const int64_t* const* strides, int* perm);
int  orc_fold(int ndim, int64_t* len, char* idx, int ntensor, int64_t* const* strides);
char orc_free_idx(int n, const char* labels);
int  orc_unit_dim(const int64_t* stride, const int* reorder, int n);

int  orc_canonicalize(int ndim_A, const int64_t* len_A, const int64_t* stride_A, const char* idx_A,
                      int ndim_B, const int64_t* len_B, const int64_t* stride_B, const char* idx_B,
                      int ndim_C, const int64_t* len_C, const int64_t* stride_C, const char* idx_C,
                      orc_groups* g);
void orc_blis_reorder(const orc_groups* g, orc_tmat* t);

/* ------------------------------------------------------------------ */
/* Block-scatter vectors (src/matrix/block_scatter_matrix.hpp:18-258)   */
/* ------------------------------------------------------------------ */
int64_t orc_gcd(int64_t a, int64_t b);
int64_t orc_lcm(int64_t a, int64_t b);
void orc_fill_scatter(int ndim, const int64_t* len, const int64_t* stride, int64_t BS,
                      int64_t off, int64_t size, int64_t* scat, int pack_3d, int elem_size);
void orc_fill_block_stride(int64_t BS, int64_t size, const int64_t* scat, int64_t* bs);

/* ------------------------------------------------------------------ */
/* Reference contraction (src/internal/3t/dense/mult.cxx:653-720, 777-837) */
/* ------------------------------------------------------------------ */
/* Full tblis_tensor_mult semantics on host memory, FP64 (acc = 0) or
 * extended-precision accumulation (acc = 1, long double). */
void orc_tensor_mult_d(int acc,
                       double alpha, int ndim_A, const int64_t* len_A, const int64_t* stride_A,
                       const char* idx_A, const double* A,
                       int ndim_B, const int64_t* len_B, const int64_t* stride_B,
                       const char* idx_B, const double* B,
                       double beta, int ndim_C, const int64_t* len_C, const int64_t* stride_C,
                       const char* idx_C, double* C);
void orc_tensor_mult_s(double alpha, int ndim_A, const int64_t* len_A, const int64_t* stride_A,
                       const char* idx_A, const float* A,
                       int ndim_B, const int64_t* len_B, const int64_t* stride_B,
                       const char* idx_B, const float* B,
                       double beta, int ndim_C, const int64_t* len_C, const int64_t* stride_C,
                       const char* idx_C, double* C_out, const float* C_in);
/* Selected entries of C = alpha*A.B + beta*C0 for a canonical contraction,
 * entries given as flat (l, m, n) logical indices; extended-precision sums. */
void orc_mult_entries_d(const orc_groups* g, double alpha, const void* A, const void* B,
                        double beta, const void* C0, int is_float, int64_t nent,
                        const int64_t* ent_l, const int64_t* ent_m, const int64_t* ent_n,
                        double* out);

/* Multithreaded blocked GEMM over scatter vectors (Goto loop nest of
 * src/nodes/gemm.hpp:178-187 with the pack semantics of
 * src/kernels/3m/packm.hpp:115-404).  CPU baseline only. */
double orc_blis_contract_d(int nthreads, double alpha, int ndim_A, const int64_t* len_A,
                           const int64_t* stride_A, const char* idx_A, const double* A,
                           int ndim_B, const int64_t* len_B, const int64_t* stride_B,
                           const char* idx_B, const double* B,
                           double beta, int ndim_C, const int64_t* len_C, const int64_t* stride_C,
                           const char* idx_C, double* C, int64_t n_limit);

/* ------------------------------------------------------------------ */
/* DPD storage (src/external/marray/include/dpd_marray_base.hpp:31-228) */
/* ------------------------------------------------------------------ */
typedef struct orc_dpd
{
    int ndim, nirrep, irrep, row_major;
    int64_t len[ORC_MAXDIM][8];                 /* as given: [dim][irrep] */
    int64_t ext[ORC_MAXDIM][8];                 /* after set_size's row-major swap */
    int leaf[ORC_MAXDIM], parent[2 * ORC_MAXDIM], perm[ORC_MAXDIM], depth[ORC_MAXDIM];
    int64_t size[2 * ORC_MAXDIM][8];
} orc_dpd;

/* layout: 0..5 = BALANCED_COL, BALANCED_ROW, BLOCKED_COL, BLOCKED_ROW,
 * PREFIX_COL, PREFIX_ROW (src/external/marray/include/utility.hpp:100-122) */
void    orc_dpd_init(orc_dpd* d, int irrep, int nirrep, int ndim, const int64_t* len, int layout);
int64_t orc_dpd_size(const orc_dpd* d);
void    orc_dpd_get_block(const orc_dpd* d, const int* irreps, int64_t* len, int64_t* offset,
                          int64_t* stride);
/* DPD contraction C = alpha*A.B + beta*C over symmetry-allowed blocks,
 * block by block (semantics of src/internal/3t/dpd/mult.cxx:29-74/671-847). */
void orc_dpd_mult_d(int acc, double alpha, const orc_dpd* dA, const char* idx_A, const double* A,
                    const orc_dpd* dB, const char* idx_B, const double* B,
                    double beta, const orc_dpd* dC, const char* idx_C, double* C);
void orc_dpd_mult_s(double alpha, const orc_dpd* dA, const char* idx_A, const float* A,
                    const orc_dpd* dB, const char* idx_B, const float* B,
                    double beta, const orc_dpd* dC, const char* idx_C, double* C_out,
                    const float* C_in);
/* DPD patch scatter for one irrep_AB (src/matrix/dpd_tensor_matrix.hpp:41-101,
 * src/matrix/patch_block_scatter_matrix.hpp:74-229), whole matrix (offset 0).
 * which: 0 = A (AC x AB), 1 = B (AB x BC), 2 = C (AC x BC).
 * Writes per-patch row/col scatter concatenated in patch order (row patch
 * major); returns the number of patches, fills npatch[2]. */
int orc_dpd_patch_scatter(const orc_dpd* dA, const char* idx_A, const orc_dpd* dB,
                          const char* idx_B, const orc_dpd* dC, const char* idx_C,
                          int which, int irrep_ab, int64_t BS_row, int64_t BS_col,
                          int* npatch, int64_t* patch_rows, int64_t* patch_cols,
                          int64_t* patch_base, int64_t* rscat, int64_t* cscat,
                          int64_t max_scat);

/* ------------------------------------------------------------------ */
/* Level-1 operations (src/iface/1t/add.cxx:24-110, scale.cxx:20-55,    */
/* set.cxx:20-51, src/internal/1t/dense/add.cxx:15-160)                 */
/* ------------------------------------------------------------------ */
/* mode 0 (add):   B <- alpha * sum_{labels only in A} A + beta * B
 * mode 1 (scale): B <- beta * B            (A unused)
 * mode 2 (set):   B <- alpha               (A unused)
 * over the distinct labels of B (a repeated label addresses a diagonal).
 * Length-1 dimensions are dropped first, as the reference's diagonal()
 * does, so a label of length 1 in one tensor only replicates or sums over
 * nothing.  beta == 0 writes B without reading it.  The value type is
 * double; the _s variant reads FP32 A / B_in and writes double B_out. */
void orc_level1_d(int mode, double alpha, int ndA, const int64_t* lenA, const int64_t* strA,
                  const char* idxA, const double* A, double beta, int ndB, const int64_t* lenB,
                  const int64_t* strB, const char* idxB, double* B);
void orc_level1_s(int mode, double alpha, int ndA, const int64_t* lenA, const int64_t* strA,
                  const char* idxA, const float* A, double beta, int ndB, const int64_t* lenB,
                  const int64_t* strB, const char* idxB, const float* B_in, double* B_out);

/* ------------------------------------------------------------------ */
/* Deterministic inputs                                                 */
/* ------------------------------------------------------------------ */
void orc_fill_uniform_d(double* x, int64_t n, uint64_t seed);

#ifdef __cplusplus
}
#endif

#endif
