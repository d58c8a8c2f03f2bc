/*
 * tblis.h -- C ABI of the MI355X tensor-contraction library (libtblis_amd.so).
 *
 * Drop-in for TBLIS 1.2.0's tensor-contraction entry point.  The types and
 * functions in the first half keep the reference's names, layouts and
 * argument meaning (citations are into the TBLIS source tree); the second
 * half holds the extensions this library adds (DPD and batched C entry
 * points, stream control, introspection for tests).
 *
 * Memory contract: tensor data pointers are device pointers (hipMalloc /
 * torch CUDA tensors) or managed memory on the current HIP device.  Host
 * pointers are rejected with an error.  Length, stride and label arrays are
 * host memory, read during the call only.
 */
#ifndef TBLIS_AMD_TBLIS_H
#define TBLIS_AMD_TBLIS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
#include <complex>
extern "C" {
#endif

/* ---------------------------------------------------------------------- */
/* Types (src/util/basic_types.h:53-77, 154-208, 471-545)                  */
/* ---------------------------------------------------------------------- */

typedef ptrdiff_t len_type;    /* TBLIS_LEN_TYPE default (configure.ac:91-123) */
typedef ptrdiff_t stride_type; /* TBLIS_STRIDE_TYPE default */
typedef char label_type;       /* TBLIS_LABEL_TYPE default */

typedef enum
{
    TYPE_SINGLE   = 0,
    TYPE_FLOAT    = TYPE_SINGLE,
    TYPE_DOUBLE   = 1,
    TYPE_SCOMPLEX = 2,
    TYPE_DCOMPLEX = 3
} type_t;

#ifdef __cplusplus
typedef std::complex<float> scomplex;
typedef std::complex<double> dcomplex;
#else
typedef float _Complex scomplex;
typedef double _Complex dcomplex;
#endif

/* 24 bytes: value union at 0, type at 16 (basic_types.h:154-208). */
typedef struct tblis_scalar
{
    union
    {
        float s;
        double d;
        float c[2];
        double z[2];
    } data;
    type_t type;
} tblis_scalar;

/* 64 bytes: type@0 conj@4 scalar@8 data@32 ndim@40 len@48 stride@56
 * (basic_types.h:471-545). */
typedef struct tblis_tensor
{
    type_t type;
    int conj;
    tblis_scalar scalar;
    void* data;
    unsigned ndim;
    len_type* len;
    stride_type* stride;
} tblis_tensor;

/* 72 bytes: type@0 conj@4 scalar@8 data@32 m@40 n@48 rs@56 cs@64
 * (src/util/basic_types.h:352-428). */
typedef struct tblis_matrix
{
    type_t type;
    int conj;
    tblis_scalar scalar;
    void* data;
    len_type m, n;
    stride_type rs, cs;
} tblis_matrix;

/* tci_comm, 24 bytes (src/external/tci/tci/communicator.h:8-15). */
typedef struct tblis_comm
{
    void* context;
    unsigned ngang;
    unsigned gid;
    unsigned nthread;
    unsigned tid;
} tblis_comm;

/* Opaque (basic_types.h:51); here it selects the contraction algorithm. */
typedef struct tblis_config_s tblis_config;

/* Single-thread communicator (src/util/thread.h:14). */
extern const tblis_comm* const tblis_single;

/* ---------------------------------------------------------------------- */
/* Descriptor setup (basic_types.h:549-579; basic_types.cxx:184-262)       */
/* ---------------------------------------------------------------------- */

void tblis_init_scalar_s(tblis_scalar* s, float value);
void tblis_init_scalar_d(tblis_scalar* s, double value);
void tblis_init_scalar_c(tblis_scalar* s, scomplex value);
void tblis_init_scalar_z(tblis_scalar* s, dcomplex value);

void tblis_init_tensor_scaled_s(tblis_tensor* t, float scalar, unsigned ndim, len_type* len,
                                float* data, stride_type* stride);
void tblis_init_tensor_scaled_d(tblis_tensor* t, double scalar, unsigned ndim, len_type* len,
                                double* data, stride_type* stride);
void tblis_init_tensor_scaled_c(tblis_tensor* t, scomplex scalar, unsigned ndim, len_type* len,
                                scomplex* data, stride_type* stride);
void tblis_init_tensor_scaled_z(tblis_tensor* t, dcomplex scalar, unsigned ndim, len_type* len,
                                dcomplex* data, stride_type* stride);
void tblis_init_tensor_s(tblis_tensor* t, unsigned ndim, len_type* len, float* data,
                         stride_type* stride);
void tblis_init_tensor_d(tblis_tensor* t, unsigned ndim, len_type* len, double* data,
                         stride_type* stride);
void tblis_init_tensor_c(tblis_tensor* t, unsigned ndim, len_type* len, scomplex* data,
                         stride_type* stride);
void tblis_init_tensor_z(tblis_tensor* t, unsigned ndim, len_type* len, dcomplex* data,
                         stride_type* stride);

/* ---------------------------------------------------------------------- */
/* The contraction (src/iface/3t/mult.h:17-20, mult.cxx:24-165)             */
/* ---------------------------------------------------------------------- */

/* C <- alpha_A*alpha_B * A.B + alpha_C * C over label-matched indices.
 * Repeated labels in one tensor take a diagonal, length-1 dimensions are
 * ignored, labels present in one tensor only are an error.  beta == 0
 * overwrites C without reading it; alpha == 0 reads neither A nor B.  After
 * the call C->scalar is 1 and C->conj is 0.  comm may be NULL or any
 * communicator (the GPU grid replaces the thread team).  cfg NULL selects
 * the default algorithm for the type (FP64 and complex double: FP64 MFMA;
 * FP32 and complex float: 3-term split-BF16 MFMA). */
void tblis_tensor_mult(const tblis_comm* comm, const tblis_config* cfg,
                       const tblis_tensor* A, const label_type* idx_A,
                       const tblis_tensor* B, const label_type* idx_B,
                       tblis_tensor* C, const label_type* idx_C);

/* Matrix descriptors (src/util/basic_types.h:430-465,
 * src/util/basic_types.cxx:86-184). */
void tblis_init_matrix_scaled_s(tblis_matrix* mat, float scalar, len_type m, len_type n, float* data,
                                stride_type rs, stride_type cs);
void tblis_init_matrix_scaled_d(tblis_matrix* mat, double scalar, len_type m, len_type n, double* data,
                                stride_type rs, stride_type cs);
void tblis_init_matrix_scaled_c(tblis_matrix* mat, scomplex scalar, len_type m, len_type n,
                                scomplex* data, stride_type rs, stride_type cs);
void tblis_init_matrix_scaled_z(tblis_matrix* mat, dcomplex scalar, len_type m, len_type n,
                                dcomplex* data, stride_type rs, stride_type cs);
void tblis_init_matrix_s(tblis_matrix* mat, len_type m, len_type n, float* data, stride_type rs,
                         stride_type cs);
void tblis_init_matrix_d(tblis_matrix* mat, len_type m, len_type n, double* data, stride_type rs,
                         stride_type cs);
void tblis_init_matrix_c(tblis_matrix* mat, len_type m, len_type n, scomplex* data, stride_type rs,
                         stride_type cs);
void tblis_init_matrix_z(tblis_matrix* mat, len_type m, len_type n, dcomplex* data, stride_type rs,
                         stride_type cs);

/* C (m x n) <- alpha_A*alpha_B * A (m x k) . B (k x n) + alpha_C * C
 * (src/iface/3m/mult.h:17-19, src/iface/3m/mult.cxx:19-191): the same panel
 * GEMM as tblis_tensor_mult with rs/cs as the two dimensions' strides.
 * Shapes must agree (A.m == C.m, B.n == C.n, A.n == B.m) and conj must be 0;
 * alpha == 0 or k == 0 scales C by beta (beta == 0 writes zeros); an empty
 * C is left alone; afterwards C->scalar = 1 and C->conj = 0. */
void tblis_matrix_mult(const tblis_comm* comm, const tblis_config* cfg, const tblis_matrix* A,
                       const tblis_matrix* B, tblis_matrix* C);

/* Level-1 operations around a contraction (src/iface/1t/add.h,
 * scale.h, set.h):
 *   add    B <- alpha_A * A + alpha_B * B over label-matched indices; labels
 *          only in A are summed, labels only in B are replicated (not both),
 *          repeated labels take a diagonal; alpha_A == 0 does not read A and
 *          alpha_B == 0 writes B without reading it
 *          (src/iface/1t/add.cxx:24-110)
 *   scale  A <- alpha_A * A (src/iface/1t/scale.cxx:20-55)
 *   set    A <- *alpha on every element the labels address
 *          (src/iface/1t/set.cxx:20-51)
 * Afterwards the written tensor's scalar is 1 and conj is 0. */
void tblis_tensor_add(const tblis_comm* comm, const tblis_config* cfg, const tblis_tensor* A,
                      const label_type* idx_A, tblis_tensor* B, const label_type* idx_B);
void tblis_tensor_scale(const tblis_comm* comm, const tblis_config* cfg, tblis_tensor* A,
                        const label_type* idx_A);
void tblis_tensor_set(const tblis_comm* comm, const tblis_config* cfg, const tblis_scalar* alpha,
                      tblis_tensor* A, const label_type* idx_A);

/* Thread-count controls (src/util/thread.h:21-23).  The GPU path does not
 * use host threads; the value is stored and returned for compatibility. */
unsigned tblis_get_num_threads(void);
void tblis_set_num_threads(unsigned num_threads);

/* Configuration lookup by name (src/util/configs.h:11).  Names:
 *   "mi355x"        default: FP64 MFMA for double, split-BF16 x3 for float
 *   "reference"     the reference's portable configuration: here the same
 *                   kernels as the default
 *   "fp64_mfma"     float contractions widened to FP64 MFMA
 *   "bf16x3"        3-term split-BF16 (6 products), FP32 accumulation
 *   "bf16x3_fast"   3-term split-BF16 without the compensated k-blocking
 *   "ozaki"         FP64-accurate for double and float: each row of A and B
 *                   scaled by a power of two and cut into 8 signed 7-bit
 *                   digits (6 or 7 with TBLIS_HIP_OZAKI_SLICES), digit-pair
 *                   products on int8 MFMA with exact int32 sums, combined in
 *                   FP64 (no counterpart in the reference; finite inputs)
 * Unknown names abort with a message, as in the reference. */
const tblis_config* tblis_get_config(const char* name);

/* ---------------------------------------------------------------------- */
/* Extension: DPD (irrep block-sparse) contraction.                        */
/* The reference exposes this only as the C++ template                     */
/* mult<T>(alpha, dpd_varray_view A, idx_A, ..., beta, dpd_varray_view C)  */
/* (src/iface/3t/mult.h:53-72, mult.cxx:170-290).  This C descriptor       */
/* carries what a freshly laid out MArray dpd_varray holds: irrep,         */
/* nirrep, per-irrep lengths [dim][irrep] and the layout code              */
/* (src/external/marray/include/utility.hpp:100-122).                      */
/* ---------------------------------------------------------------------- */

typedef enum
{
    TBLIS_DPD_BALANCED_COLUMN_MAJOR = 0,
    TBLIS_DPD_BALANCED_ROW_MAJOR = 1,
    TBLIS_DPD_BLOCKED_COLUMN_MAJOR = 2,
    TBLIS_DPD_BLOCKED_ROW_MAJOR = 3,
    TBLIS_DPD_PREFIX_COLUMN_MAJOR = 4,
    TBLIS_DPD_PREFIX_ROW_MAJOR = 5
} tblis_dpd_layout;

typedef struct tblis_dpd_tensor
{
    type_t type;
    int conj;
    tblis_scalar scalar;
    void* data;
    unsigned ndim;
    unsigned irrep;
    unsigned nirrep;      /* 1, 2, 4 or 8 */
    int layout;           /* tblis_dpd_layout */
    const len_type* len;  /* ndim x nirrep, row major: len[dim*nirrep + irrep] */
} tblis_dpd_tensor;

void tblis_init_dpd_tensor_d(tblis_dpd_tensor* t, double scalar, unsigned irrep, unsigned nirrep,
                             unsigned ndim, const len_type* len, double* data, int layout);
void tblis_init_dpd_tensor_s(tblis_dpd_tensor* t, float scalar, unsigned irrep, unsigned nirrep,
                             unsigned ndim, const len_type* len, float* data, int layout);
/* Complex DPD tensors: interleaved (re, im) elements in the same layout. */
void tblis_init_dpd_tensor_c(tblis_dpd_tensor* t, scomplex scalar, unsigned irrep, unsigned nirrep,
                             unsigned ndim, const len_type* len, scomplex* data, int layout);
void tblis_init_dpd_tensor_z(tblis_dpd_tensor* t, dcomplex scalar, unsigned irrep, unsigned nirrep,
                             unsigned ndim, const len_type* len, dcomplex* data, int layout);

/* Number of stored elements of a DPD tensor (dpd_varray_base::size). */
stride_type tblis_dpd_tensor_size(const tblis_dpd_tensor* t);

/* DPD contraction; same semantics as the dense call, block by block, for
 * all four types (complex: two real GEMMs per irrep GEMM over the
 * interleaved parts, conj honoured).  Each
 * contracted irrep is one GEMM over its symmetry-allowed blocks (when every
 * block address splits into a row part and a column part, the six MArray
 * layouts for the usual label orders, one GEMM with explicit row and k
 * offsets).  The GEMMs of all irreps run in one GEMM launch per kernel
 * variant (one workgroup per C tile), after one pack launch for the
 * operands that are not read straight from the DPD storage; C blocks are
 * written from the GEMM epilogue (split-BF16: through a C panel and one
 * unpack launch). */
void tblis_dpd_tensor_mult(const tblis_comm* comm, const tblis_config* cfg,
                           const tblis_dpd_tensor* A, const label_type* idx_A,
                           const tblis_dpd_tensor* B, const label_type* idx_B,
                           tblis_dpd_tensor* C, const label_type* idx_C);

/* ---------------------------------------------------------------------- */
/* Indexed tensors: C entry point for what the reference offers as the     */
/* C++ template tblis::mult<T> over MArray indexed_varray_view             */
/* (src/iface/3t/mult.h:74-86, src/iface/3t/mult.cxx:299-402).             */
/* ---------------------------------------------------------------------- */

/* The layout of MArray's indexed_varray (src/external/marray/include/
 * indexed_varray_base.hpp:19-45): the first dense_ndim dimensions are dense
 * and shared by every entry (lengths len[0..dense_ndim), strides
 * dense_stride); the others are indexed.  Entry i is the dense sub-tensor at
 * data[i] (device memory), placed at the index values
 * indices[i * (ndim - dense_ndim) + j] of the indexed dimensions and
 * multiplied by factor[i] (NULL: every factor is 1; of the tensor's own
 * type, so complex for scomplex / dcomplex entries).  Positions that no
 * entry lists are zero.  Entries must list distinct index values.  All four
 * types are accepted, as the reference instantiates them
 * (src/iface/3t/mult.cxx:405-409); conj conjugates a complex operand's
 * entries (not its factors). */
typedef struct tblis_indexed_tensor
{
    type_t type;
    int conj;
    tblis_scalar scalar;
    unsigned ndim;
    unsigned dense_ndim;
    const len_type* len;
    const stride_type* dense_stride;
    len_type num_indices;
    void* const* data;
    const len_type* indices;
    const void* factor;
} tblis_indexed_tensor;

void tblis_init_indexed_tensor_d(tblis_indexed_tensor* t, double scalar, unsigned ndim,
                                 unsigned dense_ndim, const len_type* len,
                                 const stride_type* dense_stride, len_type num_indices,
                                 double* const* data, const len_type* indices, const double* factor);
void tblis_init_indexed_tensor_s(tblis_indexed_tensor* t, float scalar, unsigned ndim,
                                 unsigned dense_ndim, const len_type* len,
                                 const stride_type* dense_stride, len_type num_indices,
                                 float* const* data, const len_type* indices, const float* factor);
void tblis_init_indexed_tensor_c(tblis_indexed_tensor* t, scomplex scalar, unsigned ndim,
                                 unsigned dense_ndim, const len_type* len,
                                 const stride_type* dense_stride, len_type num_indices,
                                 scomplex* const* data, const len_type* indices, const scomplex* factor);
void tblis_init_indexed_tensor_z(tblis_indexed_tensor* t, dcomplex scalar, unsigned ndim,
                                 unsigned dense_ndim, const len_type* len,
                                 const stride_type* dense_stride, len_type num_indices,
                                 dcomplex* const* data, const len_type* indices, const dcomplex* factor);

/* Indexed contraction C <- alpha A.B + beta C with alpha = A.scalar *
 * B.scalar and beta = C.scalar, following the reference
 * (src/iface/3t/mult.cxx:299-402, src/internal/3t/indexed/mult.cxx:274-318):
 * labels are unique within a tensor; first every C entry is scaled
 * (beta == 0: set to 0; beta != 1: data *= factor * beta, or 0 when its
 * factor is 0); then each C entry's data gains
 * alpha * factor_C * (A.B) at its index values, where A and B are the
 * full tensors their entries describe. */
void tblis_indexed_tensor_mult(const tblis_comm* comm, const tblis_config* cfg,
                               const tblis_indexed_tensor* A, const label_type* idx_A,
                               const tblis_indexed_tensor* B, const label_type* idx_B,
                               tblis_indexed_tensor* C, const label_type* idx_C);

/* Indexed DPD tensors (MArray indexed_dpd_varray,
 * src/external/marray/include/indexed_dpd_varray_base.hpp:21-213): the
 * first dense_ndim dimensions form a DPD tensor of irrep
 * irrep ^ idx_irrep[0] ^ ... in the given layout (lengths len[d*nirrep + r]);
 * indexed dimension j has the fixed irrep idx_irrep[j] and its index values
 * count within that irrep.  Entry i is the DPD storage at data[i] (device
 * memory), at indices[i * (ndim - dense_ndim) + j], times factor[i] (of
 * the tensor's type).  Real and complex types, as the reference
 * instantiates them (src/iface/3t/mult.cxx:539-543). */
typedef struct tblis_indexed_dpd_tensor
{
    type_t type;
    int conj;
    tblis_scalar scalar;
    unsigned ndim;
    unsigned dense_ndim;
    unsigned irrep;
    unsigned nirrep;
    int layout;                 /* tblis_dpd_layout */
    const len_type* len;        /* ndim x nirrep, row major */
    const unsigned* idx_irrep;  /* ndim - dense_ndim */
    len_type num_indices;
    void* const* data;
    const len_type* indices;
    const void* factor;
} tblis_indexed_dpd_tensor;

/* Indexed DPD contraction (src/iface/3t/mult.h:95-114,
 * src/iface/3t/mult.cxx:413-540): the same rules as
 * tblis_indexed_tensor_mult, block by block.  Computed as the reference's
 * FULL algorithm (src/internal/3t/indexed_dpd/mult.cxx:68-118): the
 * operands' DPD parts are expanded to dense entries on the device, the
 * indexed contraction runs on them, and C's symmetry-allowed blocks are
 * written back. */
void tblis_indexed_dpd_tensor_mult(const tblis_comm* comm, const tblis_config* cfg,
                                   const tblis_indexed_dpd_tensor* A, const label_type* idx_A,
                                   const tblis_indexed_dpd_tensor* B, const label_type* idx_B,
                                   tblis_indexed_dpd_tensor* C, const label_type* idx_C);

/* ---------------------------------------------------------------------- */
/* Extension: a batch of independent contractions of uneven shapes.  The */
/* tiles of all instances run in one GEMM launch per kernel variant (one  */
/* workgroup per C tile, grouped by which operands are read straight from */
/* the tensors), after one pack launch for the other operands; in FP64    */
/* every operand the GEMM can read directly is read so (re-read by at     */
/* most 64 tiles), and a batch of small instances is a single launch.     */
/* The reference has no counterpart; its equivalent is a host loop of     */
/* tblis_tensor_mult calls.  Every instance shares alpha, beta, the type  */
/* and the conj flags; complex types run as two real GEMMs per instance,  */
/* a complex alpha or beta adds one scaling pass per instance's C.        */
/* ---------------------------------------------------------------------- */
void tblis_tensor_mult_batch(const tblis_comm* comm, const tblis_config* cfg, len_type count,
                             const tblis_tensor* const* A, const label_type* const* idx_A,
                             const tblis_tensor* const* B, const label_type* const* idx_B,
                             tblis_tensor* const* C, const label_type* const* idx_C);

/* ---------------------------------------------------------------------- */
/* Extension: HIP execution control and error reporting.                   */
/* ---------------------------------------------------------------------- */

/* Stream used by the calls of this host thread (NULL = default stream). */
void tblis_hip_set_stream(void* hip_stream);
void* tblis_hip_get_stream(void);
/* 1: calls return after enqueueing; 0 (default): calls return when done. */
void tblis_hip_set_async(int async);
/* 0 (default): invalid arguments abort with a message, like the
 * reference's debug build.  1: they set the error string and return. */
void tblis_hip_set_error_mode(int return_on_error);
/* Last error of this host thread ("" if none); cleared by the next call. */
const char* tblis_hip_get_last_error(void);
/* Drop all cached contraction plans (device buffers are freed). */
void tblis_hip_clear_plan_cache(void);
/* Algorithm selected by the last contraction on this thread (config name). */
const char* tblis_hip_last_algorithm(void);
/* Autotuner (off unless TBLIS_HIP_AUTOTUNE=1): when on, building a plan for
 * a new set of GEMM shapes times every candidate tile of the panel GEMM on
 * scratch buffers (a tile that reads an operand straight from its tensor
 * also with every operand packed) and keeps the fastest; results are cached
 * per process and,
 * with TBLIS_HIP_TUNE_FILE=<path>, appended to / read from that file.  No
 * measurement happens while the stream is being captured. */
void tblis_hip_set_autotune(int on);
/* Number of tuning measurements this process has made. */
int tblis_hip_autotune_runs(void);
/* Panel GEMM tile of the last contraction on this thread. */
int tblis_hip_last_tile(void);
/* Operands that contraction's panel GEMM read straight from the tensors
 * instead of packing them: bit 0 A, bit 1 B (set when every GEMM task reads
 * it so); 0 for other kernels. */
int tblis_hip_last_direct(void);

/* No counterpart in the reference: tblis_tensor_mult for callers that
 * cannot allocate device memory (serving loops and harnesses that forbid
 * hipMalloc).  The same contraction, with its plan tables and its pack and
 * C-panel workspace in a 64 MiB device array of the library (one per GPU,
 * held from the library's first use on it until the process ends); the
 * call makes no hipMalloc or hipFree.  Real types (float, double) on the
 * panel path; a contraction whose tables and workspace exceed 64 MiB fails
 * with an error naming the sizes.  No plan cache and no autotuning: every
 * call plans on the host and uploads its tables in stream order into the
 * array.  Calls on one GPU share the array, which keeps the last call's
 * contents: calls that overlap on one GPU (other streams or host threads)
 * race. */
void tblis_hip_tensor_mult_noalloc(const tblis_comm* comm, const tblis_config* cfg,
                                   const tblis_tensor* A, const label_type* idx_A,
                                   const tblis_tensor* B, const label_type* idx_B,
                                   tblis_tensor* C, const label_type* idx_C);
/* Device allocations (plan tables, workspace, tuning scratch) this process
 * has made through the library. */
int64_t tblis_hip_device_alloc_count(void);

/* Introspection (host only, no device work): the canonical form of a
 * contraction as computed by this library's front end.
 * out_groups receives, for AB, AC, BC, ABC in order: ndim, then the folded
 * lengths and the per-tensor strides, as
 *   [n, len[n], sA[n], sB[n]]       for AB
 *   [n, len[n], sA[n], sC[n]]       for AC
 *   [n, len[n], sB[n], sC[n]]       for BC
 *   [n, len[n], sA[n], sB[n], sC[n]] for ABC
 * followed by the m/n/k/l matrix views after the stride sort and pack_3d
 * rotation ([n, len[n], s1[n], s2[n]] for m: A, C; n: B, C; k: A, B;
 * l: [n, len, sA, sB, sC]) and the three pack_3d flags.  Returns the number
 * of int64 written, or -1 for an invalid contraction. */
int64_t tblis_hip_debug_canonical(const tblis_tensor* A, const label_type* idx_A,
                                  const tblis_tensor* B, const label_type* idx_B,
                                  const tblis_tensor* C, const label_type* idx_C,
                                  int64_t* out, int64_t max_out);

/* Block-scatter vector and block strides of one matrix dimension given as
 * (len, stride) lists, at block size BS, rows [off, off+size), with the
 * pack_3d cache-line order for element size elem_size (same meaning as the
 * reference's block_scatter_matrix::fill_block_scatter). */
void tblis_hip_debug_fill_block_scatter(int ndim, const len_type* len, const stride_type* stride,
                                        len_type BS, len_type off, len_type size,
                                        stride_type* scat, stride_type* bs, int pack_3d,
                                        int elem_size);

/* The panel planner's tables of one matrix side given as (len, stride)
 * lists: its element offsets in odometer order (first dimension fastest;
 * the reference's plain scatter vector) into offs, and the (base, stride)
 * pair of every fragment group of size G with the side placed at panel row
 * r0 into tab (stride INT64_MIN where a group is not affine; a group of one
 * element has stride 1, as the reference's block strides give it).
 * Returns the number of offsets, -2 if max_out is too small, -1 on error. */
int64_t tblis_hip_debug_panel_side(int ndim, const len_type* len, const stride_type* stride, len_type G,
                                   len_type r0, stride_type* offs, stride_type* tab, int64_t max_out);

/* DPD block address (offset from the data pointer, lengths, strides) of the
 * block with the given irreps, from this library's DPD layout code. */
void tblis_hip_debug_dpd_block(const tblis_dpd_tensor* t, const unsigned* irreps,
                               len_type* len, stride_type* offset, stride_type* stride);

/* Patch list of the DPD GEMM of one irrep_AB, for matrix `which`
 * (0: A, 1: B, 2: C), block sizes BS_row/BS_col: number of patches and, per
 * patch (row-patch major), rows, cols, base offset and the concatenated
 * row/col scatter vectors.  Returns the number of patches, or -1. */
int tblis_hip_debug_dpd_patches(const tblis_dpd_tensor* A, const label_type* idx_A,
                                const tblis_dpd_tensor* B, const label_type* idx_B,
                                const tblis_dpd_tensor* C, const label_type* idx_C,
                                int which, unsigned irrep_ab, len_type BS_row, len_type BS_col,
                                int* npatch, len_type* rows, len_type* cols, stride_type* base,
                                stride_type* rscat, stride_type* cscat, int64_t max_patch,
                                int64_t max_scat);

/* One 16x16x4 FP64 MFMA in the operand layout the GEMM kernels assume:
 * A[i][k] = 1000*i + k + 1, B[k][j] = (k+1)*(j+1); writes the 16x16 result
 * row-major to out_device (256 doubles, device memory) using the output
 * lane layout the kernels assume.  Returns 0 on success. */
int tblis_hip_debug_mfma_probe(double* out_device);

/* One 32x32x16 BF16 MFMA in the operand and output layouts the split-BF16
 * kernels assume: A (32x16) and B (16x32) row-major FP32 in device memory
 * (values exactly representable in BF16), D (32x32) row-major FP32 out. */
int tblis_hip_debug_mfma_probe_bf16(const float* A, const float* B, float* D);

#ifdef __cplusplus
}
#endif

#endif
