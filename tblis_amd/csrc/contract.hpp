// High-level contraction entry points behind the C ABI.
#pragma once

#include <memory>

#include "common.hpp"
#include "level1.hpp"

namespace tblis_amd
{

struct GemmDesc;

void tensor_mult(const tblis_config* cfg, const tblis_tensor* A, const label_type* idx_A,
                 const tblis_tensor* B, const label_type* idx_B, tblis_tensor* C,
                 const label_type* idx_C);

// tensor_mult without device allocation: plan tables and workspace in the
// device scratch of noalloc.hip (real types, panel path; no plan cache,
// no autotuning).  Calls on one GPU share the scratch and must not overlap.
void tensor_mult_noalloc(const tblis_config* cfg, const tblis_tensor* A, const label_type* idx_A,
                         const tblis_tensor* B, const label_type* idx_B, tblis_tensor* C,
                         const label_type* idx_C);
// The current device's scratch (noalloc.hip) and its size.
char* noalloc_scratch(size_t& bytes);

void tensor_mult_batch(const tblis_config* cfg, int64_t count, const tblis_tensor* const* A,
                       const label_type* const* idx_A, const tblis_tensor* const* B,
                       const label_type* const* idx_B, tblis_tensor* const* C,
                       const label_type* const* idx_C);

// tblis_matrix_mult as a two-dimensional tensor contraction.
void matrix_mult(const tblis_config* cfg, const tblis_matrix* A, const tblis_matrix* B,
                 tblis_matrix* C);

void dpd_tensor_mult(const tblis_config* cfg, const tblis_dpd_tensor* A, const label_type* idx_A,
                     const tblis_dpd_tensor* B, const label_type* idx_B, tblis_dpd_tensor* C,
                     const label_type* idx_C);

int64_t debug_canonical(const tblis_tensor* A, const label_type* idx_A, const tblis_tensor* B,
                        const label_type* idx_B, const tblis_tensor* C, const label_type* idx_C,
                        int64_t* out, int64_t max_out);

int debug_dpd_patches(const tblis_dpd_tensor* A, const label_type* idx_A,
                      const tblis_dpd_tensor* B, const label_type* idx_B,
                      const tblis_dpd_tensor* C, const label_type* idx_C, int which,
                      unsigned irrep_ab, int64_t BS_row, int64_t BS_col, int* npatch,
                      int64_t* rows, int64_t* cols, int64_t* base, int64_t* rscat,
                      int64_t* cscat, int64_t max_patch, int64_t max_scat);

// Complex (scomplex / dcomplex) contraction: two real panel GEMMs per
// complex GEMM (complex.cpp).
struct PanelGemm;
struct Plan;
// One complex scaling pass of run_complex_plan: C elements (complex)
// described by desc at data.
struct CPass
{
    void* data;
    Level1Desc desc;
};
// The elements of C a contraction writes (C's diagonal for repeated
// labels) as a level-1 descriptor in complex elements.
Level1Desc c_desc(const tblis_tensor& C, const label_type* idx_C);
// The two real GEMMs (Re C, Im C) of each complex GEMM of in (offsets in
// complex elements; sa / sb = -1 conjugates A / B).
void complex_split_gemms(const std::vector<PanelGemm>& in, double sa, double sb, std::vector<PanelGemm>& out);
// Runs a plan of split complex GEMMs with complex alpha and beta: real
// ones go to the GEMM epilogue, complex ones to level-1 passes over the C
// elements each pass describes (dt 0 double, 1 float).
void run_complex_plan(const std::shared_ptr<Plan>& plan, const void* A, const void* B, void* C, double alpha_re,
                      double alpha_im, double beta_re, double beta_im, const std::vector<CPass>& cps, int dt);
void complex_tensor_mult(const tblis_config* cfg, const tblis_tensor* A, const label_type* idx_A,
                         const tblis_tensor* B, const label_type* idx_B, tblis_tensor* C,
                         const label_type* idx_C);

// The batched GEMM of a canonical contraction (m = AC, n = BC, k = AB,
// batch = ABC) at the given element bases; false when C is empty.
bool dense_gemm(const Canonical& c, GemmDesc& g, int64_t base_a, int64_t base_b, int64_t base_c);

// Descriptor helpers shared by the entry points.
int elem_size(type_t t);
void check_real(type_t t);                                  // complex types: error
double scalar_value(const tblis_scalar& s, type_t t);       // as double
double product_scalar(const tblis_scalar& a, const tblis_scalar& b, type_t t);
void reset_scalar(tblis_tensor* C);                         // scalar = 1, conj = 0
void tensor_extent(const tblis_tensor& t, int64_t& lo, int64_t& hi);
void check_data(const tblis_tensor* t, const char* name);   // device memory, in bounds
void check_tensor(const tblis_tensor* t, const label_type* idx, const char* name);

// Selected algorithm for a type under a config.
Algo resolve_algo(const tblis_config* cfg, type_t type);
// Kernel family (plan_types.hpp FAMILY_*) of an algorithm.
int algo_family(Algo a);

}  // namespace tblis_amd
