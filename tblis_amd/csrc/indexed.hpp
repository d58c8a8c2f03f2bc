// Indexed-tensor contraction (MArray indexed_varray operands).
#pragma once

#include "common.hpp"

namespace tblis_amd
{

// tblis::mult over indexed_varray_view (src/iface/3t/mult.cxx:299-402,
// src/internal/3t/indexed/mult.cxx:63-318).
void indexed_tensor_mult(const tblis_config* cfg, const tblis_indexed_tensor* A, const label_type* idx_A,
                         const tblis_indexed_tensor* B, const label_type* idx_B,
                         tblis_indexed_tensor* C, const label_type* idx_C);

// tblis::mult over indexed_dpd_varray_view (src/iface/3t/mult.cxx:413-540):
// the blocked algorithm (src/internal/3t/indexed_dpd/mult.cxx:117-994, the
// reference's default), or with TBLIS_HIP_INDEXED_DPD=full the FULL one
// (src/internal/3t/indexed_dpd/mult.cxx:68-118; the reference's
// dpd_impl == FULL).
void indexed_dpd_tensor_mult(const tblis_config* cfg, const tblis_indexed_dpd_tensor* A,
                             const label_type* idx_A, const tblis_indexed_dpd_tensor* B,
                             const label_type* idx_B, tblis_indexed_dpd_tensor* C, const label_type* idx_C);

// The blocked algorithm on operands indexed_dpd_tensor_mult has checked:
// one panel GEMM per C region (entry, slice and C block) whose k blocks are
// the matching entry pairs over every irrep assignment of the contracted
// labels, read straight from the DPD storage (indexed.cpp).
void indexed_dpd_blocked(const tblis_config* cfg, const tblis_indexed_dpd_tensor* A, const label_type* idx_A,
                         const tblis_indexed_dpd_tensor* B, const label_type* idx_B,
                         tblis_indexed_dpd_tensor* C, const label_type* idx_C);

}  // namespace tblis_amd
