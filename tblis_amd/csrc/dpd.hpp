// DPD (irrep block-sparse) storage layout, as laid out by MArray's
// dpd_varray for a freshly allocated tensor
// (src/external/marray/include/dpd_marray_base.hpp:31-228).
#pragma once

#include <cstdint>

#include "common.hpp"

namespace tblis_amd
{

constexpr int DPD_MAXDIM = 16;

struct DpdLayout
{
    int ndim = 0, nirrep = 1, irrep = 0;
    bool row_major = true;
    int64_t len[DPD_MAXDIM][8] = {};      // [dim][irrep]
    int leaf[DPD_MAXDIM] = {};            // tree node of each (layout-order) dim
    int parent[2 * DPD_MAXDIM] = {};      // parent node of each tree node
    int perm[DPD_MAXDIM] = {};            // tensor dim -> layout-order dim
    int64_t size[2 * DPD_MAXDIM][8] = {}; // elements under each node, per irrep

    void init(unsigned irrep, unsigned nirrep, unsigned ndim, const len_type* len, int layout);
    int64_t total() const;
    // Offset, lengths and strides of the block with the given irreps.
    void block(const unsigned* irreps, int64_t* blen, int64_t* off, int64_t* bstride) const;
    // Dense-equivalent row/column-major strides over the summed extents
    // (src/internal/1t/dpd/util.hpp:136-173).
    void dense_strides(int64_t* stride) const;
};

}  // namespace tblis_amd
