// DPD layout: the storage order MArray gives a dpd_varray.
//
// The irreps of the dimensions are combined pairwise along a binary tree
// whose shape is fixed by the layout family (default_depth,
// dpd_marray_base.hpp:33-63); each interior node stores, per irrep of the
// combined index, the number of elements below it (set_size, :117-160); a
// block's address is the running sum over the tree of the sizes of the
// irrep combinations that precede it (get_block, :162-228).  Row-major
// layouts are column-major layouts of the reversed dimension order.
#include "dpd.hpp"

#include <cstring>

namespace tblis_amd
{
namespace
{

void default_depth(int layout, int ndim, int* depth)
{
    if (layout == TBLIS_DPD_BALANCED_COLUMN_MAJOR || layout == TBLIS_DPD_BALANCED_ROW_MAJOR)
    {
        unsigned dl = 31 - __builtin_clz((unsigned)ndim);
        unsigned du = dl + ((ndim & (ndim - 1)) ? 1 : 0);
        unsigned o = ndim - (1u << dl);
        unsigned k = 2 * ((o + 1) / 2);
        unsigned l = 2 * (o / 2);
        unsigned h = (ndim + 1) / 2;
        for (unsigned i = 0; i < (unsigned)ndim; i++)
        {
            if (i < k)
                depth[i] = du;
            else if (i < h)
                depth[i] = dl;
            else if (i < h + l)
                depth[i] = du;
            else
                depth[i] = dl;
        }
    }
    else if (layout == TBLIS_DPD_PREFIX_COLUMN_MAJOR || layout == TBLIS_DPD_PREFIX_ROW_MAJOR)
    {
        depth[0] = ndim - 1;
        for (int i = 1; i < ndim; i++) depth[i] = ndim - i;
    }
    else
    {
        for (int i = 0; i < ndim - 1; i++) depth[i] = i + 1;
        depth[ndim - 1] = ndim - 1;
    }
}

}  // namespace

void DpdLayout::init(unsigned irrep_, unsigned nirrep_, unsigned ndim_, const len_type* lens,
                     int layout)
{
    if (ndim_ == 0 || ndim_ > (unsigned)DPD_MAXDIM) fail("DPD tensor: ndim must be 1..16");
    if (!(nirrep_ == 1 || nirrep_ == 2 || nirrep_ == 4 || nirrep_ == 8))
        fail("DPD tensor: nirrep must be 1, 2, 4 or 8");
    if (irrep_ >= nirrep_) fail("DPD tensor: irrep out of range");
    if (layout < 0 || layout > 5) fail("DPD tensor: unknown layout");
    ndim = ndim_;
    nirrep = nirrep_;
    irrep = irrep_;
    row_major = layout & 1;
    std::memset(len, 0, sizeof(len));
    std::memset(size, 0, sizeof(size));
    for (int i = 0; i < ndim; i++)
        for (int r = 0; r < nirrep; r++)
        {
            if (lens[i * nirrep + r] < 0) fail("DPD tensor: negative length");
            len[i][r] = lens[i * nirrep + r];
        }

    // Tree from the depth list: leaves at the deepest level pair up into
    // interior nodes level by level; nodes are numbered in creation order.
    int depth[DPD_MAXDIM], node[DPD_MAXDIM], leaf_of[DPD_MAXDIM];
    default_depth(layout, ndim, depth);
    int n = ndim;
    for (int i = 0; i < n; i++) leaf_of[i] = i;
    int pos = 0;
    for (int d = ndim; d-- > 0;)
    {
        for (int i = 0; i < n; i++)
        {
            if (depth[i] == d + 1)
            {
                parent[node[i]] = parent[node[i + 1]] = pos;
                for (int j = i + 1; j + 1 < n; j++)
                {
                    depth[j] = depth[j + 1];
                    node[j] = node[j + 1];
                    leaf_of[j] = leaf_of[j + 1];
                }
                n--;
                depth[i]--;
                node[i] = pos++;
                leaf_of[i] = -1;
            }
            else if (depth[i] == d)
            {
                node[i] = pos++;
                if (leaf_of[i] != -1) leaf[leaf_of[i]] = node[i];
            }
        }
    }

    // Sizes: leaves carry the per-irrep lengths (reversed order for row
    // major), interior nodes the irrep-convolution of their children.
    for (int i = 0; i < ndim; i++)
    {
        int src = row_major ? ndim - 1 - i : i;
        perm[i] = src;
        for (int r = 0; r < nirrep; r++) size[leaf[src]][r] = len[i][r];
    }
    for (int i = 0; i < ndim - 1; i++)
    {
        int up = parent[2 * i];
        for (int r1 = 0; r1 < nirrep; r1++)
        {
            int64_t s = 0;
            for (int r2 = 0; r2 < nirrep; r2++) s += size[2 * i][r1 ^ r2] * size[2 * i + 1][r2];
            size[up][r1] = s;
        }
    }
}

int64_t DpdLayout::total() const { return size[2 * ndim - 2][irrep]; }

void DpdLayout::block(const unsigned* irreps, int64_t* blen, int64_t* off, int64_t* bstride) const
{
    int nirr[2 * DPD_MAXDIM];
    int64_t nstride[2 * DPD_MAXDIM];
    for (int i = 0; i < ndim; i++) nirr[leaf[perm[i]]] = irreps[i];
    for (int i = 0; i < ndim - 1; i++) nirr[parent[2 * i]] = nirr[2 * i] ^ nirr[2 * i + 1];
    nstride[2 * ndim - 2] = 1;
    int64_t o = 0;
    for (int i = ndim - 1; i-- > 0;)
    {
        int up = parent[2 * i];
        nstride[2 * i] = nstride[up];
        nstride[2 * i + 1] = nstride[2 * i] * size[2 * i][nirr[2 * i]];
        int64_t before = 0;
        for (int r = 0; r < nirr[2 * i + 1]; r++) before += size[2 * i][r ^ nirr[up]] * size[2 * i + 1][r];
        o += before * nstride[2 * i];
    }
    for (int i = 0; i < ndim; i++)
    {
        bstride[i] = nstride[leaf[perm[i]]];
        blen[i] = len[i][irreps[i]];
    }
    *off = o;
}

void DpdLayout::dense_strides(int64_t* stride) const
{
    int64_t tot[DPD_MAXDIM];
    for (int j = 0; j < ndim; j++)
    {
        tot[j] = 0;
        for (int r = 0; r < nirrep; r++) tot[j] += len[j][r];
    }
    int64_t s = 1;
    for (int q = 0; q < ndim; q++)
    {
        // perm is its own inverse (identity or reversal)
        int d = perm[q];
        stride[d] = s;
        s *= tot[d];
    }
}

}  // namespace tblis_amd
