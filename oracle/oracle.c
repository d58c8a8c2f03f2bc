/*
 * oracle.c -- CPU restatement of TBLIS 1.2.0's tensor-contraction path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * The reference cannot be built here from its own sources without the
 * configure-generated tblis_config.h, so this restatement is pinned by:
 *   - orc_dpd_init/get_block: the known-answer table of
 *     src/external/marray/test/dpd_varray.cxx:7-55 (6 layouts, 8 blocks:
 *     offsets, lengths, strides), transcribed as tests/golden/dpd_varray_kat.json,
 *     and the closed-form total size of dpd_marray_base.hpp:460-496;
 *   - orc_fill_scatter (plain order): the closed form "offset of the
 *     odometer-ordered multi-index off+i"; (pack_3d order): permutation of
 *     the plain range, block structure, brute-force on tiny shapes;
 *   - orc_fill_block_stride: brute force on arithmetic progressions;
 *   - orc_diagonal/fold/canonicalize/blis_reorder: the worked canonical forms
 *     of the dense workloads (hand-derived from the reference), invariance of
 *     the contraction result under relabelling and dimension permutation, and
 *     the element-offset identity (fold preserves every element's address);
 *   - orc_tensor_mult_d/_s: textbook contraction (numpy.einsum) on small
 *     random shapes, including the beta==0 overwrite rule (NaN in C is not
 *     read) and alpha==0 / k==0 / empty-m rules;
 *   - orc_mult_entries_d: every (l, m, n) entry of random canonical forms
 *     (batch labels, diagonals, FP32 and FP64 data) against numpy.einsum
 *     mapped through the canonical C strides, and beta==0 over NaN;
 *   - orc_blis_contract_d: numpy.einsum on random strided contractions
 *     (batch labels, several MC/KC/NR blocks, beta==0 over NaN) and the
 *     column limit of the bounded CPU sample;
 *   - orc_dpd_mult_d: block-wise vs full dense expansion (mode 2), which are
 *     two different algorithms for src/internal/3t/dpd/mult.cxx:29-74, and
 *     (through pyoracle.complex_dpd_mult, four block-wise calls) numpy's
 *     complex einsum over dense expansions built from orc_dpd_get_block.
 * tests/test_oracle.py holds each of these checks.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ==================================================================== */
/* Sorted label-set helpers: stl_ext intersection/exclusion/union        */
/* (src/external/stl_ext/include/algorithm.hpp:221-420).                 */
/* ==================================================================== */

static void lab_sort(char* v, int n)
{
    for (int i = 1; i < n; i++)
    {
        char x = v[i];
        int j = i - 1;
        while (j >= 0 && v[j] > x) { v[j + 1] = v[j]; j--; }
        v[j + 1] = x;
    }
}

static int lab_intersect(const char* a_, int na, const char* b_, int nb, char* out)
{
    char a[3 * ORC_MAXDIM], b[3 * ORC_MAXDIM];
    memcpy(a, a_, na); memcpy(b, b_, nb);
    lab_sort(a, na); lab_sort(b, nb);
    int i = 0, j = 0, n = 0;
    while (i < na && j < nb)
    {
        if (a[i] < b[j]) i++;
        else if (b[j] < a[i]) j++;
        else { out[n++] = a[i]; i++; j++; }
    }
    return n;
}

static int lab_exclude(const char* a_, int na, const char* b_, int nb, char* out)
{
    char a[3 * ORC_MAXDIM], b[3 * ORC_MAXDIM];
    memcpy(a, a_, na); memcpy(b, b_, nb);
    lab_sort(a, na); lab_sort(b, nb);
    int i = 0, j = 0, n = 0;
    while (i < na && j < nb)
    {
        if (a[i] < b[j]) out[n++] = a[i++];
        else if (b[j] < a[i]) j++;
        else { i++; j++; }
    }
    while (i < na) out[n++] = a[i++];
    return n;
}

static int lab_union(const char* a_, int na, const char* b_, int nb, char* out)
{
    char a[3 * ORC_MAXDIM], b[3 * ORC_MAXDIM];
    memcpy(a, a_, na); memcpy(b, b_, nb);
    lab_sort(a, na); lab_sort(b, nb);
    int i = 0, j = 0, n = 0;
    while (i < na && j < nb)
    {
        if (a[i] < b[j]) out[n++] = a[i++];
        else if (b[j] < a[i]) out[n++] = b[j++];
        else { out[n++] = a[i]; i++; j++; }
    }
    while (i < na) out[n++] = a[i++];
    while (j < nb) out[n++] = b[j++];
    return n;
}

/* select_from(v, s, match): v[i] for the first i with s[i]==m, per m in match
 * (algorithm.hpp:527-545). */
static void select_by_label(const int64_t* v, const char* s, int ns, const char* match, int nm,
                            int64_t* out)
{
    for (int k = 0; k < nm; k++)
        for (int i = 0; i < ns; i++)
            if (s[i] == match[k]) { out[k] = v[i]; break; }
}

/* ==================================================================== */
/* src/util/tensor.hpp                                                   */
/* ==================================================================== */

/* free_idx: smallest label not present (tensor.hpp:33-46). */
char orc_free_idx(int n, const char* labels)
{
    if (n == 0) return 0;
    char s[3 * ORC_MAXDIM + 2];
    memcpy(s, labels, n);
    lab_sort(s, n);
    if (s[0] > 0) return 0;
    for (int i = 1; i < n; i++)
        if (s[i] > s[i - 1] + 1) return (char)(s[i - 1] + 1);
    return (char)(s[n - 1] + 1);
}

/* sort_by_stride (tensor.hpp:95-148): key = min stride over tensors, ties
 * broken by each tensor's stride in argument order.  libstdc++'s std::sort on
 * at most 16 elements is an insertion sort, i.e. stable; so is this one. */
static int stride_less(int i, int j, int ntensor, const int64_t* const* s)
{
    int64_t mi = s[0][i], mj = s[0][j];
    for (int k = 1; k < ntensor; k++)
    {
        if (s[k][i] < mi) mi = s[k][i];
        if (s[k][j] < mj) mj = s[k][j];
    }
    if (mi < mj) return 1;
    if (mi > mj) return 0;
    for (int k = 0; k < ntensor; k++)
    {
        if (s[k][i] < s[k][j]) return 1;
        if (s[k][i] > s[k][j]) return 0;
    }
    return 0;
}

void orc_sort_by_stride(int n, int ntensor, const int64_t* const* strides, int* perm)
{
    for (int i = 0; i < n; i++) perm[i] = i;
    for (int i = 1; i < n; i++)
    {
        int x = perm[i], j = i - 1;
        while (j >= 0 && stride_less(x, perm[j], ntensor, strides)) { perm[j + 1] = perm[j]; j--; }
        perm[j + 1] = x;
    }
}

/* unit_dim (tensor.hpp:484-491). */
int orc_unit_dim(const int64_t* stride, const int* reorder, int n)
{
    for (int i = 0; i < n; i++)
        if (stride[reorder[i]] == 1) return i;
    return n;
}

/* diagonal (tensor.hpp:375-412): stable sort by label, drop length-1 dims,
 * a repeated label adds its stride to the first occurrence.  Returns the new
 * ndim, or -1 if a repeated label has a different length. */
int orc_diagonal(int ndim, const int64_t* len, const int64_t* stride, const char* idx,
                 int64_t* len_out, int64_t* stride_out, char* idx_out)
{
    int inds[ORC_MAXDIM];
    for (int i = 0; i < ndim; i++) inds[i] = i;
    for (int i = 1; i < ndim; i++)
    {
        int x = inds[i], j = i - 1;
        while (j >= 0 && idx[x] < idx[inds[j]]) { inds[j + 1] = inds[j]; j--; }
        inds[j + 1] = x;
    }
    int n = 0;
    for (int i = 0; i < ndim; i++)
    {
        int d = inds[i];
        if (i == 0 || idx[d] != idx[inds[i - 1]])
        {
            if (len[d] != 1)
            {
                len_out[n] = len[d];
                stride_out[n] = stride[d];
                idx_out[n] = idx[d];
                n++;
            }
        }
        else if (len[d] != 1)
        {
            if (n == 0 || len_out[n - 1] != len[d] || idx_out[n - 1] != idx[d]) return -1;
            stride_out[n - 1] += stride[d];
        }
    }
    return n;
}

/* fold (tensor.hpp:340-373): order dims by the FIRST tensor's strides and
 * merge dim inds[i] into the previous output dim when, for every tensor,
 * stride[inds[i]] == stride[inds[i-1]] * len[inds[i-1]]. */
int orc_fold(int ndim, int64_t* len, char* idx, int ntensor, int64_t* const* strides)
{
    int inds[ORC_MAXDIM];
    const int64_t* first[1] = {strides[0]};
    orc_sort_by_stride(ndim, 1, first, inds);

    int64_t olen[ORC_MAXDIM], ostr[4][ORC_MAXDIM];
    char oidx[ORC_MAXDIM];
    memcpy(olen, len, sizeof(int64_t) * ndim);
    memcpy(oidx, idx, ndim);
    for (int t = 0; t < ntensor; t++) memcpy(ostr[t], strides[t], sizeof(int64_t) * ndim);

    int n = 0;
    for (int i = 0; i < ndim; i++)
    {
        int contig = i != 0;
        for (int t = 0; contig && t < ntensor; t++)
            contig = ostr[t][inds[i]] == ostr[t][inds[i - 1]] * olen[inds[i - 1]];
        if (contig)
        {
            len[n - 1] *= olen[inds[i]];
        }
        else
        {
            idx[n] = oidx[inds[i]];
            len[n] = olen[inds[i]];
            for (int t = 0; t < ntensor; t++) strides[t][n] = ostr[t][inds[i]];
            n++;
        }
    }
    return n;
}

/* Front end of tblis_tensor_mult (src/iface/3t/mult.cxx:24-116). */
int orc_canonicalize(int ndim_A, const int64_t* len_A_, const int64_t* stride_A_, const char* idx_A_,
                     int ndim_B, const int64_t* len_B_, const int64_t* stride_B_, const char* idx_B_,
                     int ndim_C, const int64_t* len_C_, const int64_t* stride_C_, const char* idx_C_,
                     orc_groups* g)
{
    memset(g, 0, sizeof(*g));
    int64_t lA[ORC_MAXDIM + 2], sA[ORC_MAXDIM + 2], lB[ORC_MAXDIM + 2], sB[ORC_MAXDIM + 2];
    int64_t lC[ORC_MAXDIM + 2], sC[ORC_MAXDIM + 2];
    char iA[ORC_MAXDIM + 2], iB[ORC_MAXDIM + 2], iC[ORC_MAXDIM + 2];

    int nA = orc_diagonal(ndim_A, len_A_, stride_A_, idx_A_, lA, sA, iA);
    int nB = orc_diagonal(ndim_B, len_B_, stride_B_, idx_B_, lB, sB, iB);
    int nC = orc_diagonal(ndim_C, len_C_, stride_C_, idx_C_, lC, sC, iC);
    if (nA < 0 || nB < 0 || nC < 0) return 0;

    char t1[3 * ORC_MAXDIM], t2[3 * ORC_MAXDIM];
    int n1 = lab_intersect(iA, nA, iB, nB, t1);
    int nabc = lab_intersect(t1, n1, iC, nC, t2);

    /* mult.cxx:52-70: a tensor holding only ABC labels gets dummy dims */
    if (nA == nabc || nB == nabc || nC == nabc)
    {
        char u1[3 * ORC_MAXDIM + 2], u2[3 * ORC_MAXDIM + 2];
        int nu = lab_union(iA, nA, iB, nB, u1);
        nu = lab_union(u1, nu, iC, nC, u2);
        char d1 = orc_free_idx(nu, u2);
        lA[nA] = 1; sA[nA] = 0; iA[nA++] = d1;
        lC[nC] = 1; sC[nC] = 0; iC[nC++] = d1;
        nu = lab_union(iA, nA, iB, nB, u1);
        nu = lab_union(u1, nu, iC, nC, u2);
        char d2 = orc_free_idx(nu, u2);
        lB[nB] = 1; sB[nB] = 0; iB[nB++] = d2;
        lC[nC] = 1; sC[nC] = 0; iC[nC++] = d2;
    }

    char abc[ORC_MAXDIM], ab[ORC_MAXDIM], ac[ORC_MAXDIM], bc[ORC_MAXDIM];
    n1 = lab_intersect(iA, nA, iB, nB, t1);
    int n_abc = lab_intersect(t1, n1, iC, nC, abc);
    int n_ab = lab_exclude(t1, n1, abc, n_abc, ab);
    n1 = lab_intersect(iA, nA, iC, nC, t1);
    int n_ac = lab_exclude(t1, n1, abc, n_abc, ac);
    n1 = lab_intersect(iB, nB, iC, nC, t1);
    int n_bc = lab_exclude(t1, n1, abc, n_abc, bc);

    /* A-, B- and C-only labels are invalid (mult.cxx:98-104) */
    if (n_ab + n_ac + n_abc != nA || n_ab + n_bc + n_abc != nB || n_ac + n_bc + n_abc != nC)
        return 0;

    int64_t chk1[ORC_MAXDIM], chk2[ORC_MAXDIM];
    g->n_abc = n_abc; memcpy(g->idx_abc, abc, n_abc);
    select_by_label(lA, iA, nA, abc, n_abc, g->len_abc);
    select_by_label(lB, iB, nB, abc, n_abc, chk1);
    select_by_label(lC, iC, nC, abc, n_abc, chk2);
    for (int i = 0; i < n_abc; i++)
        if (chk1[i] != g->len_abc[i] || chk2[i] != g->len_abc[i]) return 0;
    select_by_label(sA, iA, nA, abc, n_abc, g->sA_abc);
    select_by_label(sB, iB, nB, abc, n_abc, g->sB_abc);
    select_by_label(sC, iC, nC, abc, n_abc, g->sC_abc);

    g->n_ab = n_ab; memcpy(g->idx_ab, ab, n_ab);
    select_by_label(lA, iA, nA, ab, n_ab, g->len_ab);
    select_by_label(lB, iB, nB, ab, n_ab, chk1);
    for (int i = 0; i < n_ab; i++) if (chk1[i] != g->len_ab[i]) return 0;
    select_by_label(sA, iA, nA, ab, n_ab, g->sA_ab);
    select_by_label(sB, iB, nB, ab, n_ab, g->sB_ab);

    g->n_ac = n_ac; memcpy(g->idx_ac, ac, n_ac);
    select_by_label(lA, iA, nA, ac, n_ac, g->len_ac);
    select_by_label(lC, iC, nC, ac, n_ac, chk1);
    for (int i = 0; i < n_ac; i++) if (chk1[i] != g->len_ac[i]) return 0;
    select_by_label(sA, iA, nA, ac, n_ac, g->sA_ac);
    select_by_label(sC, iC, nC, ac, n_ac, g->sC_ac);

    g->n_bc = n_bc; memcpy(g->idx_bc, bc, n_bc);
    select_by_label(lB, iB, nB, bc, n_bc, g->len_bc);
    select_by_label(lC, iC, nC, bc, n_bc, chk1);
    for (int i = 0; i < n_bc; i++) if (chk1[i] != g->len_bc[i]) return 0;
    select_by_label(sB, iB, nB, bc, n_bc, g->sB_bc);
    select_by_label(sC, iC, nC, bc, n_bc, g->sC_bc);

    /* mult.cxx:113-116 */
    int64_t* f_abc[3] = {g->sA_abc, g->sB_abc, g->sC_abc};
    g->n_abc = orc_fold(g->n_abc, g->len_abc, g->idx_abc, 3, f_abc);
    int64_t* f_ab[2] = {g->sA_ab, g->sB_ab};
    g->n_ab = orc_fold(g->n_ab, g->len_ab, g->idx_ab, 2, f_ab);
    int64_t* f_ac[2] = {g->sA_ac, g->sC_ac};
    g->n_ac = orc_fold(g->n_ac, g->len_ac, g->idx_ac, 2, f_ac);
    int64_t* f_bc[2] = {g->sB_bc, g->sC_bc};
    g->n_bc = orc_fold(g->n_bc, g->len_bc, g->idx_bc, 2, f_bc);

    g->valid = 1;
    return 1;
}

/* mult_blis reorder (src/internal/3t/dense/mult.cxx:212-263 for AB+AC+BC,
 * :479-556 for AB+AC+BC+ABC): stride sort per group, unit-stride detection,
 * pack_3d rotation; the outputs are the tensor_matrix (len, stride) lists. */
static void rotate_to_second(int* r, int n, int u)
{
    /* std::rotate(r+1, r+u, r+n) */
    int tmp[ORC_MAXDIM], k = 0;
    for (int i = u; i < n; i++) tmp[k++] = r[i];
    for (int i = 1; i < u; i++) tmp[k++] = r[i];
    for (int i = 1; i < n; i++) r[i] = tmp[i - 1];
}

void orc_blis_reorder(const orc_groups* g, orc_tmat* t)
{
    memset(t, 0, sizeof(*t));
    int rAC[ORC_MAXDIM], rBC[ORC_MAXDIM], rAB[ORC_MAXDIM], rABC[ORC_MAXDIM];
    const int64_t* s_ac[2] = {g->sC_ac, g->sA_ac};
    const int64_t* s_bc[2] = {g->sC_bc, g->sB_bc};
    const int64_t* s_ab[2] = {g->sA_ab, g->sB_ab};
    const int64_t* s_abc[3] = {g->sC_abc, g->sA_abc, g->sB_abc};
    orc_sort_by_stride(g->n_ac, 2, s_ac, rAC);
    orc_sort_by_stride(g->n_bc, 2, s_bc, rBC);
    orc_sort_by_stride(g->n_ab, 2, s_ab, rAB);
    orc_sort_by_stride(g->n_abc, 3, s_abc, rABC);

    int uA_AC = orc_unit_dim(g->sA_ac, rAC, g->n_ac);
    int uB_BC = orc_unit_dim(g->sB_bc, rBC, g->n_bc);
    int uA_AB = orc_unit_dim(g->sA_ab, rAB, g->n_ab);
    int uB_AB = orc_unit_dim(g->sB_ab, rAB, g->n_ab);

    t->pack_m_3d = uA_AC > 0 && uA_AC < g->n_ac;
    t->pack_n_3d = uB_BC > 0 && uB_BC < g->n_bc;
    t->pack_k_3d = (uA_AB > 0 && uA_AB < g->n_ab) || (uB_AB > 0 && uB_AB < g->n_ab);

    if (t->pack_m_3d) rotate_to_second(rAC, g->n_ac, uA_AC);
    if (t->pack_n_3d) rotate_to_second(rBC, g->n_bc, uB_BC);
    if (t->pack_k_3d) rotate_to_second(rAB, g->n_ab, uA_AB > uB_AB ? uA_AB : uB_AB);

    t->nm = g->n_ac; t->nn = g->n_bc; t->nk = g->n_ab; t->nl = g->n_abc;
    for (int i = 0; i < g->n_ac; i++)
    {
        t->len_m[i] = g->len_ac[rAC[i]];
        t->sA_m[i] = g->sA_ac[rAC[i]];
        t->sC_m[i] = g->sC_ac[rAC[i]];
    }
    for (int i = 0; i < g->n_bc; i++)
    {
        t->len_n[i] = g->len_bc[rBC[i]];
        t->sB_n[i] = g->sB_bc[rBC[i]];
        t->sC_n[i] = g->sC_bc[rBC[i]];
    }
    for (int i = 0; i < g->n_ab; i++)
    {
        t->len_k[i] = g->len_ab[rAB[i]];
        t->sA_k[i] = g->sA_ab[rAB[i]];
        t->sB_k[i] = g->sB_ab[rAB[i]];
    }
    for (int i = 0; i < g->n_abc; i++)
    {
        t->len_l[i] = g->len_abc[rABC[i]];
        t->sA_l[i] = g->sA_abc[rABC[i]];
        t->sB_l[i] = g->sB_abc[rABC[i]];
        t->sC_l[i] = g->sC_abc[rABC[i]];
    }
}

/* ==================================================================== */
/* src/matrix/block_scatter_matrix.hpp                                  */
/* ==================================================================== */

/* gcd/lcm (block_scatter_matrix.hpp:18-50): binary gcd. */
int64_t orc_gcd(int64_t a, int64_t b)
{
    a = a < 0 ? -a : a;
    b = b < 0 ? -b : b;
    if (a == 0) return b;
    if (b == 0) return a;
    unsigned d = __builtin_ctzl((unsigned long)(a | b));
    a >>= __builtin_ctzl((unsigned long)a);
    b >>= __builtin_ctzl((unsigned long)b);
    while (a != b)
    {
        if (a > b) a = (a - b) >> 1;
        else b = (b - a) >> 1;
    }
    return a << d;
}

/* lcm = a*b/gcd (block_scatter_matrix.hpp:46-49).  The reference's binary
 * gcd above shifts once per subtraction, so it returns 0 for some pairs
 * (e.g. gcd(5, 8)) and the reference then divides by zero; 0 is returned
 * here instead and the pack_3d fill aborts on it. */
int64_t orc_lcm(int64_t a, int64_t b)
{
    int64_t g = orc_gcd(a, b);
    return g ? a * b / g : 0;
}

/* odometer position -> offset (viterator::position, viterator.hpp:106-137) */
static int64_t odometer_offset(int n, const int64_t* len, const int64_t* stride, int64_t pos,
                               int64_t* digits)
{
    int64_t off = 0;
    for (int i = 0; i < n; i++)
    {
        digits[i] = pos % len[i];
        pos /= len[i];
        off += digits[i] * stride[i];
    }
    return off;
}

/* viterator::next after the first call (viterator.hpp:43-80): returns 0 when
 * the iteration wraps. */
static int odometer_next(int n, const int64_t* len, const int64_t* stride, int64_t* digits,
                         int64_t* off)
{
    for (int i = 0; i < n; i++)
    {
        if (digits[i] == len[i] - 1)
        {
            *off -= digits[i] * stride[i];
            digits[i] = 0;
            if (i == n - 1) return 0;
        }
        else
        {
            *off += stride[i];
            digits[i]++;
            return 1;
        }
    }
    return 0;
}

/* fill_scatter (block_scatter_matrix.hpp:68-230). */
void orc_fill_scatter(int ndim, const int64_t* len_, const int64_t* stride_, int64_t BS,
                      int64_t off, int64_t size, int64_t* scat, int pack_3d, int elem_size)
{
    if (size == 0) return;
    const int64_t CL = 64 / elem_size;
    if (ndim == 0) { scat[0] = 0; return; }

    int64_t m0 = len_[0], s0 = stride_[0];
    const int64_t* len = len_ + 1;
    const int64_t* stride = stride_ + 1;
    int n = ndim - 1;
    int empty = 0;
    for (int i = 0; i < n; i++) if (len[i] == 0) empty = 1;
    int64_t digits[ORC_MAXDIM];

    if (!pack_3d)
    {
        if (empty) return;
        int64_t p0 = off % m0;
        int64_t pos = odometer_offset(n, len, stride, off / m0, digits);
        int first = 1;
        for (int64_t idx = 0; idx < size;)
        {
            if (!first && !odometer_next(n, len, stride, digits, &pos)) break;
            first = 0;
            int64_t pos2 = pos + p0 * s0;
            int64_t imax = m0 - p0 < size - idx ? m0 - p0 : size - idx;
            for (int64_t i = 0; i < imax; i++) { scat[idx++] = pos2; pos2 += s0; }
            p0 = 0;
        }
        return;
    }

    int64_t m1 = len[0], s1 = stride[0];
    len++; stride++; n--;
    empty = 0;
    for (int i = 0; i < n; i++) if (len[i] == 0) empty = 1;

    int64_t r0m = m0 - m0 % BS; if (r0m < BS) r0m = BS;
    int64_t r1m = m1 - m1 % BS; if (r1m < BS) r1m = BS;
    int64_t lc = orc_lcm(BS, CL);
    if (lc == 0) abort(); /* the reference divides by zero here */
    int64_t BS0 = lc < r0m ? lc : r0m;
    int64_t BS1 = CL < r1m ? CL : r1m;

    int64_t p01 = off % (m0 * m1);
    int64_t p0 = p01 % m0;
    int64_t p1 = p01 / m0;
    int64_t q01 = (off + size - 1) % (m0 * m1);
    int64_t q0 = q01 % m0;
    int64_t q1 = q01 / m0;
    int64_t r0 = q0 - q0 % BS0;
    int64_t b_min = off / (m0 * m1);
    int64_t b_max = (off + size - 1) / (m0 * m1) + 1;

    if (p0 == 0) p1--;
    if (q0 == m0 - 1) q1++;

    if (empty) return;
    int64_t pos0 = odometer_offset(n, len, stride, b_min, digits);
    int64_t digits0[ORC_MAXDIM];
    memcpy(digits0, digits, sizeof(digits));

    int64_t idx = 0;
    if (r0 > 0)
    {
        int64_t pos = pos0;
        int first = 1;
        for (int64_t b = b_min; b < b_max; b++)
        {
            if (!first && !odometer_next(n, len, stride, digits, &pos)) break;
            first = 0;
            int64_t min1 = b == b_min ? p1 + 1 : 0;
            int64_t max1 = b == b_max - 1 ? q1 : m1;
            for (int64_t i1 = min1; i1 < max1; i1 += BS1)
                for (int64_t i0 = 0; i0 < r0; i0 += BS0)
                {
                    int64_t e1 = BS1 < max1 - i1 ? BS1 : max1 - i1;
                    for (int64_t j1 = 0; j1 < e1; j1++)
                        for (int64_t j0 = 0; j0 < BS0; j0++)
                            scat[idx++] = pos + (i0 + j0) * s0 + (i1 + j1) * s1;
                }
        }
        memcpy(digits, digits0, sizeof(digits));
    }

    int64_t pos = pos0;
    int first = 1;
    for (int64_t b = b_min; b < b_max; b++)
    {
        if (!first && !odometer_next(n, len, stride, digits, &pos)) break;
        first = 0;
        int64_t min1 = b == b_min ? p1 + 1 : 0;
        int64_t max1 = b == b_max - 1 ? q1 : m1;

        for (int64_t j1 = min1; j1 < max1; j1++)
            for (int64_t j0 = r0; j0 < m0; j0++)
                scat[idx++] = pos + j0 * s0 + j1 * s1;

        int do_p = b == b_min && p0 > 0;
        int do_q = b == b_max - 1 && q0 < m0 - 1;

        if (do_p && do_q && p1 == q1)
        {
            for (int64_t j0 = p0; j0 <= q0; j0++) scat[idx++] = pos + j0 * s0 + p1 * s1;
        }
        else
        {
            if (do_p)
                for (int64_t j0 = p0; j0 < m0; j0++) scat[idx++] = pos + j0 * s0 + p1 * s1;
            if (do_q)
                for (int64_t j0 = 0; j0 <= q0; j0++) scat[idx++] = pos + j0 * s0 + q1 * s1;
        }
    }
}

/* fill_block_stride (block_scatter_matrix.hpp:232-247). */
void orc_fill_block_stride(int64_t BS, int64_t size, const int64_t* scat, int64_t* bs)
{
    if (size == 0) return;
    for (int64_t i = 0; i < size; i += BS)
    {
        int64_t bl = size - i < BS ? size - i : BS;
        int64_t s = bl > 1 ? scat[i + 1] - scat[i] : 1;
        for (int64_t j = i + 1; j + 1 < i + bl; j++)
            if (scat[j + 1] - scat[j] != s) s = 0;
        bs[i] = s;
    }
}

/* ==================================================================== */
/* Reference contraction: mult_ref (src/internal/3t/dense/mult.cxx:653-720) */
/* with the early-out rules of :777-816 and the front end's alpha==0 rule  */
/* (src/iface/3t/mult.cxx:133-148).                                       */
/* ==================================================================== */

static int64_t prod(const int64_t* v, int n)
{
    int64_t p = 1;
    for (int i = 0; i < n; i++) p *= v[i];
    return p;
}

/* offset of flat index f over (len, stride) lists, first dim fastest */
static int64_t flat_offset(int64_t f, const int64_t* len, const int64_t* stride, int n)
{
    int64_t off = 0;
    for (int i = 0; i < n; i++)
    {
        off += (f % len[i]) * stride[i];
        f /= len[i];
    }
    return off;
}

typedef struct
{
    int64_t nl, nm, nn, nk;
    int64_t *oAl, *oBl, *oCl, *oAm, *oCm, *oBn, *oCn, *oAk, *oBk;
} flat_plan;

static void make_flat_plan(const orc_groups* g, flat_plan* p)
{
    p->nl = prod(g->len_abc, g->n_abc);
    p->nm = prod(g->len_ac, g->n_ac);
    p->nn = prod(g->len_bc, g->n_bc);
    p->nk = prod(g->len_ab, g->n_ab);
    int64_t nl = p->nl > 0 ? p->nl : 1, nm = p->nm > 0 ? p->nm : 1;
    int64_t nn = p->nn > 0 ? p->nn : 1, nk = p->nk > 0 ? p->nk : 1;
    p->oAl = malloc(sizeof(int64_t) * nl); p->oBl = malloc(sizeof(int64_t) * nl);
    p->oCl = malloc(sizeof(int64_t) * nl);
    p->oAm = malloc(sizeof(int64_t) * nm); p->oCm = malloc(sizeof(int64_t) * nm);
    p->oBn = malloc(sizeof(int64_t) * nn); p->oCn = malloc(sizeof(int64_t) * nn);
    p->oAk = malloc(sizeof(int64_t) * nk); p->oBk = malloc(sizeof(int64_t) * nk);
    for (int64_t i = 0; i < p->nl; i++)
    {
        p->oAl[i] = flat_offset(i, g->len_abc, g->sA_abc, g->n_abc);
        p->oBl[i] = flat_offset(i, g->len_abc, g->sB_abc, g->n_abc);
        p->oCl[i] = flat_offset(i, g->len_abc, g->sC_abc, g->n_abc);
    }
    for (int64_t i = 0; i < p->nm; i++)
    {
        p->oAm[i] = flat_offset(i, g->len_ac, g->sA_ac, g->n_ac);
        p->oCm[i] = flat_offset(i, g->len_ac, g->sC_ac, g->n_ac);
    }
    for (int64_t i = 0; i < p->nn; i++)
    {
        p->oBn[i] = flat_offset(i, g->len_bc, g->sB_bc, g->n_bc);
        p->oCn[i] = flat_offset(i, g->len_bc, g->sC_bc, g->n_bc);
    }
    for (int64_t i = 0; i < p->nk; i++)
    {
        p->oAk[i] = flat_offset(i, g->len_ab, g->sA_ab, g->n_ab);
        p->oBk[i] = flat_offset(i, g->len_ab, g->sB_ab, g->n_ab);
    }
}

static void free_flat_plan(flat_plan* p)
{
    free(p->oAl); free(p->oBl); free(p->oCl); free(p->oAm); free(p->oCm);
    free(p->oBn); free(p->oCn); free(p->oAk); free(p->oBk);
}

static void contract_flat(int acc, const flat_plan* p, double alpha, const void* A, const void* B,
                          double beta, const void* Cin, double* Cout, int is_float,
                          int out_is_float)
{
    const double* Ad = A; const double* Bd = B;
    const float* Af = A; const float* Bf = B;
    #pragma omp parallel for collapse(2) schedule(static)
    for (int64_t l = 0; l < p->nl; l++)
        for (int64_t m = 0; m < p->nm; m++)
        {
            for (int64_t n = 0; n < p->nn; n++)
            {
                int64_t ca = p->oAl[l] + p->oAm[m], cb = p->oBl[l] + p->oBn[n];
                int64_t cc = p->oCl[l] + p->oCm[m] + p->oCn[n];
                double r;
                if (acc)
                {
                    long double s = 0;
                    for (int64_t k = 0; k < p->nk; k++)
                    {
                        long double a = is_float ? Af[ca + p->oAk[k]] : Ad[ca + p->oAk[k]];
                        long double b = is_float ? Bf[cb + p->oBk[k]] : Bd[cb + p->oBk[k]];
                        s += a * b;
                    }
                    long double t = (long double)alpha * s;
                    if (beta != 0.0)
                    {
                        long double c0 = out_is_float ? ((const float*)Cin)[cc] : ((const double*)Cin)[cc];
                        t += (long double)beta * c0;
                    }
                    r = (double)t;
                }
                else
                {
                    double s = 0;
                    for (int64_t k = 0; k < p->nk; k++)
                    {
                        double a = is_float ? Af[ca + p->oAk[k]] : Ad[ca + p->oAk[k]];
                        double b = is_float ? Bf[cb + p->oBk[k]] : Bd[cb + p->oBk[k]];
                        s += a * b;
                    }
                    double t = alpha * s;
                    if (beta != 0.0)
                    {
                        double c0 = out_is_float ? ((const float*)Cin)[cc] : ((const double*)Cin)[cc];
                        t += beta * c0;
                    }
                    r = t;
                }
                Cout[cc] = r;
            }
        }
}

/* C <- beta*C over the C part of a canonical form (set/scale paths). */
static void scale_c(const orc_groups* g, double beta, const void* Cin, double* Cout,
                    int out_is_float)
{
    int64_t lens[3 * ORC_MAXDIM], strides[3 * ORC_MAXDIM];
    int n = 0;
    for (int i = 0; i < g->n_ac; i++) { lens[n] = g->len_ac[i]; strides[n++] = g->sC_ac[i]; }
    for (int i = 0; i < g->n_bc; i++) { lens[n] = g->len_bc[i]; strides[n++] = g->sC_bc[i]; }
    for (int i = 0; i < g->n_abc; i++) { lens[n] = g->len_abc[i]; strides[n++] = g->sC_abc[i]; }
    int64_t tot = prod(lens, n);
    for (int64_t f = 0; f < tot; f++)
    {
        int64_t o = flat_offset(f, lens, strides, n);
        double c0 = out_is_float ? ((const float*)Cin)[o] : ((const double*)Cin)[o];
        Cout[o] = beta == 0.0 ? 0.0 : beta * c0;
    }
}

static void tensor_mult_any(int acc, double alpha, int ndim_A, const int64_t* len_A,
                            const int64_t* stride_A, const char* idx_A, const void* A,
                            int ndim_B, const int64_t* len_B, const int64_t* stride_B,
                            const char* idx_B, const void* B,
                            double beta, int ndim_C, const int64_t* len_C, const int64_t* stride_C,
                            const char* idx_C, const void* Cin, double* Cout, int is_float)
{
    orc_groups g;
    if (!orc_canonicalize(ndim_A, len_A, stride_A, idx_A, ndim_B, len_B, stride_B, idx_B,
                          ndim_C, len_C, stride_C, idx_C, &g))
        abort();

    if (alpha == 0.0)
    {
        if (beta != 1.0) scale_c(&g, beta, Cin, Cout, is_float);
        return;
    }
    int64_t nac = prod(g.len_ac, g.n_ac), nbc = prod(g.len_bc, g.n_bc);
    int64_t nabc = prod(g.len_abc, g.n_abc), nab = prod(g.len_ab, g.n_ab);
    if (nac == 0 || nbc == 0 || nabc == 0) return;
    if (nab == 0)
    {
        if (beta != 1.0) scale_c(&g, beta, Cin, Cout, is_float);
        return;
    }
    flat_plan p;
    make_flat_plan(&g, &p);
    contract_flat(acc, &p, alpha, A, B, beta, Cin, Cout, is_float, is_float);
    free_flat_plan(&p);
}

void orc_tensor_mult_d(int acc, double alpha, int ndim_A, const int64_t* len_A,
                       const int64_t* stride_A, const char* idx_A, const double* A,
                       int ndim_B, const int64_t* len_B, const int64_t* stride_B,
                       const char* idx_B, const double* B,
                       double beta, int ndim_C, const int64_t* len_C, const int64_t* stride_C,
                       const char* idx_C, double* C)
{
    tensor_mult_any(acc, alpha, ndim_A, len_A, stride_A, idx_A, A, ndim_B, len_B, stride_B, idx_B, B,
                    beta, ndim_C, len_C, stride_C, idx_C, C, C, 0);
}

void orc_tensor_mult_s(double alpha, int ndim_A, const int64_t* len_A, const int64_t* stride_A,
                       const char* idx_A, const float* A,
                       int ndim_B, const int64_t* len_B, const int64_t* stride_B,
                       const char* idx_B, const float* B,
                       double beta, int ndim_C, const int64_t* len_C, const int64_t* stride_C,
                       const char* idx_C, double* C_out, const float* C_in)
{
    tensor_mult_any(1, alpha, ndim_A, len_A, stride_A, idx_A, A, ndim_B, len_B, stride_B, idx_B, B,
                    beta, ndim_C, len_C, stride_C, idx_C, C_in, C_out, 1);
}

void orc_mult_entries_d(const orc_groups* g, double alpha, const void* A, const void* B,
                        double beta, const void* C0, int is_float, int64_t nent,
                        const int64_t* ent_l, const int64_t* ent_m, const int64_t* ent_n,
                        double* out)
{
    int64_t nk = prod(g->len_ab, g->n_ab);
    int64_t* oAk = malloc(sizeof(int64_t) * (nk ? nk : 1));
    int64_t* oBk = malloc(sizeof(int64_t) * (nk ? nk : 1));
    for (int64_t k = 0; k < nk; k++)
    {
        oAk[k] = flat_offset(k, g->len_ab, g->sA_ab, g->n_ab);
        oBk[k] = flat_offset(k, g->len_ab, g->sB_ab, g->n_ab);
    }
    #pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < nent; e++)
    {
        int64_t ca = flat_offset(ent_l[e], g->len_abc, g->sA_abc, g->n_abc) +
                     flat_offset(ent_m[e], g->len_ac, g->sA_ac, g->n_ac);
        int64_t cb = flat_offset(ent_l[e], g->len_abc, g->sB_abc, g->n_abc) +
                     flat_offset(ent_n[e], g->len_bc, g->sB_bc, g->n_bc);
        int64_t cc = flat_offset(ent_l[e], g->len_abc, g->sC_abc, g->n_abc) +
                     flat_offset(ent_m[e], g->len_ac, g->sC_ac, g->n_ac) +
                     flat_offset(ent_n[e], g->len_bc, g->sC_bc, g->n_bc);
        long double s = 0;
        for (int64_t k = 0; k < nk; k++)
        {
            long double a = is_float ? ((const float*)A)[ca + oAk[k]] : ((const double*)A)[ca + oAk[k]];
            long double b = is_float ? ((const float*)B)[cb + oBk[k]] : ((const double*)B)[cb + oBk[k]];
            s += a * b;
        }
        long double t = (long double)alpha * s;
        if (beta != 0.0)
            t += (long double)beta * (is_float ? ((const float*)C0)[cc] : ((const double*)C0)[cc]);
        out[e] = (double)t;
    }
    free(oAk); free(oBk);
}

/* ==================================================================== */
/* CPU baseline: blocked GEMM over scatter vectors.  Loop nest of         */
/* TensorGEMM (src/nodes/gemm.hpp:178-187: nc -> kc -> pack B -> mc ->     */
/* pack A -> micro-kernel), packing as pack_nb/pack_sb                     */
/* (src/kernels/3m/packm.hpp:293-404), micro-kernel semantics of           */
/* gemm_ukr_def (src/kernels/3m/gemm.hpp:30-67).                           */
/* ==================================================================== */

#define BL_MR 6
#define BL_NR 8
#define BL_MC 72
#define BL_KC 256
#define BL_NC 4080

#ifdef __AVX2__
#include <immintrin.h>
#endif

/* ab (MR x NR, row major) = sum_p a[p*MR + i] * b[p*NR + j]: the register
 * tile of the reference's haswell_d6x8 configuration (MR=6, NR=8,
 * src/configs/haswell/config.hpp:82-100), written here with AVX2 FMA
 * intrinsics: 12 accumulators of 4 doubles, one broadcast of A and two
 * loads of B per row step. */
static void bl_ukr(int64_t k, const double* restrict a, const double* restrict b, double* restrict ab)
{
#ifdef __AVX2__
    __m256d c[BL_MR][2];
    for (int i = 0; i < BL_MR; i++) c[i][0] = c[i][1] = _mm256_setzero_pd();
    for (int64_t p = 0; p < k; p++)
    {
        const __m256d b0 = _mm256_load_pd(b + p * BL_NR), b1 = _mm256_load_pd(b + p * BL_NR + 4);
        for (int i = 0; i < BL_MR; i++)
        {
            const __m256d ai = _mm256_broadcast_sd(a + p * BL_MR + i);
            c[i][0] = _mm256_fmadd_pd(ai, b0, c[i][0]);
            c[i][1] = _mm256_fmadd_pd(ai, b1, c[i][1]);
        }
    }
    for (int i = 0; i < BL_MR; i++)
    {
        _mm256_storeu_pd(ab + i * BL_NR, c[i][0]);
        _mm256_storeu_pd(ab + i * BL_NR + 4, c[i][1]);
    }
#else
    double acc[BL_MR][BL_NR];
    for (int i = 0; i < BL_MR; i++)
        for (int j = 0; j < BL_NR; j++) acc[i][j] = 0;
    for (int64_t p = 0; p < k; p++)
        for (int i = 0; i < BL_MR; i++)
        {
            double ai = a[p * BL_MR + i];
            for (int j = 0; j < BL_NR; j++) acc[i][j] += ai * b[p * BL_NR + j];
        }
    for (int i = 0; i < BL_MR; i++)
        for (int j = 0; j < BL_NR; j++) ab[i * BL_NR + j] = acc[i][j];
#endif
}

/* Thread split of the GEMM loops (make_gemm_thread_config,
 * src/util/gemm_thread.hpp:22-57, with the default ratios of
 * src/configs/config_builder.hpp:111-117: m counts twice n, at most 3
 * threads on the jr loop and 1 on the ir loop): each prime factor of the
 * thread count goes to the ic loop unless the n work per thread is at least
 * the m work (tci_partition_2x2, src/external/tci/tci/parallel.c:280-325);
 * a factor 3 or 2 of the n part then moves to the jr loop. */
static void bl_threads(int nt, int64_t m, int64_t n, int* ic_nt, int* jc_nt, int* jr_nt)
{
    uint64_t w1 = (uint64_t)m * 2, w2 = (uint64_t)n;
    int n1 = 1, n2 = 1;
    if (nt < 4)
    {
        if (w1 >= w2) n1 = nt; else n2 = nt;
    }
    else
    {
        int x = nt;
        for (int f = 2; x > 1; )
        {
            if (x % f) { f++; continue; }
            x /= f;
            if (w2 >= w1) { w2 /= (uint64_t)f; n2 *= f; }
            else { w1 /= (uint64_t)f; n1 *= f; }
        }
    }
    int jr = 1;
    for (int r = 3; r > 1; r--)
        if (n2 % r == 0) { jr = r; n2 /= r; break; }
    *ic_nt = n1;
    *jc_nt = n2;
    *jr_nt = jr;
}

/* The loop nest of TensorGEMM (src/nodes/gemm.hpp:178-187): NC columns ->
 * KC of k -> pack B (KC x NC, NR-wide panels, from the scatter vectors as
 * pack_nb/pack_sb do, src/kernels/3m/packm.hpp:293-404) -> MC rows -> pack
 * A (MC x KC, MR-wide panels) -> micro-kernel over the NR x MR grid, C
 * updated through its scatter vectors (beta == 0 writes without reading,
 * src/kernels/3m/gemm.hpp:54-58; beta becomes 1 after the first KC block,
 * src/nodes/partm.hpp:83).  Threads are split over the jc, ic and jr loops
 * as bl_threads gives; all of them pack the shared B panel, each packs the
 * A blocks of its ic range. */
double orc_blis_contract_d(int nthreads, double alpha, int ndim_A, const int64_t* len_A,
                           const int64_t* stride_A, const char* idx_A, const double* A,
                           int ndim_B, const int64_t* len_B, const int64_t* stride_B,
                           const char* idx_B, const double* B,
                           double beta, int ndim_C, const int64_t* len_C, const int64_t* stride_C,
                           const char* idx_C, double* C, int64_t n_limit)
{
    orc_groups g;
    orc_tmat t;
    if (!orc_canonicalize(ndim_A, len_A, stride_A, idx_A, ndim_B, len_B, stride_B, idx_B,
                          ndim_C, len_C, stride_C, idx_C, &g))
        abort();
    orc_blis_reorder(&g, &t);
    int64_t m = prod(t.len_m, t.nm), n = prod(t.len_n, t.nn), k = prod(t.len_k, t.nk);
    int64_t L = prod(t.len_l, t.nl);
    if (n_limit > 0 && n_limit < n) n = n_limit;
    int64_t *rA = malloc(8 * (m + 1)), *cA = malloc(8 * (k + 1)), *rB = malloc(8 * (k + 1));
    int64_t *cB = malloc(8 * (n + 1)), *rC = malloc(8 * (m + 1)), *cC = malloc(8 * (n + 1));
    orc_fill_scatter(t.nm, t.len_m, t.sA_m, BL_MR, 0, m, rA, t.pack_m_3d, 8);
    orc_fill_scatter(t.nk, t.len_k, t.sA_k, BL_KC, 0, k, cA, t.pack_k_3d, 8);
    orc_fill_scatter(t.nk, t.len_k, t.sB_k, BL_KC, 0, k, rB, t.pack_k_3d, 8);
    orc_fill_scatter(t.nn, t.len_n, t.sB_n, BL_NR, 0, n, cB, t.pack_n_3d, 8);
    orc_fill_scatter(t.nm, t.len_m, t.sC_m, BL_MR, 0, m, rC, t.pack_m_3d, 8);
    orc_fill_scatter(t.nn, t.len_n, t.sC_n, BL_NR, 0, n, cC, t.pack_n_3d, 8);
    int nt = 1;
#ifdef _OPENMP
    nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
    int ic_nt, jc_nt, jr_nt;
    bl_threads(nt, m, n, &ic_nt, &jc_nt, &jr_nt);
    const int64_t ncp_max = ((n < BL_NC ? n : BL_NC) + BL_NR - 1) / BL_NR * BL_NR;
    double* bp = aligned_alloc(64, sizeof(double) * (size_t)(ncp_max * BL_KC + 8));
    const int64_t nic = (m + BL_MC - 1) / BL_MC;

    for (int64_t l = 0; l < L && m > 0 && n > 0; l++)
    {
        const int64_t oA = flat_offset(l, t.len_l, t.sA_l, t.nl);
        const int64_t oB = flat_offset(l, t.len_l, t.sB_l, t.nl);
        const int64_t oC = flat_offset(l, t.len_l, t.sC_l, t.nl);
        for (int64_t jc = 0; jc < n; jc += BL_NC)
        {
            const int64_t nc = n - jc < BL_NC ? n - jc : BL_NC;
            const int64_t ncp = (nc + BL_NR - 1) / BL_NR * BL_NR;
            const int64_t npan = ncp / BL_NR;
            for (int64_t pc = 0; pc < k || (k == 0 && pc == 0); pc += BL_KC)
            {
                const int64_t kc = k - pc < BL_KC ? k - pc : BL_KC;
                const double beta_l = pc == 0 ? beta : 1.0;
                #pragma omp parallel num_threads(nt)
                {
                    #pragma omp for schedule(static)
                    for (int64_t jr = 0; jr < ncp; jr += BL_NR)
                        for (int64_t p = 0; p < kc; p++)
                            for (int j = 0; j < BL_NR; j++)
                                bp[jr * kc + p * BL_NR + j] = jr + j < nc ?
                                    B[oB + rB[pc + p] + cB[jc + jr + j]] : 0.0;
                    int tid = 0;
#ifdef _OPENMP
                    tid = omp_get_thread_num();
#endif
                    /* this thread's column range (jc and jr parts, whole
                     * NR panels) and MC blocks (ic part) */
                    const int wc = (tid / (ic_nt * jr_nt)) * jr_nt + tid % jr_nt, ncol = jc_nt * jr_nt;
                    const int ig = (tid / jr_nt) % ic_nt;
                    const int64_t p0 = npan * wc / ncol, p1 = npan * (wc + 1) / ncol;
                    const int64_t b0 = nic * ig / ic_nt, b1 = nic * (ig + 1) / ic_nt;
                    double ap[BL_MC * BL_KC] __attribute__((aligned(64)));
                    double ab[BL_MR * BL_NR] __attribute__((aligned(64)));
                    for (int64_t ib = b0; ib < b1 && p0 < p1; ib++)
                    {
                        const int64_t ic = ib * BL_MC;
                        const int64_t mc = m - ic < BL_MC ? m - ic : BL_MC;
                        for (int64_t ir = 0; ir < mc; ir += BL_MR)
                            for (int64_t p = 0; p < kc; p++)
                                for (int i = 0; i < BL_MR; i++)
                                    ap[ir * kc + p * BL_MR + i] = ir + i < mc ?
                                        A[oA + rA[ic + ir + i] + cA[pc + p]] : 0.0;
                        for (int64_t jr = p0 * BL_NR; jr < p1 * BL_NR && jr < nc; jr += BL_NR)
                            for (int64_t ir = 0; ir < mc; ir += BL_MR)
                            {
                                bl_ukr(kc, ap + ir * kc, bp + jr * kc, ab);
                                for (int i = 0; i < BL_MR && ir + i < mc; i++)
                                {
                                    double* crow = C + oC + rC[ic + ir + i];
                                    for (int j = 0; j < BL_NR && jr + j < nc; j++)
                                    {
                                        double* c = crow + cC[jc + jr + j];
                                        *c = beta_l == 0.0 ? alpha * ab[i * BL_NR + j]
                                                           : alpha * ab[i * BL_NR + j] + beta_l * *c;
                                    }
                                }
                            }
                    }
                }
            }
        }
    }
    free(bp);
    free(rA); free(cA); free(rB); free(cB); free(rC); free(cC);
    return 2.0 * (double)m * (double)n * (double)k * (double)L;
}

/* ==================================================================== */
/* DPD storage: src/external/marray/include/dpd_marray_base.hpp:31-228    */
/* ==================================================================== */

/* default_depth (dpd_marray_base.hpp:33-63). layout: 0/1 balanced, 2/3
 * blocked, 4/5 prefix; odd = row major. */
static void dpd_default_depth(int layout, int ndim, int* depth)
{
    if (layout == 0 || layout == 1)
    {
        unsigned dl = sizeof(unsigned) * 8 - __builtin_clz((unsigned)ndim) - 1;
        unsigned du = dl + ((ndim & (ndim - 1)) ? 1 : 0);
        unsigned o = ndim - (1u << dl);
        unsigned k = 2 * ((o + 1) / 2);
        unsigned l = 2 * (o / 2);
        unsigned h = (ndim + 1) / 2;
        for (unsigned i = 0; i < k; i++) depth[i] = du;
        for (unsigned i = k; i < h; i++) depth[i] = dl;
        for (unsigned i = h; i < h + l; i++) depth[i] = du;
        for (unsigned i = h + l; i < (unsigned)ndim; i++) depth[i] = dl;
    }
    else if (layout == 4 || layout == 5)
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

/* set_tree (dpd_marray_base.hpp:65-115). */
static void dpd_set_tree(orc_dpd* d)
{
    int ndim = d->ndim;
    int depth[ORC_MAXDIM], node[ORC_MAXDIM], leaf_idx[ORC_MAXDIM];
    int n = ndim;
    for (int i = 0; i < ndim; i++) { depth[i] = d->depth[i]; node[i] = 0; leaf_idx[i] = i; }
    int pos = 0;
    for (int dd = ndim; dd-- > 0;)
    {
        for (int i = 0; i < n; i++)
        {
            if (depth[i] == dd + 1)
            {
                d->parent[node[i]] = d->parent[node[i + 1]] = pos;
                for (int j = i + 1; j + 1 < n; j++)
                {
                    depth[j] = depth[j + 1];
                    node[j] = node[j + 1];
                    leaf_idx[j] = leaf_idx[j + 1];
                }
                n--;
                depth[i]--;
                node[i] = pos++;
                leaf_idx[i] = -1;
            }
            else if (depth[i] == dd)
            {
                node[i] = pos++;
                if (leaf_idx[i] != -1) d->leaf[leaf_idx[i]] = node[i];
            }
        }
    }
}

/* set_size (dpd_marray_base.hpp:117-160). */
static void dpd_set_size(orc_dpd* d)
{
    int ndim = d->ndim, nirrep = d->nirrep;
    for (int i = 0; i < ndim; i++)
        for (int j = 0; j < 8; j++) d->ext[i][j] = d->len[i][j];
    if (!d->row_major)
    {
        for (int i = 0; i < ndim; i++)
        {
            for (int j = 0; j < nirrep; j++) d->size[d->leaf[i]][j] = d->ext[i][j];
            d->perm[i] = i;
        }
    }
    else
    {
        for (int i = 0; i < ndim; i++)
        {
            for (int j = 0; j < nirrep; j++) d->size[d->leaf[ndim - 1 - i]][j] = d->ext[i][j];
            d->perm[i] = ndim - 1 - i;
        }
        for (int i = 0; i < ndim / 2; i++)
            for (int j = 0; j < nirrep; j++)
            {
                int64_t t = d->ext[i][j];
                d->ext[i][j] = d->ext[ndim - 1 - i][j];
                d->ext[ndim - 1 - i][j] = t;
            }
    }
    for (int i = 0; i < ndim - 1; i++)
    {
        int next = d->parent[2 * i];
        for (int irr1 = 0; irr1 < nirrep; irr1++)
        {
            d->size[next][irr1] = 0;
            for (int irr2 = 0; irr2 < nirrep; irr2++)
                d->size[next][irr1] += d->size[2 * i][irr1 ^ irr2] * d->size[2 * i + 1][irr2];
        }
    }
}

void orc_dpd_init(orc_dpd* d, int irrep, int nirrep, int ndim, const int64_t* len, int layout)
{
    memset(d, 0, sizeof(*d));
    d->ndim = ndim; d->nirrep = nirrep; d->irrep = irrep;
    d->row_major = layout & 1;
    for (int i = 0; i < ndim; i++)
        for (int j = 0; j < nirrep; j++) d->len[i][j] = len[i * nirrep + j];
    dpd_default_depth(layout, ndim, d->depth);
    dpd_set_tree(d);
    dpd_set_size(d);
}

int64_t orc_dpd_size(const orc_dpd* d)
{
    if (d->ndim == 1) return d->len[0][d->irrep];
    return d->size[2 * d->ndim - 2][d->irrep];
}

/* get_block (dpd_marray_base.hpp:162-228), for a freshly laid out tensor
 * (per-irrep offsets 0 and leaps 1). */
void orc_dpd_get_block(const orc_dpd* d, const int* irreps, int64_t* len, int64_t* offset,
                       int64_t* stride)
{
    int ndim = d->ndim;
    int dirrep[2 * ORC_MAXDIM];
    int64_t dstride[2 * ORC_MAXDIM];
    dstride[2 * ndim - 2] = 1;
    for (int i = 0; i < ndim; i++) dirrep[d->leaf[d->perm[i]]] = irreps[i];
    for (int i = 0; i < ndim - 1; i++) dirrep[d->parent[2 * i]] = dirrep[2 * i] ^ dirrep[2 * i + 1];
    int64_t data = 0;
    for (int i = ndim - 1; i-- > 0;)
    {
        int irrep = dirrep[d->parent[2 * i]];
        dstride[2 * i] = dstride[d->parent[2 * i]];
        dstride[2 * i + 1] = dstride[2 * i] * d->size[2 * i][dirrep[2 * i]];
        int64_t off = 0;
        for (int irr1 = 0; irr1 < dirrep[2 * i + 1]; irr1++)
            off += d->size[2 * i][irr1 ^ irrep] * d->size[2 * i + 1][irr1];
        data += off * dstride[2 * i];
    }
    for (int i = 0; i < ndim; i++)
    {
        stride[i] = dstride[d->leaf[d->perm[i]]];
        len[i] = d->ext[d->perm[i]][irreps[i]];
    }
    *offset = data;
}

/* ==================================================================== */
/* DPD contraction oracle                                                */
/* ==================================================================== */

static int find_label(const char* idx, int n, char c)
{
    for (int i = 0; i < n; i++) if (idx[i] == c) return i;
    return -1;
}

/* dense expansion: per-dim irrep offsets (block_to_full, util.hpp:66-98) */
static void dpd_dense_meta(const orc_dpd* d, int64_t* tot, int64_t off[][8])
{
    for (int i = 0; i < d->ndim; i++)
    {
        tot[i] = 0;
        for (int r = 0; r < d->nirrep; r++) { off[i][r] = tot[i]; tot[i] += d->len[i][r]; }
    }
}

static void dpd_mult_impl(int acc, double alpha, const orc_dpd* dA, const char* idx_A,
                          const void* A, const orc_dpd* dB, const char* idx_B, const void* B,
                          double beta, const orc_dpd* dC, const char* idx_C, const void* Cin,
                          double* Cout, int is_float)
{
    int nirrep = dA->nirrep;
    /* gather the set of labels: C's labels then A-only-with-B (contracted) */
    char lab[3 * ORC_MAXDIM];
    int nlab = 0;
    for (int i = 0; i < dC->ndim; i++) lab[nlab++] = idx_C[i];
    for (int i = 0; i < dA->ndim; i++)
        if (find_label(lab, nlab, idx_A[i]) < 0) lab[nlab++] = idx_A[i];
    int nc = dC->ndim;
    int nk = nlab - nc;
    int64_t lablen[3 * ORC_MAXDIM][8];
    for (int q = 0; q < nlab; q++)
    {
        int a = find_label(idx_A, dA->ndim, lab[q]);
        int b = find_label(idx_B, dB->ndim, lab[q]);
        int c = find_label(idx_C, dC->ndim, lab[q]);
        for (int r = 0; r < nirrep; r++)
            lablen[q][r] = c >= 0 ? dC->len[c][r] : a >= 0 ? dA->len[a][r] : dB->len[b][r];
    }

    if (acc == 2)
    {
        /* full dense expansion (mult_full, dpd/mult.cxx:29-74) */
        int64_t totA[ORC_MAXDIM], totB[ORC_MAXDIM], totC[ORC_MAXDIM];
        int64_t offA[ORC_MAXDIM][8], offB[ORC_MAXDIM][8], offC[ORC_MAXDIM][8];
        dpd_dense_meta(dA, totA, offA); dpd_dense_meta(dB, totB, offB); dpd_dense_meta(dC, totC, offC);
        int64_t nA = prod(totA, dA->ndim), nB = prod(totB, dB->ndim), nC = prod(totC, dC->ndim);
        double* FA = calloc(nA, sizeof(double));
        double* FB = calloc(nB, sizeof(double));
        double* FC = calloc(nC, sizeof(double));
        const orc_dpd* ds[3] = {dA, dB, dC};
        const void* src[3] = {A, B, Cin};
        double* dst[3] = {FA, FB, FC};
        int64_t* tots[3] = {totA, totB, totC};
        int64_t (*offs[3])[8] = {offA, offB, offC};
        for (int t = 0; t < 3; t++)
        {
            const orc_dpd* d = ds[t];
            int ndim = d->ndim;
            int64_t fstr[ORC_MAXDIM];
            fstr[0] = 1;
            for (int i = 1; i < ndim; i++) fstr[i] = fstr[i - 1] * tots[t][i - 1];
            int64_t nb = 1;
            for (int i = 1; i < ndim; i++) nb *= nirrep;
            for (int64_t b = 0; b < nb; b++)
            {
                int irr[ORC_MAXDIM];
                int64_t bb = b;
                irr[0] = d->irrep;
                for (int i = 1; i < ndim; i++) { irr[i] = bb % nirrep; bb /= nirrep; irr[0] ^= irr[i]; }
                int64_t blen[ORC_MAXDIM], bstr[ORC_MAXDIM], boff;
                orc_dpd_get_block(d, irr, blen, &boff, bstr);
                int64_t ne = prod(blen, ndim);
                for (int64_t e = 0; e < ne; e++)
                {
                    int64_t f = e, so = boff, fo = 0;
                    for (int i = 0; i < ndim; i++)
                    {
                        int64_t x = f % blen[i]; f /= blen[i];
                        so += x * bstr[i];
                        fo += (offs[t][i][irr[i]] + x) * fstr[i];
                    }
                    dst[t][fo] = is_float ? ((const float*)src[t])[so] : ((const double*)src[t])[so];
                }
            }
        }
        /* dense contraction via the dense oracle (column-major full arrays) */
        int64_t sAf[ORC_MAXDIM], sBf[ORC_MAXDIM], sCf[ORC_MAXDIM];
        sAf[0] = sBf[0] = sCf[0] = 1;
        for (int i = 1; i < dA->ndim; i++) sAf[i] = sAf[i - 1] * totA[i - 1];
        for (int i = 1; i < dB->ndim; i++) sBf[i] = sBf[i - 1] * totB[i - 1];
        for (int i = 1; i < dC->ndim; i++) sCf[i] = sCf[i - 1] * totC[i - 1];
        orc_tensor_mult_d(0, alpha, dA->ndim, totA, sAf, idx_A, FA, dB->ndim, totB, sBf, idx_B, FB,
                          beta, dC->ndim, totC, sCf, idx_C, FC);
        /* full_to_block for C */
        {
            const orc_dpd* d = dC;
            int ndim = d->ndim;
            int64_t nb = 1;
            for (int i = 1; i < ndim; i++) nb *= nirrep;
            for (int64_t b = 0; b < nb; b++)
            {
                int irr[ORC_MAXDIM];
                int64_t bb = b;
                irr[0] = d->irrep;
                for (int i = 1; i < ndim; i++) { irr[i] = bb % nirrep; bb /= nirrep; irr[0] ^= irr[i]; }
                int64_t blen[ORC_MAXDIM], bstr[ORC_MAXDIM], boff;
                orc_dpd_get_block(d, irr, blen, &boff, bstr);
                int64_t ne = prod(blen, ndim);
                for (int64_t e = 0; e < ne; e++)
                {
                    int64_t f = e, so = boff, fo = 0;
                    for (int i = 0; i < ndim; i++)
                    {
                        int64_t x = f % blen[i]; f /= blen[i];
                        so += x * bstr[i];
                        fo += (offC[i][irr[i]] + x) * sCf[i];
                    }
                    Cout[so] = FC[fo];
                }
            }
        }
        free(FA); free(FB); free(FC);
        return;
    }

    /* block-wise: for each C block, sum over irreps of the contracted labels */
    int64_t nbC = 1;
    for (int i = 1; i < nc; i++) nbC *= nirrep;
    int64_t nbK = 1;
    for (int i = 0; i < nk; i++) nbK *= nirrep;

    for (int64_t b = 0; b < nbC; b++)
    {
        int irr[3 * ORC_MAXDIM];
        int64_t bb = b;
        irr[0] = dC->irrep;
        for (int i = 1; i < nc; i++) { irr[i] = bb % nirrep; bb /= nirrep; irr[0] ^= irr[i]; }
        int irrC[ORC_MAXDIM];
        for (int i = 0; i < nc; i++) irrC[i] = irr[i];
        int64_t clen[ORC_MAXDIM], cstr[ORC_MAXDIM], coff;
        orc_dpd_get_block(dC, irrC, clen, &coff, cstr);
        int64_t ncel = prod(clen, nc);
        if (ncel == 0) continue;
        long double* sum = calloc(ncel, sizeof(long double));
        int any = 0;
        for (int64_t kb = 0; kb < nbK; kb++)
        {
            int64_t kk = kb;
            for (int i = 0; i < nk; i++) { irr[nc + i] = kk % nirrep; kk /= nirrep; }
            int irrA[ORC_MAXDIM], irrB[ORC_MAXDIM], xa = 0, xb = 0;
            for (int i = 0; i < dA->ndim; i++)
            {
                irrA[i] = irr[find_label(lab, nlab, idx_A[i])];
                xa ^= irrA[i];
            }
            for (int i = 0; i < dB->ndim; i++)
            {
                irrB[i] = irr[find_label(lab, nlab, idx_B[i])];
                xb ^= irrB[i];
            }
            if (xa != dA->irrep || xb != dB->irrep) continue;
            int64_t alen[ORC_MAXDIM], astr[ORC_MAXDIM], aoff, blen[ORC_MAXDIM], bstr[ORC_MAXDIM], boff;
            orc_dpd_get_block(dA, irrA, alen, &aoff, astr);
            orc_dpd_get_block(dB, irrB, blen, &boff, bstr);
            int64_t klen[3 * ORC_MAXDIM];
            int64_t nke = 1;
            for (int i = 0; i < nk; i++) { klen[i] = lablen[nc + i][irr[nc + i]]; nke *= klen[i]; }
            if (nke == 0) continue;
            any = 1;
            /* strides of every label in A and B blocks */
            int64_t sAl[3 * ORC_MAXDIM], sBl[3 * ORC_MAXDIM];
            for (int q = 0; q < nlab; q++)
            {
                int a = find_label(idx_A, dA->ndim, lab[q]);
                int bq = find_label(idx_B, dB->ndim, lab[q]);
                sAl[q] = a >= 0 ? astr[a] : 0;
                sBl[q] = bq >= 0 ? bstr[bq] : 0;
            }
            #pragma omp parallel for schedule(static)
            for (int64_t e = 0; e < ncel; e++)
            {
                int64_t f = e, oa = aoff, ob = boff;
                for (int i = 0; i < nc; i++)
                {
                    int64_t x = f % clen[i]; f /= clen[i];
                    oa += x * sAl[i]; ob += x * sBl[i];
                }
                long double s = 0;
                for (int64_t ke = 0; ke < nke; ke++)
                {
                    int64_t g2 = ke, oa2 = oa, ob2 = ob;
                    for (int i = 0; i < nk; i++)
                    {
                        int64_t x = g2 % klen[i]; g2 /= klen[i];
                        oa2 += x * sAl[nc + i]; ob2 += x * sBl[nc + i];
                    }
                    long double av = is_float ? ((const float*)A)[oa2] : ((const double*)A)[oa2];
                    long double bv = is_float ? ((const float*)B)[ob2] : ((const double*)B)[ob2];
                    if (acc) s += av * bv;
                    else s = (double)((double)s + (double)av * (double)bv);
                }
                sum[e] += s;
            }
        }
        (void)any;
        for (int64_t e = 0; e < ncel; e++)
        {
            int64_t f = e, oc = coff;
            for (int i = 0; i < nc; i++) { int64_t x = f % clen[i]; f /= clen[i]; oc += x * cstr[i]; }
            long double t = (long double)alpha * sum[e];
            if (beta != 0.0)
                t += (long double)beta * (is_float ? ((const float*)Cin)[oc] : ((const double*)Cin)[oc]);
            Cout[oc] = (double)t;
        }
        free(sum);
    }
}

void orc_dpd_mult_d(int acc, double alpha, const orc_dpd* dA, const char* idx_A, const double* A,
                    const orc_dpd* dB, const char* idx_B, const double* B,
                    double beta, const orc_dpd* dC, const char* idx_C, double* C)
{
    if (acc == 2)
    {
        dpd_mult_impl(2, alpha, dA, idx_A, A, dB, idx_B, B, beta, dC, idx_C, C, C, 0);
        return;
    }
    /* block-wise reads C before writing each block only once: in-place OK */
    dpd_mult_impl(acc, alpha, dA, idx_A, A, dB, idx_B, B, beta, dC, idx_C, C, C, 0);
}

void orc_dpd_mult_s(double alpha, const orc_dpd* dA, const char* idx_A, const float* A,
                    const orc_dpd* dB, const char* idx_B, const float* B,
                    double beta, const orc_dpd* dC, const char* idx_C, double* C_out,
                    const float* C_in)
{
    dpd_mult_impl(1, alpha, dA, idx_A, A, dB, idx_B, B, beta, dC, idx_C, C_in, C_out, 1);
}

/* DPD patch scatter: dense_total_lengths_and_strides (dpd/util.hpp:136-173),
 * the reorder of dpd/mult.cxx:256-306, dpd_tensor_matrix block enumeration
 * (dpd_tensor_matrix.hpp:41-101) and the per-patch fill of
 * patch_block_scatter_matrix.hpp:133-228 for the whole matrix (offset 0). */
static void dpd_dense_strides(const orc_dpd* d, int64_t* stride)
{
    int ndim = d->ndim;
    int64_t len[ORC_MAXDIM];
    for (int j = 0; j < ndim; j++)
    {
        len[j] = 0;
        for (int r = 0; r < d->nirrep; r++) len[j] += d->len[j][r];
    }
    int iperm[ORC_MAXDIM] = {0};
    for (int i = 0; i < ndim; i++) iperm[d->perm[i]] = i;
    stride[iperm[0]] = 1;
    for (int j = 1; j < ndim; j++) stride[iperm[j]] = stride[iperm[j - 1]] * len[iperm[j - 1]];
}

int orc_dpd_patch_scatter(const orc_dpd* dA, const char* idx_A, const orc_dpd* dB,
                          const char* idx_B, const orc_dpd* dC, const char* idx_C,
                          int which, int irrep_ab, int64_t BS_row, int64_t BS_col,
                          int* npatch, int64_t* patch_rows, int64_t* patch_cols,
                          int64_t* patch_base, int64_t* rscat, int64_t* cscat,
                          int64_t max_scat)
{
    int nirrep = dA->nirrep;
    /* dimension index lists (iface/3t/mult.cxx:240-250): positions in each
     * tensor, in sorted-label order */
    char lab_ab[ORC_MAXDIM], lab_ac[ORC_MAXDIM], lab_bc[ORC_MAXDIM], t1[3 * ORC_MAXDIM];
    int n1 = lab_intersect(idx_A, dA->ndim, idx_B, dB->ndim, t1);
    char abc[ORC_MAXDIM];
    int nabc = lab_intersect(t1, n1, idx_C, dC->ndim, abc);
    if (nabc) return -1;
    int nab = lab_exclude(t1, n1, abc, 0, lab_ab);
    n1 = lab_intersect(idx_A, dA->ndim, idx_C, dC->ndim, t1);
    int nac = lab_exclude(t1, n1, abc, 0, lab_ac);
    n1 = lab_intersect(idx_B, dB->ndim, idx_C, dC->ndim, t1);
    int nbc = lab_exclude(t1, n1, abc, 0, lab_bc);

    int iA_AB[ORC_MAXDIM], iB_AB[ORC_MAXDIM], iA_AC[ORC_MAXDIM], iC_AC[ORC_MAXDIM];
    int iB_BC[ORC_MAXDIM], iC_BC[ORC_MAXDIM];
    for (int i = 0; i < nab; i++)
    {
        iA_AB[i] = find_label(idx_A, dA->ndim, lab_ab[i]);
        iB_AB[i] = find_label(idx_B, dB->ndim, lab_ab[i]);
    }
    for (int i = 0; i < nac; i++)
    {
        iA_AC[i] = find_label(idx_A, dA->ndim, lab_ac[i]);
        iC_AC[i] = find_label(idx_C, dC->ndim, lab_ac[i]);
    }
    for (int i = 0; i < nbc; i++)
    {
        iB_BC[i] = find_label(idx_B, dB->ndim, lab_bc[i]);
        iC_BC[i] = find_label(idx_C, dC->ndim, lab_bc[i]);
    }

    int64_t stA[ORC_MAXDIM], stB[ORC_MAXDIM], stC[ORC_MAXDIM];
    dpd_dense_strides(dA, stA); dpd_dense_strides(dB, stB); dpd_dense_strides(dC, stC);

    int64_t kAC[2][ORC_MAXDIM], kBC[2][ORC_MAXDIM], kAB[2][ORC_MAXDIM];
    for (int i = 0; i < nac; i++) { kAC[0][i] = stC[iC_AC[i]]; kAC[1][i] = stA[iA_AC[i]]; }
    for (int i = 0; i < nbc; i++) { kBC[0][i] = stC[iC_BC[i]]; kBC[1][i] = stB[iB_BC[i]]; }
    for (int i = 0; i < nab; i++) { kAB[0][i] = stA[iA_AB[i]]; kAB[1][i] = stB[iB_AB[i]]; }
    int pAC[ORC_MAXDIM], pBC[ORC_MAXDIM], pAB[ORC_MAXDIM];
    const int64_t* sac[2] = {kAC[0], kAC[1]};
    const int64_t* sbc[2] = {kBC[0], kBC[1]};
    const int64_t* sab[2] = {kAB[0], kAB[1]};
    orc_sort_by_stride(nac, 2, sac, pAC);
    orc_sort_by_stride(nbc, 2, sbc, pBC);
    orc_sort_by_stride(nab, 2, sab, pAB);

    int t[ORC_MAXDIM];
#define PERMUTE(v, p, n) do { for (int q_ = 0; q_ < (n); q_++) t[q_] = v[p[q_]]; \
                              for (int q_ = 0; q_ < (n); q_++) v[q_] = t[q_]; } while (0)
    PERMUTE(iA_AC, pAC, nac); PERMUTE(iC_AC, pAC, nac);
    PERMUTE(iA_AB, pAB, nab); PERMUTE(iB_AB, pAB, nab);
    PERMUTE(iB_BC, pBC, nbc); PERMUTE(iC_BC, pBC, nbc);
#undef PERMUTE

    /* unit_dim over tensor dims (dpd/mult.cxx:277-282) */
    int uA_AC = nac, uB_BC = nbc, uA_AB = nab, uB_AB = nab;
    for (int i = 0; i < nac; i++) if (stA[iA_AC[i]] == 1) { uA_AC = i; break; }
    for (int i = 0; i < nbc; i++) if (stB[iB_BC[i]] == 1) { uB_BC = i; break; }
    for (int i = 0; i < nab; i++) if (stA[iA_AB[i]] == 1) { uA_AB = i; break; }
    for (int i = 0; i < nab; i++) if (stB[iB_AB[i]] == 1) { uB_AB = i; break; }
    int pm = uA_AC > 0 && uA_AC < nac;
    int pn = uB_BC > 0 && uB_BC < nbc;
    int pk = (uA_AB > 0 && uA_AB < nab) || (uB_AB > 0 && uB_AB < nab);
    if (pm) { rotate_to_second(iA_AC, nac, uA_AC); rotate_to_second(iC_AC, nac, uA_AC); }
    if (pn) { rotate_to_second(iB_BC, nbc, uB_BC); rotate_to_second(iC_BC, nbc, uB_BC); }
    if (pk)
    {
        int u = uA_AB > uB_AB ? uA_AB : uB_AB;
        rotate_to_second(iA_AB, nab, u); rotate_to_second(iB_AB, nab, u);
    }

    const orc_dpd* d;
    const int *rows, *cols;
    int nr, ncl, col_irrep, prow, pcol;
    if (which == 0)
    {
        d = dA; rows = iA_AC; nr = nac; cols = iA_AB; ncl = nab; col_irrep = irrep_ab;
        prow = pm; pcol = pk;
    }
    else if (which == 1)
    {
        d = dB; rows = iB_AB; nr = nab; cols = iB_BC; ncl = nbc; col_irrep = dB->irrep ^ irrep_ab;
        prow = pk; pcol = pn;
    }
    else
    {
        d = dC; rows = iC_AC; nr = nac; cols = iC_BC; ncl = nbc; col_irrep = dB->irrep ^ irrep_ab;
        prow = pm; pcol = pn;
    }
    int irrep2[2] = {col_irrep ^ d->irrep, col_irrep};
    const int* dims[2] = {rows, cols};
    int nd[2] = {nr, ncl};
    int pack3d[2] = {prow, pcol};
    int64_t BS[2] = {BS_row, BS_col};

    /* block enumeration (dpd_tensor_matrix.hpp:72-101) */
    int64_t bsize[2][4096];
    int bidx[2][4096];
    int nblk[2] = {0, 0};
    for (int dm = 0; dm < 2; dm++)
    {
        if (nd[dm] == 0)
        {
            bsize[dm][0] = irrep2[dm] == 0 ? 1 : 0;
            bidx[dm][0] = 0;
            nblk[dm] = 1;
            continue;
        }
        int64_t ncombo = 1;
        for (int i = 1; i < nd[dm]; i++) ncombo *= nirrep;
        for (int64_t idx = 0; idx < ncombo; idx++)
        {
            int64_t x = idx;
            int irr0 = irrep2[dm];
            int64_t size = 1;
            for (int i = 1; i < nd[dm]; i++)
            {
                int ri = x % nirrep; x /= nirrep;
                irr0 ^= ri;
                size *= d->len[dims[dm][i]][ri];
            }
            size *= d->len[dims[dm][0]][irr0];
            if (size == 0) continue;
            bsize[dm][nblk[dm]] = size;
            bidx[dm][nblk[dm]] = (int)idx;
            nblk[dm]++;
        }
    }
    npatch[0] = nblk[0];
    npatch[1] = nblk[1];
    int irrep_mask = nirrep - 1;
    int irrep_bits = __builtin_popcount(irrep_mask);
    int64_t rpos = 0, cpos = 0;
    int np = 0;
    for (int p0 = 0; p0 < nblk[0]; p0++)
        for (int p1 = 0; p1 < nblk[1]; p1++)
        {
            int irr[ORC_MAXDIM] = {0};
            int blk[2] = {p0, p1};
            for (int dm = 0; dm < 2; dm++)
            {
                if (nd[dm] == 0) continue;
                int idx = bidx[dm][blk[dm]];
                irr[dims[dm][0]] = irrep2[dm];
                for (int i = 1; i < nd[dm]; i++)
                {
                    irr[dims[dm][i]] = idx & irrep_mask;
                    irr[dims[dm][0]] ^= irr[dims[dm][i]];
                    idx >>= irrep_bits;
                }
            }
            int64_t blen[ORC_MAXDIM], bstr[ORC_MAXDIM], boff;
            orc_dpd_get_block(d, irr, blen, &boff, bstr);
            patch_rows[np] = bsize[0][p0];
            patch_cols[np] = bsize[1][p1];
            patch_base[np] = boff;
            if (rpos + bsize[0][p0] > max_scat || cpos + bsize[1][p1] > max_scat) return -2;
            for (int dm = 0; dm < 2; dm++)
            {
                int64_t l[ORC_MAXDIM], s[ORC_MAXDIM];
                for (int i = 0; i < nd[dm]; i++) { l[i] = blen[dims[dm][i]]; s[i] = bstr[dims[dm][i]]; }
                int64_t* out = dm == 0 ? rscat + rpos : cscat + cpos;
                orc_fill_scatter(nd[dm], l, s, BS[dm], 0, bsize[dm][blk[dm]], out, pack3d[dm], 8);
            }
            rpos += bsize[0][p0];
            cpos += bsize[1][p1];
            np++;
        }
    return np;
}

/* ==================================================================== */
/* Deterministic inputs: mt19937_64, u = (x >> 11) * 2^-53, v = 2u - 1.  */
/* ==================================================================== */
void orc_fill_uniform_d(double* x, int64_t n, uint64_t seed)
{
    uint64_t mt[312];
    int mti;
    mt[0] = seed;
    for (mti = 1; mti < 312; mti++)
        mt[mti] = 6364136223846793005ULL * (mt[mti - 1] ^ (mt[mti - 1] >> 62)) + (uint64_t)mti;
    const uint64_t UM = 0xFFFFFFFF80000000ULL, LM = 0x7FFFFFFFULL;
    const uint64_t MAG[2] = {0ULL, 0xB5026F5AA96619E9ULL};
    for (int64_t k = 0; k < n; k++)
    {
        if (mti >= 312)
        {
            int i;
            for (i = 0; i < 312 - 156; i++)
            {
                uint64_t y = (mt[i] & UM) | (mt[i + 1] & LM);
                mt[i] = mt[i + 156] ^ (y >> 1) ^ MAG[(int)(y & 1ULL)];
            }
            for (; i < 311; i++)
            {
                uint64_t y = (mt[i] & UM) | (mt[i + 1] & LM);
                mt[i] = mt[i + (156 - 312)] ^ (y >> 1) ^ MAG[(int)(y & 1ULL)];
            }
            uint64_t y = (mt[311] & UM) | (mt[0] & LM);
            mt[311] = mt[155] ^ (y >> 1) ^ MAG[(int)(y & 1ULL)];
            mti = 0;
        }
        uint64_t y = mt[mti++];
        y ^= (y >> 29) & 0x5555555555555555ULL;
        y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
        y ^= (y << 37) & 0xFFF7EEE000000000ULL;
        y ^= (y >> 43);
        x[k] = 2.0 * ((double)(y >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
    }
}


/* ------------------------------------------------------------------ */
/* Level-1 operations                                                   */
/* ------------------------------------------------------------------ */

/* Distinct labels of the kept (length != 1) dimensions of B, then those of
 * A not in B; B's elements are enumerated in an odometer over the first
 * group, A's sum over the second. */
static void level1(int mode, double alpha, int ndA, const int64_t* lenA, const int64_t* strA,
                   const char* idxA, const double* Ad, const float* As, double beta, int ndB,
                   const int64_t* lenB, const int64_t* strB, const char* idxB, const double* Bd_in,
                   const float* Bs_in, double* B_out)
{
    char lab[64];
    int64_t len[64];
    int nb = 0, nt = 0;
    for (int i = 0; i < ndB; i++) {
        if (lenB[i] == 1) continue;
        int seen = 0;
        for (int j = 0; j < nb; j++) seen |= lab[j] == idxB[i];
        if (!seen) { lab[nb] = idxB[i]; len[nb] = lenB[i]; nb++; }
    }
    nt = nb;
    if (mode == 0)
        for (int i = 0; i < ndA; i++) {
            if (lenA[i] == 1) continue;
            int seen = 0;
            for (int j = 0; j < nt; j++) seen |= lab[j] == idxA[i];
            if (!seen) { lab[nt] = idxA[i]; len[nt] = lenA[i]; nt++; }
        }
    int64_t nB = 1, nS = 1;
    for (int j = 0; j < nb; j++) nB *= len[j];
    for (int j = nb; j < nt; j++) nS *= len[j];
    int64_t val[64];
    for (int64_t e = 0; e < nB; e++) {
        int64_t r = e;
        for (int j = 0; j < nb; j++) { val[j] = r % len[j]; r /= len[j]; }
        int64_t offB = 0;
        for (int i = 0; i < ndB; i++) {
            if (lenB[i] == 1) continue;
            for (int j = 0; j < nb; j++) if (lab[j] == idxB[i]) offB += val[j] * strB[i];
        }
        double v;
        if (mode == 2) {
            v = alpha;
        } else {
            double old = 0.0;
            if (beta != 0.0) old = Bd_in ? Bd_in[offB] : (double)Bs_in[offB];
            if (mode == 1) {
                v = beta * old;
            } else {
                long double sum = 0.0L;
                for (int64_t f = 0; f < nS; f++) {
                    int64_t q = f;
                    for (int j = nb; j < nt; j++) { val[j] = q % len[j]; q /= len[j]; }
                    int64_t offA = 0;
                    for (int i = 0; i < ndA; i++) {
                        if (lenA[i] == 1) continue;
                        for (int j = 0; j < nt; j++) if (lab[j] == idxA[i]) offA += val[j] * strA[i];
                    }
                    sum += Ad ? (long double)Ad[offA] : (long double)As[offA];
                }
                v = (double)((long double)alpha * sum);
                if (beta != 0.0) v += beta * old;
            }
        }
        B_out[offB] = v;
    }
}

void orc_level1_d(int mode, double alpha, int ndA, const int64_t* lenA, const int64_t* strA,
                  const char* idxA, const double* A, double beta, int ndB, const int64_t* lenB,
                  const int64_t* strB, const char* idxB, double* B)
{
    level1(mode, alpha, ndA, lenA, strA, idxA, A, NULL, beta, ndB, lenB, strB, idxB, B, NULL, B);
}

void orc_level1_s(int mode, double alpha, int ndA, const int64_t* lenA, const int64_t* strA,
                  const char* idxA, const float* A, double beta, int ndB, const int64_t* lenB,
                  const int64_t* strB, const char* idxB, const float* B_in, double* B_out)
{
    level1(mode, alpha, ndA, lenA, strA, idxA, NULL, A, beta, ndB, lenB, strB, idxB, NULL, B_in,
           B_out);
}
