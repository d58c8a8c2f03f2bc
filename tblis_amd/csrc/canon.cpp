// Front end of the contraction: descriptor -> canonical groups -> matrix views.
//
// Restates the integer canonicalisation of TBLIS so that the GEMM this
// library launches sees exactly the (length, stride) lists the reference
// hands to TensorGEMM:
//   diagonal            src/util/tensor.hpp:375-412
//   free_idx            src/util/tensor.hpp:33-62
//   sort_by_stride      src/util/tensor.hpp:95-148
//   fold                src/util/tensor.hpp:340-373
//   unit_dim            src/util/tensor.hpp:484-491
//   group split         src/iface/3t/mult.cxx:24-116
//   mult_blis reorder   src/internal/3t/dense/mult.cxx:212-263, 479-556
#include <algorithm>
#include <cstring>

#include "common.hpp"

namespace tblis_amd
{
namespace
{

struct TDim
{
    char label;
    int64_t len;
    int64_t stride;
};

// Label order is the built-in < on (signed) char, as in the reference's
// std::sort over label vectors.
inline bool label_less(char a, char b) { return (signed char)a < (signed char)b; }

// Repeated labels take the diagonal (strides add), length-1 dimensions drop.
bool take_diagonal(const tblis_tensor& t, const label_type* idx, std::vector<TDim>& out,
                   std::string& why)
{
    std::vector<TDim> in;
    for (unsigned i = 0; i < t.ndim; i++) in.push_back({idx[i], t.len[i], t.stride[i]});
    std::stable_sort(in.begin(), in.end(),
                     [](const TDim& x, const TDim& y) { return label_less(x.label, y.label); });
    out.clear();
    for (size_t i = 0; i < in.size(); i++)
    {
        if (in[i].len < 0)
        {
            why = "negative tensor length";
            return false;
        }
        bool repeat = i > 0 && in[i].label == in[i - 1].label;
        if (!repeat)
        {
            if (in[i].len != 1) out.push_back(in[i]);
        }
        else if (in[i].len != 1)
        {
            if (out.empty() || out.back().label != in[i].label || out.back().len != in[i].len)
            {
                why = std::string("repeated label '") + in[i].label + "' with unequal lengths";
                return false;
            }
            out.back().stride += in[i].stride;
        }
    }
    return true;
}

std::vector<char> sorted_labels(const std::vector<TDim>& v)
{
    std::vector<char> l;
    for (auto& d : v) l.push_back(d.label);
    std::sort(l.begin(), l.end(), label_less);
    return l;
}

std::vector<char> set_and(const std::vector<char>& a, const std::vector<char>& b)
{
    std::vector<char> r;
    std::set_intersection(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(r),
                          label_less);
    return r;
}

std::vector<char> set_minus(const std::vector<char>& a, const std::vector<char>& b)
{
    std::vector<char> r;
    std::set_difference(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(r), label_less);
    return r;
}

std::vector<char> set_or(const std::vector<char>& a, const std::vector<char>& b)
{
    std::vector<char> r;
    std::set_union(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(r), label_less);
    return r;
}

// Smallest label value not in the (sorted) set: 0 if every label is
// positive, else the first gap, else one past the largest.
char first_free_label(const std::vector<char>& s)
{
    if (s.empty() || (signed char)s[0] > 0) return 0;
    for (size_t i = 1; i < s.size(); i++)
        if ((signed char)s[i] > (signed char)s[i - 1] + 1) return (char)(s[i - 1] + 1);
    return (char)(s.back() + 1);
}

const TDim* find(const std::vector<TDim>& v, char l)
{
    for (auto& d : v)
        if (d.label == l) return &d;
    return nullptr;
}

// Stable order by (min stride over the listed slots, then each slot's
// stride in turn).
void sort_by_strides(Group& g, std::initializer_list<int> slots)
{
    std::vector<int> sl(slots);
    std::stable_sort(g.begin(), g.end(), [&](const Dim& x, const Dim& y) {
        int64_t mx = x.s[sl[0]], my = y.s[sl[0]];
        for (size_t i = 1; i < sl.size(); i++)
        {
            mx = std::min(mx, x.s[sl[i]]);
            my = std::min(my, y.s[sl[i]]);
        }
        if (mx != my) return mx < my;
        for (int k : sl)
            if (x.s[k] != y.s[k]) return x.s[k] < y.s[k];
        return false;
    });
}

// Merge dims that are contiguous in every tensor of the group, scanning in
// the order of the first tensor's strides.
void fold_group(Group& g, int ntensor)
{
    Group sorted = g;
    sort_by_strides(sorted, {0});
    Group out;
    for (size_t i = 0; i < sorted.size(); i++)
    {
        bool contig = i > 0;
        for (int t = 0; contig && t < ntensor; t++)
            contig = sorted[i].s[t] == sorted[i - 1].s[t] * sorted[i - 1].len;
        if (contig)
            out.back().len *= sorted[i].len;
        else
            out.push_back(sorted[i]);
    }
    g = out;
}

int unit_position(const Group& g, int slot)
{
    for (size_t i = 0; i < g.size(); i++)
        if (g[i].s[slot] == 1) return (int)i;
    return (int)g.size();
}

void rotate_to_second(Group& g, int u) { std::rotate(g.begin() + 1, g.begin() + u, g.end()); }

}  // namespace

bool canonicalize_add(const tblis_tensor& A, const label_type* idx_A, const tblis_tensor& B,
                      const label_type* idx_B, Level1Form& out, std::string& why)
{
    std::vector<TDim> da, db;
    if (!take_diagonal(A, idx_A, da, why) || !take_diagonal(B, idx_B, db, why)) return false;
    const std::vector<char> la = sorted_labels(da), lb = sorted_labels(db);
    const std::vector<char> ab = set_and(la, lb), ao = set_minus(la, ab), bo = set_minus(lb, ab);
    if (!ao.empty() && !bo.empty())
    {
        why = "labels present only in A and labels present only in B";
        return false;
    }
    out = Level1Form();
    for (char l : ab)
    {
        const TDim *x = find(da, l), *y = find(db, l);
        if (x->len != y->len)
        {
            why = std::string("label '") + l + "' has different lengths in A and B";
            return false;
        }
        Dim d{};
        d.label = l;
        d.len = x->len;
        d.s[0] = x->stride;
        d.s[1] = y->stride;
        out.ab.push_back(d);
    }
    for (char l : ao)
    {
        const TDim* x = find(da, l);
        Dim d{};
        d.label = l;
        d.len = x->len;
        d.s[0] = x->stride;
        d.s[1] = 0;
        out.a_only.push_back(d);
    }
    for (char l : bo)
    {
        const TDim* y = find(db, l);
        Dim d{};
        d.label = l;
        d.len = y->len;
        d.s[0] = 0;
        d.s[1] = y->stride;
        out.b_only.push_back(d);
    }
    // fold (src/iface/1t/add.cxx:78-80): AB by A's strides over both
    // tensors, A-only by A's, B-only by B's
    fold_group(out.ab, 2);
    fold_group(out.a_only, 1);
    for (auto& d : out.b_only) std::swap(d.s[0], d.s[1]);
    fold_group(out.b_only, 1);
    for (auto& d : out.b_only) std::swap(d.s[0], d.s[1]);
    return true;
}

bool canonicalize_one(const tblis_tensor& A, const label_type* idx_A, Group& out, std::string& why)
{
    std::vector<TDim> da;
    if (!take_diagonal(A, idx_A, da, why)) return false;
    out.clear();
    for (auto& x : da)
    {
        Dim d{};
        d.label = x.label;
        d.len = x.len;
        d.s[0] = x.stride;
        out.push_back(d);
    }
    fold_group(out, 1);
    for (auto& d : out) std::swap(d.s[0], d.s[1]);
    return true;
}

bool canonicalize(const tblis_tensor& A, const label_type* idx_A, const tblis_tensor& B,
                  const label_type* idx_B, const tblis_tensor& C, const label_type* idx_C,
                  Canonical& out, std::string& why)
{
    std::vector<TDim> a, b, c;
    if (!take_diagonal(A, idx_A, a, why) || !take_diagonal(B, idx_B, b, why) ||
        !take_diagonal(C, idx_C, c, why))
        return false;

    auto la = sorted_labels(a), lb = sorted_labels(b), lc = sorted_labels(c);
    size_t nabc = set_and(set_and(la, lb), lc).size();

    // A tensor holding only batch (ABC) labels gets one dummy label shared
    // with C (first free label), then B gets another.
    if (la.size() == nabc || lb.size() == nabc || lc.size() == nabc)
    {
        char d1 = first_free_label(set_or(set_or(la, lb), lc));
        a.push_back({d1, 1, 0});
        c.push_back({d1, 1, 0});
        la = sorted_labels(a);
        lc = sorted_labels(c);
        char d2 = first_free_label(set_or(set_or(la, lb), lc));
        b.push_back({d2, 1, 0});
        c.push_back({d2, 1, 0});
        lb = sorted_labels(b);
        lc = sorted_labels(c);
    }

    auto abc = set_and(set_and(la, lb), lc);
    auto ab = set_minus(set_and(la, lb), abc);
    auto ac = set_minus(set_and(la, lc), abc);
    auto bc = set_minus(set_and(lb, lc), abc);

    if (ab.size() + ac.size() + abc.size() != la.size() ||
        ab.size() + bc.size() + abc.size() != lb.size() ||
        ac.size() + bc.size() + abc.size() != lc.size())
    {
        why = "a label appears in only one tensor";
        return false;
    }

    auto build = [&](const std::vector<char>& labels, const std::vector<TDim>* t0,
                     const std::vector<TDim>* t1, const std::vector<TDim>* t2, Group& g) {
        g.clear();
        for (char l : labels)
        {
            Dim d;
            d.label = l;
            const std::vector<TDim>* ts[3] = {t0, t1, t2};
            for (int k = 0; k < 3; k++)
            {
                if (!ts[k]) continue;
                const TDim* x = find(*ts[k], l);
                if (k == 0)
                    d.len = x->len;
                else if (x->len != d.len)
                {
                    why = std::string("label '") + l + "' has unequal lengths";
                    return false;
                }
                d.s[k] = x->stride;
            }
            g.push_back(d);
        }
        return true;
    };

    if (!build(abc, &a, &b, &c, out.abc) || !build(ab, &a, &b, nullptr, out.ab) ||
        !build(ac, &a, &c, nullptr, out.ac) || !build(bc, &b, &c, nullptr, out.bc))
        return false;

    fold_group(out.abc, 3);
    fold_group(out.ab, 2);
    fold_group(out.ac, 2);
    fold_group(out.bc, 2);
    return true;
}

MatrixViews blis_views(const Canonical& c)
{
    MatrixViews v;
    v.m = c.ac;   // slots A, C
    v.n = c.bc;   // slots B, C
    v.k = c.ab;   // slots A, B
    v.l = c.abc;  // slots A, B, C

    sort_by_strides(v.m, {1, 0});     // (C, A)
    sort_by_strides(v.n, {1, 0});     // (C, B)
    sort_by_strides(v.k, {0, 1});     // (A, B)
    sort_by_strides(v.l, {2, 0, 1});  // (C, A, B)

    int uA_m = unit_position(v.m, 0);
    int uB_n = unit_position(v.n, 0);
    int uA_k = unit_position(v.k, 0);
    int uB_k = unit_position(v.k, 1);
    int nm = (int)v.m.size(), nn = (int)v.n.size(), nk = (int)v.k.size();

    v.pack_m_3d = uA_m > 0 && uA_m < nm;
    v.pack_n_3d = uB_n > 0 && uB_n < nn;
    v.pack_k_3d = (uA_k > 0 && uA_k < nk) || (uB_k > 0 && uB_k < nk);

    if (v.pack_m_3d) rotate_to_second(v.m, uA_m);
    if (v.pack_n_3d) rotate_to_second(v.n, uB_n);
    if (v.pack_k_3d) rotate_to_second(v.k, std::max(uA_k, uB_k));
    return v;
}

}  // namespace tblis_amd
