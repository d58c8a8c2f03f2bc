// Host side of indexed-tensor contractions.
//
// An indexed tensor is a list of dense sub-tensors ("entries") placed at
// index values of its indexed dimensions, each with a factor
// (src/external/marray/include/indexed_varray_base.hpp:19-45).  The
// contraction C <- alpha A.B + beta C follows src/iface/3t/mult.cxx:299-402
// and src/internal/3t/indexed/mult.cxx:274-318: every C entry is scaled
// first, then for every triple of entries (a, b, c) whose index values agree
// on the labels they share, c's data gains
//     alpha * f_a * f_b * f_c * (a . b)
// restricted to the slices that the fixed index values select in the
// tensors where those labels are dense (the reference's "mixed" indices,
// src/internal/1t/indexed/util.hpp:98-140,272-330).  Labels dense in every
// tensor that has them are the GEMM dimensions (AB: k, AC: m, BC: n, ABC:
// batch).
//
// On the GPU every C block (entry and slice, and batch position) is one
// panel GEMM whose k blocks are the matching (a, b) pairs, each scaled by
// its coefficient while it is packed; the sum over pairs is the GEMM's k
// loop, so no two launches ever accumulate into the same C element.  The
// beta pass over the C entries runs first, in one launch.
#include "indexed.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <unordered_map>

#include "contract.hpp"
#include "dpd.hpp"
#include "level1.hpp"
#include "panel.hpp"
#include "plan.hpp"

namespace tblis_amd
{

namespace
{

template <typename T>
void put(std::string& k, const T& v)
{
    k.append((const char*)&v, sizeof(T));
}

void hip_ok(hipError_t e, const char* what)
{
    if (e != hipSuccess) fail(std::string(what) + ": " + hipGetErrorString(e));
}

// Host view of one operand: a plain indexed tensor (one irrep; an entry's
// dense part is one strided block) or an indexed DPD tensor (an entry's
// dense part is DPD storage whose blocks follow the irreps of its dense
// dimensions; index values count within their dimension's irrep).
struct Ix
{
    const label_type* idx = nullptr;
    int nd = 0, ndd = 0, nix = 0;
    int64_t n = 0;
    void* const* data = nullptr;
    const len_type* indices = nullptr;
    int nir = 1;
    unsigned irrep = 0;            // irrep of the whole tensor
    std::vector<unsigned> iirr;    // irrep of each indexed dimension
    std::vector<int64_t> lens;     // nd x nir lengths
    std::vector<int64_t> dstride;  // plain: strides of the dense dimensions
    bool dpd = false;
    int layout = 0;
    DpdLayout L;                   // DPD: layout of an entry's dense part
    int64_t lo = 0, hi = -1;       // element extent of an entry's dense part
    std::vector<int64_t> off;      // element offset of each entry from data[0]
    std::vector<double> fac;
    std::vector<double> fim;       // imaginary parts of the factors (complex types)
    bool zero(int64_t e) const { return fac[e] == 0.0 && fim[e] == 0.0; }
    int64_t value(int64_t e, int pos) const { return indices[e * nix + (pos - ndd)]; }
    int64_t len(int pos, unsigned r = 0) const { return lens[pos * nir + r]; }
    int pos_of(label_type l) const
    {
        for (int i = 0; i < nd; i++)
            if (idx[i] == l) return i;
        return -1;
    }
    bool dense(int pos) const { return pos < ndd; }
    unsigned ix_irrep(int pos) const { return iirr[pos - ndd]; }
    // lengths, strides and offset of the dense block whose dimensions have
    // the irreps irr[0 .. ndd)
    void block(const unsigned* irr, int64_t* blen, int64_t* bst, int64_t& boff) const
    {
        boff = 0;
        if (dpd)
        {
            if (ndd > 0) L.block(irr, blen, &boff, bst);
            return;
        }
        for (int i = 0; i < ndd; i++) blen[i] = lens[i], bst[i] = dstride[i];
    }
};

// Entry offsets from data[0] (in elements of es bytes) and factors.
void load_entries(Ix& x, const std::string& name, void* const* data, const void* factor, type_t type, int es)
{
    x.data = data;
    x.off.resize(x.n);
    x.fac.resize(x.n);
    x.fim.assign(x.n, 0.0);
    for (int64_t e = 0; e < x.n; e++)
    {
        if (!data[e]) fail(name + ": null entry pointer");
        const intptr_t d = (const char*)data[e] - (const char*)data[0];
        if (d % es) fail(name + ": entry pointers are not aligned to the element size");
        x.off[e] = (int64_t)(d / es);
        x.fac[e] = !factor ? 1.0
                   : type == TYPE_DOUBLE ? ((const double*)factor)[e]
                   : type == TYPE_SINGLE ? (double)((const float*)factor)[e]
                   : type == TYPE_DCOMPLEX ? ((const double*)factor)[2 * e]
                                           : (double)((const float*)factor)[2 * e];
        if (factor && type == TYPE_DCOMPLEX) x.fim[e] = ((const double*)factor)[2 * e + 1];
        if (factor && type == TYPE_SCOMPLEX) x.fim[e] = (double)((const float*)factor)[2 * e + 1];
    }
}

// Offsets [lo, hi] of a plain entry's dense part.
void dense_extent(const Ix& x, int64_t& lo, int64_t& hi)
{
    lo = hi = 0;
    for (int i = 0; i < x.ndd; i++)
    {
        if (x.len(i) == 0)
        {
            lo = 0, hi = -1;
            return;
        }
        const int64_t s = (x.len(i) - 1) * x.dstride[i];
        if (s < 0)
            lo += s;
        else
            hi += s;
    }
}

void load(Ix& x, const tblis_indexed_tensor* t, const label_type* idx, const char* nm, type_t type, int es)
{
    const std::string name(nm);
    if (!t) fail(name + ": null descriptor");
    if (t->type != type) fail("A, B and C must have the same type");
    if (t->ndim > 0 && !idx) fail(name + ": null label array");
    if (t->dense_ndim > t->ndim) fail(name + ": more dense dimensions than dimensions");
    if (t->ndim > 0 && !t->len) fail(name + ": null length array");
    if (t->dense_ndim > 0 && !t->dense_stride) fail(name + ": null stride array");
    if (t->num_indices < 0) fail(name + ": negative number of entries");
    if (t->num_indices > 0 && !t->data) fail(name + ": null data array");
    x.idx = idx;
    x.nd = (int)t->ndim;
    x.ndd = (int)t->dense_ndim;
    x.nix = x.nd - x.ndd;
    x.n = t->num_indices;
    x.indices = t->indices;
    if (x.nix > 0 && x.n > 0 && !t->indices) fail(name + ": null index array");
    if (x.nix == 0 && x.n > 1) fail(name + ": several entries without indexed dimensions");
    x.lens.assign(t->len, t->len + x.nd);
    x.dstride.assign(t->dense_stride, t->dense_stride + x.ndd);
    x.iirr.assign(x.nix, 0u);
    for (int i = 0; i < x.nd; i++)
    {
        if (t->len[i] < 0) fail(name + ": negative tensor length");
        for (int j = 0; j < i; j++)
            if (idx[i] == idx[j]) fail(name + ": repeated label '" + std::string(1, idx[i]) + "'");
    }
    load_entries(x, name, t->data, t->factor, type, es);
    for (int64_t e = 0; e < x.n; e++)
        for (int p = x.ndd; p < x.nd; p++)
        {
            const int64_t v = x.value(e, p);
            if (v < 0 || v >= t->len[p]) fail(name + ": index value out of range");
        }
    dense_extent(x, x.lo, x.hi);
}

// An indexed DPD operand (checked by indexed_dpd_tensor_mult first).
void load_dpd(Ix& x, const tblis_indexed_dpd_tensor* t, const label_type* idx, const char* nm, type_t type, int es)
{
    x.idx = idx;
    x.nd = (int)t->ndim;
    x.ndd = (int)t->dense_ndim;
    x.nix = x.nd - x.ndd;
    x.n = t->num_indices;
    x.indices = t->indices;
    x.nir = (int)t->nirrep;
    x.irrep = t->irrep;
    x.dpd = true;
    x.layout = t->layout;
    x.lens.assign(t->len, t->len + (size_t)x.nd * x.nir);
    x.iirr.assign(t->idx_irrep, t->idx_irrep + x.nix);
    unsigned dirrep = t->irrep;
    for (unsigned r : x.iirr) dirrep ^= r;
    if (x.ndd > 0) x.L.init(dirrep, t->nirrep, (unsigned)x.ndd, t->len, t->layout);
    load_entries(x, nm, t->data, t->factor, type, es);
    x.lo = 0;
    x.hi = x.ndd > 0 ? x.L.total() - 1 : 0;
}

void key_of(std::string& k, const Ix& x)
{
    put(k, (int32_t)x.nd);
    put(k, (int32_t)x.ndd);
    put(k, x.n);
    put(k, (int32_t)x.nir);
    put(k, x.irrep);
    put(k, (int32_t)(x.dpd ? x.layout : -1));
    k.append((const char*)x.lens.data(), sizeof(int64_t) * x.lens.size());
    k.append((const char*)x.dstride.data(), sizeof(int64_t) * x.dstride.size());
    k.append((const char*)x.iirr.data(), sizeof(unsigned) * x.iirr.size());
    k.append(x.idx, x.nd);
    if (x.n > 0 && x.nix > 0) k.append((const char*)x.indices, sizeof(len_type) * x.n * x.nix);
    k.append((const char*)x.off.data(), sizeof(int64_t) * x.n);
    k.append((const char*)x.fac.data(), sizeof(double) * x.n);
    k.append((const char*)x.fim.data(), sizeof(double) * x.n);
}

struct VecHash
{
    size_t operator()(const std::vector<int64_t>& v) const
    {
        uint64_t h = 1469598103934665603ull;
        for (int64_t x : v) h = (h ^ (uint64_t)x) * 1099511628211ull;
        return (size_t)h;
    }
};

// Built state of one indexed contraction (cached by shape, indices,
// pointer offsets and factors).
struct IndexedPlan
{
    std::shared_ptr<Plan> gemm;  // null: no GEMM work
    Level1Desc cdesc;            // C's dense part, for the beta pass
    int64_t nent = 0;            // C entries
    int64_t nunc = 0;            // C entries no GEMM writes whole (beta == 0 zeroes only these)
    bool whole = false;          // every GEMM writes whole C entries: beta == 0 goes to the epilogue
    void* d_tab = nullptr;       // C entry offsets (int64), factors (double), then the same for
                                 // the uncovered entries; complex: offsets, real and imaginary
                                 // parts of the factors
    bool pinned = false;         // used while capturing: never freed
    ~IndexedPlan()
    {
        if (d_tab && !pinned) (void)hipFree(d_tab);
    }
};

std::mutex g_mu;
std::map<std::string, std::shared_ptr<IndexedPlan>>* g_cache = new std::map<std::string, std::shared_ptr<IndexedPlan>>();

std::shared_ptr<IndexedPlan> cache_find(const std::string& k)
{
    std::lock_guard<std::mutex> g(g_mu);
    auto it = g_cache->find(k);
    return it == g_cache->end() ? nullptr : it->second;
}

void cache_store(const std::string& k, const std::shared_ptr<IndexedPlan>& p)
{
    std::lock_guard<std::mutex> g(g_mu);
    if (g_cache->size() >= 128)
        for (auto it = g_cache->begin(); it != g_cache->end();)
            if (!it->second->pinned && it->second.use_count() == 1)
                it = g_cache->erase(it);
            else
                ++it;
    (*g_cache)[k] = p;
}

// Dense dimensions of a plain C, smallest stride first, as a level-1
// descriptor; an indexed DPD C entry's storage is one contiguous run.
Level1Desc dense_desc(const Ix& c)
{
    Level1Desc d;
    if (c.dpd)
    {
        d.nb = c.hi - c.lo + 1;
        if (d.nb > 0)
        {
            d.ndb = 1;
            d.len_b[0] = d.nb;
            d.sb[0] = 1;
            d.sa_b[0] = 0;
        }
        return d;
    }
    std::vector<int> order(c.ndd);
    for (int i = 0; i < c.ndd; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return std::llabs(c.dstride[a]) < std::llabs(c.dstride[b]);
    });
    for (int i : order)
    {
        if (c.len(i) == 1) continue;
        d.nb *= c.len(i);
        // fold into the previous dimension when contiguous with it
        if (d.ndb > 0 && c.dstride[i] == d.sb[d.ndb - 1] * d.len_b[d.ndb - 1])
        {
            d.len_b[d.ndb - 1] *= c.len(i);
            continue;
        }
        if (d.ndb >= LEVEL1_MAXDIM) fail("indexed: more than 16 dense dimensions");
        d.len_b[d.ndb] = c.len(i);
        d.sb[d.ndb] = c.dstride[i];
        d.sa_b[d.ndb] = 0;
        d.ndb++;
    }
    for (int i = 0; i < c.ndd; i++)
        if (c.len(i) == 0) d.nb = 0;
    return d;
}

// cplx: interleaved complex tensors (offsets in complex elements, es the
// size of a real part); each GEMM then runs as two real GEMMs with the
// complex coefficients f_a f_b f_c as k-block scales (sa / sb = -1
// conjugate A / B).
std::shared_ptr<IndexedPlan> build(const Ix& A, const Ix& B, const Ix& C, int es, int dtype, int family, bool cplx,
                                   double sa, double sb)
{
    auto ip = std::make_shared<IndexedPlan>();
    ip->cdesc = dense_desc(C);
    ip->nent = C.n;
    std::vector<char> covered(C.n, 0);
    // device table of the C entries (and of those no GEMM covers whole)
    auto upload = [&]() {
        if (C.n == 0) return;
        std::vector<int64_t> tab(2 * C.n);
        for (int64_t e = 0; e < C.n; e++)
        {
            tab[e] = C.off[e];
            std::memcpy(&tab[C.n + e], &C.fac[e], sizeof(double));
        }
        if (cplx)
        {
            // offsets, real parts, imaginary parts
            for (int64_t e = 0; e < C.n; e++)
            {
                int64_t bits;
                std::memcpy(&bits, &C.fim[e], sizeof(double));
                tab.push_back(bits);
            }
            hip_ok(device_alloc((void**)&ip->d_tab, tab.size() * sizeof(int64_t)), "hipMalloc (indexed table)");
            hip_ok(hipMemcpy(ip->d_tab, tab.data(), tab.size() * sizeof(int64_t), hipMemcpyHostToDevice),
                   "hipMemcpy (indexed table)");
            return;
        }
        std::vector<int64_t> unc_off;
        std::vector<double> unc_fac;
        for (int64_t e = 0; e < C.n; e++)
            if (!covered[e]) unc_off.push_back(C.off[e]), unc_fac.push_back(C.fac[e]);
        ip->nunc = (int64_t)unc_off.size();
        tab.insert(tab.end(), unc_off.begin(), unc_off.end());
        for (double f : unc_fac)
        {
            int64_t bits;
            std::memcpy(&bits, &f, sizeof(double));
            tab.push_back(bits);
        }
        hip_ok(device_alloc((void**)&ip->d_tab, tab.size() * sizeof(int64_t)), "hipMalloc (indexed table)");
        hip_ok(hipMemcpy(ip->d_tab, tab.data(), tab.size() * sizeof(int64_t), hipMemcpyHostToDevice),
               "hipMemcpy (indexed table)");
    };

    // label classes
    struct Lab
    {
        label_type l;
        int pa, pb, pc;
    };
    std::vector<Lab> labs;
    auto add = [&](label_type l) {
        for (auto& x : labs)
            if (x.l == l) return;
        labs.push_back(Lab{l, A.pos_of(l), B.pos_of(l), C.pos_of(l)});
    };
    for (int i = 0; i < A.nd; i++) add(A.idx[i]);
    for (int i = 0; i < B.nd; i++) add(B.idx[i]);
    for (int i = 0; i < C.nd; i++) add(C.idx[i]);
    auto all_dense = [&](const Lab& x) {
        return (x.pa < 0 || A.dense(x.pa)) && (x.pb < 0 || B.dense(x.pb)) && (x.pc < 0 || C.dense(x.pc));
    };
    // C is sliced when a label dense in C is fixed by an A or B entry; then
    // a GEMM writes part of a C entry, otherwise all of it
    bool sliced = false;
    for (const Lab& x : labs)
        if (x.pc >= 0 && C.dense(x.pc) && !all_dense(x)) sliced = true;

    // Irreps (indexed DPD; a plain tensor has the one irrep 0).  A label
    // indexed somewhere takes the irrep of its indexed dimension, and where
    // two tensors index it with different irreps nothing matches
    // (src/internal/3t/indexed_dpd/mult.cxx:1226-1280).  The labels dense in
    // every tensor that has them (the GEMM dimensions) take every irrep; an
    // assignment counts when each tensor's labels multiply to its irrep, and
    // then every operand's dense part is one block (the reference's blocked
    // algorithm, src/internal/3t/indexed_dpd/mult.cxx:117-994).
    const int nir = A.nir;
    std::vector<size_t> gl;  // labels dense everywhere
    std::vector<unsigned> lirr(labs.size(), 0);
    bool consistent = true;
    for (size_t i = 0; i < labs.size(); i++)
    {
        const Lab& x = labs[i];
        if (all_dense(x))
        {
            gl.push_back(i);
            continue;
        }
        int r = -1;
        for (const auto& [X, p] : {std::pair<const Ix*, int>{&A, x.pa}, {&B, x.pb}, {&C, x.pc}})
        {
            if (p < 0 || X->dense(p)) continue;
            const int q = (int)X->ix_irrep(p);
            if (r >= 0 && r != q) consistent = false;
            r = q;
        }
        lirr[i] = (unsigned)r;
    }

    // matching keys
    std::vector<const Lab*> kb, kc;  // indexed in A and B; indexed in C and in A or B
    for (const Lab& x : labs)
    {
        const bool ia = x.pa >= 0 && !A.dense(x.pa), ib = x.pb >= 0 && !B.dense(x.pb);
        const bool ic = x.pc >= 0 && !C.dense(x.pc);
        if (ia && ib) kb.push_back(&x);
        if (ic && (ia || ib)) kc.push_back(&x);
    }
    std::unordered_map<std::vector<int64_t>, std::vector<int64_t>, VecHash> bmap, cmap;
    std::vector<int64_t> key;
    for (int64_t e = 0; e < B.n; e++)
    {
        key.clear();
        for (const Lab* x : kb) key.push_back(B.value(e, x->pb));
        bmap[key].push_back(e);
    }
    for (int64_t e = 0; e < C.n; e++)
    {
        key.clear();
        for (const Lab* x : kc) key.push_back(C.value(e, x->pc));
        cmap[key].push_back(e);
    }

    // GEMM shape of one irrep assignment (its operand blocks)
    struct Shape
    {
        Side mA, mC, nB, nC, kA, kB;
        std::vector<int64_t> blen, bsa, bsb, bsc;  // ABC batch
        int64_t M = 1, N = 1, K = 1, NB = 1;
    };
    std::vector<Shape> shapes;
    // C regions (entry, slice, C block), in order of first appearance, each
    // with its k blocks
    struct KBlk
    {
        int64_t oa, ob;
        double coef, coef_im;
        int32_t sid;
    };
    struct Grp
    {
        int64_t oc;
        int32_t sid;  // shape of the first k block (C's side is the same for all)
        std::vector<KBlk> k;
    };
    std::map<int64_t, size_t> where;
    std::vector<Grp> groups;

    // The assignments of irreps to the labels dense everywhere that give each
    // tensor its irrep, depth first: a tensor's product is checked as soon
    // as its last such label is set, which prunes the search by a factor of
    // nir per tensor (the products of the other labels are fixed).
    std::vector<std::vector<unsigned>> assigns;
    if (consistent && A.n > 0 && B.n > 0 && C.n > 0)
    {
        const unsigned want[3] = {A.irrep, B.irrep, C.irrep};
        unsigned fx[3] = {0, 0, 0};
        int last[3] = {-1, -1, -1};
        std::vector<char> is_g(labs.size(), 0);
        for (size_t g : gl) is_g[g] = 1;
        for (size_t i = 0; i < labs.size(); i++)
        {
            const int pos[3] = {labs[i].pa, labs[i].pb, labs[i].pc};
            for (int x = 0; x < 3; x++)
                if (pos[x] >= 0 && !is_g[i]) fx[x] ^= lirr[i];
        }
        for (size_t d = 0; d < gl.size(); d++)
        {
            const int pos[3] = {labs[gl[d]].pa, labs[gl[d]].pb, labs[gl[d]].pc};
            for (int x = 0; x < 3; x++)
                if (pos[x] >= 0) last[x] = (int)d;
        }
        bool ok0 = true;
        for (int x = 0; x < 3; x++)
            if (last[x] < 0 && fx[x] != want[x]) ok0 = false;
        std::vector<unsigned> cur(gl.size(), 0);
        std::function<void(size_t, unsigned, unsigned, unsigned)> dfs = [&](size_t d, unsigned p0, unsigned p1,
                                                                            unsigned p2) {
            if (d == gl.size())
            {
                assigns.push_back(cur);
                return;
            }
            const Lab& x = labs[gl[d]];
            for (int r = 0; r < nir; r++)
            {
                const unsigned q[3] = {x.pa >= 0 ? p0 ^ r : p0, x.pb >= 0 ? p1 ^ r : p1, x.pc >= 0 ? p2 ^ r : p2};
                bool ok = true;
                for (int t = 0; t < 3; t++)
                    if (last[t] == (int)d && q[t] != want[t]) ok = false;
                if (!ok) continue;
                cur[d] = (unsigned)r;
                dfs(d + 1, q[0], q[1], q[2]);
            }
        };
        if (ok0) dfs(0, fx[0], fx[1], fx[2]);
    }
    std::vector<int64_t> val(labs.size());
    std::vector<char> fixed(labs.size());
    std::vector<unsigned> irr_a(A.ndd), irr_b(B.ndd), irr_c(C.ndd);
    std::vector<int64_t> alen(A.ndd), ast(A.ndd), blen_(B.ndd), bst(B.ndd), clen(C.ndd), cst(C.ndd);
    for (const std::vector<unsigned>& asg : assigns)
    {
        for (size_t d = 0; d < gl.size(); d++) lirr[gl[d]] = asg[d];
        for (size_t i = 0; i < labs.size(); i++)
        {
            const Lab& x = labs[i];
            if (x.pa >= 0 && A.dense(x.pa)) irr_a[x.pa] = lirr[i];
            if (x.pb >= 0 && B.dense(x.pb)) irr_b[x.pb] = lirr[i];
            if (x.pc >= 0 && C.dense(x.pc)) irr_c[x.pc] = lirr[i];
        }
        int64_t boa, bob, boc;
        A.block(irr_a.data(), alen.data(), ast.data(), boa);
        B.block(irr_b.data(), blen_.data(), bst.data(), bob);
        C.block(irr_c.data(), clen.data(), cst.data(), boc);
        Shape sh;
        bool empty = false;
        for (size_t g : gl)
        {
            const Lab& x = labs[g];
            const int64_t len = x.pa >= 0 ? alen[x.pa] : blen_[x.pb];
            if (len == 0) empty = true;
            if (len == 1) continue;
            if (x.pa >= 0 && x.pb >= 0 && x.pc >= 0)
            {
                sh.blen.push_back(len);
                sh.bsa.push_back(ast[x.pa]);
                sh.bsb.push_back(bst[x.pb]);
                sh.bsc.push_back(cst[x.pc]);
                sh.NB *= len;
            }
            else if (x.pa >= 0 && x.pb >= 0)
            {
                sh.kA.len.push_back(len), sh.kA.stride.push_back(ast[x.pa]);
                sh.kB.len.push_back(len), sh.kB.stride.push_back(bst[x.pb]);
                sh.K *= len;
            }
            else if (x.pa >= 0)
            {
                sh.mA.len.push_back(len), sh.mA.stride.push_back(ast[x.pa]);
                sh.mC.len.push_back(len), sh.mC.stride.push_back(cst[x.pc]);
                sh.M *= len;
            }
            else
            {
                sh.nB.len.push_back(len), sh.nB.stride.push_back(bst[x.pb]);
                sh.nC.len.push_back(len), sh.nC.stride.push_back(cst[x.pc]);
                sh.N *= len;
            }
        }
        if (empty) continue;
        const int32_t sid = (int32_t)shapes.size();
        shapes.push_back(sh);

        for (int64_t a = 0; a < A.n; a++)
        {
            if (A.zero(a)) continue;
            key.clear();
            for (const Lab* x : kb) key.push_back(A.value(a, x->pa));
            auto bit = bmap.find(key);
            if (bit == bmap.end()) continue;
            for (int64_t b : bit->second)
            {
                if (B.zero(b)) continue;
                for (size_t i = 0; i < labs.size(); i++)
                {
                    const Lab& x = labs[i];
                    fixed[i] = 0;
                    if (x.pa >= 0 && !A.dense(x.pa))
                        val[i] = A.value(a, x.pa), fixed[i] = 1;
                    else if (x.pb >= 0 && !B.dense(x.pb))
                        val[i] = B.value(b, x.pb), fixed[i] = 1;
                }
                key.clear();
                for (const Lab* x : kc) key.push_back(val[x - labs.data()]);
                auto cit = cmap.find(key);
                if (cit == cmap.end()) continue;
                for (int64_t c : cit->second)
                {
                    // f_a f_b f_c (complex for complex types)
                    const double abr = A.fac[a] * B.fac[b] - A.fim[a] * B.fim[b];
                    const double abi = A.fac[a] * B.fim[b] + A.fim[a] * B.fac[b];
                    const double coef = abr * C.fac[c] - abi * C.fim[c];
                    const double coef_im = abr * C.fim[c] + abi * C.fac[c];
                    if (coef == 0.0 && coef_im == 0.0) continue;
                    int64_t oa = A.off[a] + boa, ob = B.off[b] + bob, oc = C.off[c] + boc;
                    for (size_t i = 0; i < labs.size(); i++)
                    {
                        const Lab& x = labs[i];
                        int64_t v = val[i];
                        bool f = fixed[i];
                        if (!f && x.pc >= 0 && !C.dense(x.pc)) v = C.value(c, x.pc), f = true;
                        if (!f) continue;  // dense everywhere: a GEMM dimension
                        if (x.pa >= 0 && A.dense(x.pa)) oa += v * ast[x.pa];
                        if (x.pb >= 0 && B.dense(x.pb)) ob += v * bst[x.pb];
                        if (x.pc >= 0 && C.dense(x.pc)) oc += v * cst[x.pc];
                    }
                    auto w = where.find(oc);
                    size_t gi;
                    if (w == where.end())
                    {
                        gi = groups.size();
                        where[oc] = gi;
                        groups.push_back(Grp{oc, sid, {}});
                    }
                    else
                    {
                        gi = w->second;
                        const Shape& s0 = shapes[groups[gi].sid];
                        if (s0.M != sh.M || s0.N != sh.N || s0.NB != sh.NB)
                            fail("internal: indexed C region reached with different shapes");
                    }
                    groups[gi].k.push_back(KBlk{oa, ob, coef, coef_im, sid});
                }
            }
        }
    }
    if (!sliced && !C.dpd)
        for (int64_t e = 0; e < C.n; e++) covered[e] = where.count(C.off[e]) ? 1 : 0;
    ip->whole = !sliced && !C.dpd;
    upload();
    if (groups.empty()) return ip;

    PlanBuilder pb(es, dtype, family);
    if (cplx) pb.set_complex(sa, sb);
    for (const Grp& gr : groups)
    {
        const Shape& cs = shapes[gr.sid];
        std::vector<int64_t> bi(cs.blen.size(), 0);
        for (int64_t t = 0; t < cs.NB; t++)
        {
            int64_t dc = 0;
            for (size_t i = 0; i < bi.size(); i++) dc += bi[i] * cs.bsc[i];
            PanelGemm g;
            g.mlen = {cs.M};
            g.nlen = {cs.N};
            for (const KBlk& kbk : gr.k)
            {
                const Shape& s = shapes[kbk.sid];
                int64_t da = 0, db = 0;
                for (size_t i = 0; i < bi.size(); i++)
                {
                    da += bi[i] * s.bsa[i];
                    db += bi[i] * s.bsb[i];
                }
                g.klen.push_back(s.K);
                g.a.push_back(PBlock{kbk.oa + da, s.mA, s.kA});
                g.b.push_back(PBlock{kbk.ob + db, s.nB, s.kB});
                g.kscale.push_back(kbk.coef);
                if (cplx) g.kscale_im.push_back(kbk.coef_im);
            }
            CBlock cb;
            cb.base = gr.oc + dc;
            cb.rows = cs.mC;
            cb.cols = cs.nC;
            g.c.push_back(cb);
            pb.add_panel(g);
            for (size_t i = 0; i < bi.size(); i++)
            {
                if (++bi[i] < cs.blen[i]) break;
                bi[i] = 0;
            }
        }
    }
    ip->gemm = pb.build();
    return ip;
}

void check_entries(const Ix& x, const char* nm, int es)
{
    if (x.hi < x.lo) return;
    for (int64_t e = 0; e < x.n; e++) check_device_pointer(x.data[e], nm, x.lo, x.hi, es);
}

// A real call: C_e <- f_e beta C_e (beta == 0: zero; skipped when beta ==
// 1) in one pass over the C entries, then the GEMMs add alpha times their
// sums; when every GEMM writes whole C entries and beta == 0 the epilogue
// overwrites those and the pass zeroes only the others.
void run_real(IndexedPlan& ip, const Ix& a, const Ix& b, const Ix& c, type_t type, const tblis_scalar& sA,
              const tblis_scalar& sB, const tblis_scalar& sC, int es)
{
    const double alpha = product_scalar(sA, sB, type);
    const double beta = scalar_value(sC, type);
    ThreadState& ts = thread_state();
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    hip_ok(hipStreamIsCapturing(ts.stream, &cap), "hipStreamIsCapturing");
    const bool capturing = cap != hipStreamCaptureStatusNone;
    if (capturing) ip.pinned = true;

    const bool gemm = alpha != 0.0 && ip.gemm && ip.gemm->ntasks > 0;
    const bool epilogue_zero = gemm && beta == 0.0 && ip.whole;
    const int64_t npass = epilogue_zero ? ip.nunc : ip.nent;
    const bool scale = beta != 1.0 && npass > 0 && ip.cdesc.nb > 0;
    if (!scale && !gemm) return;
    const void* a0 = a.n ? a.data[0] : nullptr;
    const void* b0 = b.n ? b.data[0] : nullptr;
    void* c0 = c.data[0];
    Plan* vp = ip.gemm ? ip.gemm.get() : nullptr;
    if (!vp || !vp->already_verified(a0, b0, c0, gemm))
    {
        check_entries(c, "C", es);
        if (gemm)
        {
            check_entries(a, "A", es);
            check_entries(b, "B", es);
        }
        if (vp) vp->mark_verified(a0, b0, c0, gemm);
    }
    if (scale)
    {
        const int64_t* offs = (const int64_t*)ip.d_tab + (epilogue_zero ? 2 * ip.nent : 0);
        const double* fac = (const double*)(offs + npass);
        hip_ok(launch_level1_entries(type == TYPE_DOUBLE ? 0 : 1, ip.cdesc, offs, fac, npass, beta, c0, ts.stream),
               "indexed beta pass");
    }
    if (gemm)
        run_plan(ip.gemm, a0, b0, c0, alpha, epilogue_zero ? 0.0 : 1.0, true);
    else if (!ts.async && !capturing)
        hip_ok(hipStreamSynchronize(ts.stream), "hipStreamSynchronize");
}

// A complex call: C_e <- g_e C_e before the GEMMs (g_e = f_e beta, or 1
// when beta == 1, as the real rule; divided by alpha when alpha is not
// real), the GEMMs accumulating sum f_a f_b f_c a.b, then C_e <- alpha C_e
// for a complex alpha (a real alpha goes to the GEMM epilogue).
void run_complex(IndexedPlan& ip, const Ix& a, const Ix& b, const Ix& c, type_t type, const tblis_scalar& sA,
                 const tblis_scalar& sB, const tblis_scalar& sC, int es, int dt)
{
    const bool z = type == TYPE_DCOMPLEX;
    auto part = [&](const tblis_scalar& s, int i) { return z ? s.data.z[i] : (double)s.data.c[i]; };
    const double ar0 = part(sA, 0), ai0 = part(sA, 1), br0 = part(sB, 0), bi0 = part(sB, 1);
    double ar = ar0 * br0 - ai0 * bi0, ai = ar0 * bi0 + ai0 * br0;
    if (!z) ar = (double)(float)ar, ai = (double)(float)ai;
    const double betar = part(sC, 0), betai = part(sC, 1);
    ThreadState& ts = thread_state();
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    hip_ok(hipStreamIsCapturing(ts.stream, &cap), "hipStreamIsCapturing");
    const bool capturing = cap != hipStreamCaptureStatusNone;
    if (capturing) ip.pinned = true;
    const bool gemm = (ar != 0.0 || ai != 0.0) && ip.gemm && ip.gemm->ntasks > 0;
    const bool beta_one = betar == 1.0 && betai == 0.0;
    const bool calpha = gemm && ai != 0.0;
    const int64_t n = ip.nent;
    if (!gemm && (beta_one || n == 0 || ip.cdesc.nb == 0)) return;
    const void* a0 = a.n ? a.data[0] : nullptr;
    const void* b0 = b.n ? b.data[0] : nullptr;
    void* c0 = c.data[0];
    Plan* vp = ip.gemm ? ip.gemm.get() : nullptr;
    if (!vp || !vp->already_verified(a0, b0, c0, gemm))
    {
        check_entries(c, "C", es);
        if (gemm)
        {
            check_entries(a, "A", es);
            check_entries(b, "B", es);
        }
        if (vp) vp->mark_verified(a0, b0, c0, gemm);
    }
    const int64_t* offs = (const int64_t*)ip.d_tab;
    const double* fre = offs ? (const double*)(offs + n) : nullptr;
    const double* fim = fre ? fre + n : nullptr;
    // pre pass
    double sr = betar, si = betai;
    if (calpha)
    {
        const double d = ar * ar + ai * ai, ir = ar / d, ii = -ai / d;
        if (beta_one)
            sr = ir, si = ii;
        else
            sr = betar * ir - betai * ii, si = betar * ii + betai * ir;
    }
    if (calpha || !beta_one)
        hip_ok(launch_level1_centries(dt, ip.cdesc, offs, calpha && beta_one ? nullptr : fre, fim, n, sr, si, c0,
                                      ts.stream),
               "indexed beta pass");
    if (gemm)
    {
        {
            struct Restore
            {
                ThreadState& t;
                bool async;
                ~Restore() { t.async = async; }
            } restore{ts, ts.async};
            if (calpha) ts.async = true;
            run_plan(ip.gemm, a0, b0, c0, calpha ? 1.0 : ar, 1.0, true);
        }
        if (calpha)
            hip_ok(launch_level1_centries(dt, ip.cdesc, offs, nullptr, nullptr, n, ar, ai, c0, ts.stream),
                   "indexed alpha pass");
    }
    if ((!gemm || calpha) && !ts.async && !capturing) hip_ok(hipStreamSynchronize(ts.stream), "hipStreamSynchronize");
}

// Plan lookup (or build) and run of a loaded contraction; resets C's scalar.
void contract(const tblis_config* cfg, const Ix& a, const Ix& b, const Ix& c, type_t type, int conj_a, int conj_b,
              const tblis_scalar& sA, const tblis_scalar& sB, tblis_scalar& sC, int& conj_c)
{
    const bool cplx = type == TYPE_SCOMPLEX || type == TYPE_DCOMPLEX;
    const type_t rtype = type == TYPE_DOUBLE || type == TYPE_DCOMPLEX ? TYPE_DOUBLE : TYPE_SINGLE;
    const int es = elem_size(type);
    const double sa = conj_a ? -1.0 : 1.0, sb = conj_b ? -1.0 : 1.0;
    Algo algo = resolve_algo(cfg, rtype);
    if (algo == Algo::Reference) algo = rtype == TYPE_DOUBLE ? Algo::Fp64Mfma : Algo::Bf16x3;
    thread_state().last_algo = algo_name(algo);
    thread_state().last_tile = -1;
    thread_state().last_direct = 0;

    std::string key;
    put(key, (int32_t)algo);
    put(key, (int32_t)type);
    put(key, sa);
    put(key, sb);
    key_of(key, a);
    key_of(key, b);
    key_of(key, c);
    std::shared_ptr<IndexedPlan> ip = cache_find(key);
    if (!ip)
    {
        ip = build(a, b, c, cplx ? es / 2 : es, rtype == TYPE_DOUBLE ? 0 : 1, algo_family(algo), cplx, sa, sb);
        cache_store(key, ip);
    }
    if (cplx)
        run_complex(*ip, a, b, c, type, sA, sB, sC, es, rtype == TYPE_DOUBLE ? 0 : 1);
    else
        run_real(*ip, a, b, c, type, sA, sB, sC, es);
    if (type == TYPE_SINGLE)
        sC.data.s = 1.0f;
    else if (type == TYPE_DOUBLE)
        sC.data.d = 1.0;
    else if (type == TYPE_SCOMPLEX)
        sC.data.c[0] = 1.0f, sC.data.c[1] = 0.0f;
    else
        sC.data.z[0] = 1.0, sC.data.z[1] = 0.0;
    conj_c = 0;
}

}  // namespace

void indexed_tensor_mult(const tblis_config* cfg, const tblis_indexed_tensor* A, const label_type* idx_A,
                         const tblis_indexed_tensor* B, const label_type* idx_B,
                         tblis_indexed_tensor* C, const label_type* idx_C)
{
    if (!A || !B || !C) fail("indexed: null descriptor");
    const type_t type = A->type;
    // complex types (src/iface/3t/mult.cxx:405-409 instantiates all four)
    const int es = elem_size(type);
    Ix a, b, c;
    load(a, A, idx_A, "A", type, es);
    load(b, B, idx_B, "B", type, es);
    load(c, C, idx_C, "C", type, es);

    // label rules (src/iface/3t/mult.cxx:305-388)
    auto count_in = [](const Ix& x, label_type l) { return x.pos_of(l) >= 0 ? 1 : 0; };
    for (const Ix* x : {&a, &b, &c})
        for (int i = 0; i < x->nd; i++)
        {
            const label_type l = x->idx[i];
            if (count_in(a, l) + count_in(b, l) + count_in(c, l) < 2)
                fail("indexed: label '" + std::string(1, l) + "' appears in only one tensor");
            int64_t len = x->len(i);
            for (const Ix* y : {&a, &b, &c})
            {
                const int p = y->pos_of(l);
                if (p >= 0 && y->len(p) != len)
                    fail("indexed: label '" + std::string(1, l) + "' has unequal lengths");
            }
        }
    contract(cfg, a, b, c, type, A->conj, B->conj, A->scalar, B->scalar, C->scalar, C->conj);
}

void indexed_dpd_blocked(const tblis_config* cfg, const tblis_indexed_dpd_tensor* A, const label_type* idx_A,
                         const tblis_indexed_dpd_tensor* B, const label_type* idx_B,
                         tblis_indexed_dpd_tensor* C, const label_type* idx_C)
{
    const type_t type = A->type;
    const int es = elem_size(type);
    Ix a, b, c;
    load_dpd(a, A, idx_A, "A", type, es);
    load_dpd(b, B, idx_B, "B", type, es);
    load_dpd(c, C, idx_C, "C", type, es);
    contract(cfg, a, b, c, type, A->conj, B->conj, A->scalar, B->scalar, C->scalar, C->conj);
}

}  // namespace tblis_amd
