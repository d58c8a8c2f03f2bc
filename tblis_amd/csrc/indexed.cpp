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
#include <map>
#include <memory>
#include <mutex>
#include <unordered_map>

#include "contract.hpp"
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

// Host view of one operand.
struct Ix
{
    const tblis_indexed_tensor* t = nullptr;
    const label_type* idx = nullptr;
    int nd = 0, ndd = 0, nix = 0;
    int64_t n = 0;
    std::vector<int64_t> off;  // element offset of each entry from data[0]
    std::vector<double> fac;
    std::vector<double> fim;   // imaginary parts of the factors (complex types)
    bool zero(int64_t e) const { return fac[e] == 0.0 && fim[e] == 0.0; }
    int64_t value(int64_t e, int pos) const { return t->indices[e * nix + (pos - ndd)]; }
    int64_t stride(int pos) const { return t->dense_stride[pos]; }
    int pos_of(label_type l) const
    {
        for (int i = 0; i < nd; i++)
            if (idx[i] == l) return i;
        return -1;
    }
    bool dense(int pos) const { return pos < ndd; }
};

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
    x.t = t;
    x.idx = idx;
    x.nd = (int)t->ndim;
    x.ndd = (int)t->dense_ndim;
    x.nix = x.nd - x.ndd;
    x.n = t->num_indices;
    if (x.nix > 0 && x.n > 0 && !t->indices) fail(name + ": null index array");
    if (x.nix == 0 && x.n > 1) fail(name + ": several entries without indexed dimensions");
    for (int i = 0; i < x.nd; i++)
    {
        if (t->len[i] < 0) fail(name + ": negative tensor length");
        for (int j = 0; j < i; j++)
            if (idx[i] == idx[j]) fail(name + ": repeated label '" + std::string(1, idx[i]) + "'");
    }
    x.off.resize(x.n);
    x.fac.resize(x.n);
    x.fim.assign(x.n, 0.0);
    for (int64_t e = 0; e < x.n; e++)
    {
        if (!t->data[e]) fail(name + ": null entry pointer");
        const intptr_t d = (const char*)t->data[e] - (const char*)t->data[0];
        if (d % es) fail(name + ": entry pointers are not aligned to the element size");
        x.off[e] = (int64_t)(d / es);
        x.fac[e] = !t->factor ? 1.0
                   : type == TYPE_DOUBLE ? ((const double*)t->factor)[e]
                   : type == TYPE_SINGLE ? (double)((const float*)t->factor)[e]
                   : type == TYPE_DCOMPLEX ? ((const double*)t->factor)[2 * e]
                                           : (double)((const float*)t->factor)[2 * e];
        if (t->factor && type == TYPE_DCOMPLEX) x.fim[e] = ((const double*)t->factor)[2 * e + 1];
        if (t->factor && type == TYPE_SCOMPLEX) x.fim[e] = (double)((const float*)t->factor)[2 * e + 1];
        for (int p = x.ndd; p < x.nd; p++)
        {
            const int64_t v = x.value(e, p);
            if (v < 0 || v >= t->len[p]) fail(name + ": index value out of range");
        }
    }
}

void key_of(std::string& k, const Ix& x)
{
    put(k, (int32_t)x.nd);
    put(k, (int32_t)x.ndd);
    put(k, x.n);
    k.append((const char*)x.t->len, sizeof(len_type) * x.nd);
    k.append((const char*)x.t->dense_stride, sizeof(stride_type) * x.ndd);
    k.append(x.idx, x.nd);
    if (x.n > 0 && x.nix > 0) k.append((const char*)x.t->indices, sizeof(len_type) * x.n * x.nix);
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

// Dense dimensions of C, smallest stride first, as a level-1 descriptor.
Level1Desc dense_desc(const Ix& c)
{
    std::vector<int> order(c.ndd);
    for (int i = 0; i < c.ndd; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return std::llabs(c.stride(a)) < std::llabs(c.stride(b));
    });
    Level1Desc d;
    for (int i : order)
    {
        if (c.t->len[i] == 1) continue;
        d.nb *= c.t->len[i];
        // fold into the previous dimension when contiguous with it
        if (d.ndb > 0 && c.stride(i) == d.sb[d.ndb - 1] * d.len_b[d.ndb - 1])
        {
            d.len_b[d.ndb - 1] *= c.t->len[i];
            continue;
        }
        if (d.ndb >= LEVEL1_MAXDIM) fail("indexed: more than 16 dense dimensions");
        d.len_b[d.ndb] = c.t->len[i];
        d.sb[d.ndb] = c.stride(i);
        d.sa_b[d.ndb] = 0;
        d.ndb++;
    }
    for (int i = 0; i < c.ndd; i++)
        if (c.t->len[i] == 0) d.nb = 0;
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
    Side mA, mC, nB, nC, kA, kB;
    std::vector<int64_t> blen, bsa, bsb, bsc;  // ABC batch
    for (const Lab& x : labs)
    {
        if (!all_dense(x)) continue;
        const int64_t len = x.pa >= 0 ? A.t->len[x.pa] : B.t->len[x.pb];
        if (len == 1) continue;
        if (x.pa >= 0 && x.pb >= 0 && x.pc >= 0)
        {
            blen.push_back(len);
            bsa.push_back(A.stride(x.pa));
            bsb.push_back(B.stride(x.pb));
            bsc.push_back(C.stride(x.pc));
        }
        else if (x.pa >= 0 && x.pb >= 0)
        {
            kA.len.push_back(len), kA.stride.push_back(A.stride(x.pa));
            kB.len.push_back(len), kB.stride.push_back(B.stride(x.pb));
        }
        else if (x.pa >= 0)
        {
            mA.len.push_back(len), mA.stride.push_back(A.stride(x.pa));
            mC.len.push_back(len), mC.stride.push_back(C.stride(x.pc));
        }
        else
        {
            nB.len.push_back(len), nB.stride.push_back(B.stride(x.pb));
            nC.len.push_back(len), nC.stride.push_back(C.stride(x.pc));
        }
    }
    auto prod = [](const std::vector<int64_t>& v) {
        int64_t p = 1;
        for (int64_t x : v) p *= x;
        return p;
    };
    const int64_t M = prod(mA.len), N = prod(nB.len), K = prod(kA.len), NBATCH = prod(blen);
    if (M == 0 || N == 0 || K == 0 || NBATCH == 0 || A.n == 0 || B.n == 0 || C.n == 0)
    {
        upload();
        return ip;
    }
    // C is sliced when a label dense in C is fixed by an A or B entry; then
    // a GEMM writes part of a C entry, otherwise all of it
    bool sliced = false;
    for (const Lab& x : labs)
        if (x.pc >= 0 && C.dense(x.pc) && !all_dense(x)) sliced = true;

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

    // C blocks, in order of first appearance: (C offset) -> k blocks
    struct KBlk
    {
        int64_t oa, ob;
        double coef, coef_im;
    };
    std::map<int64_t, size_t> where;
    std::vector<int64_t> cbase;
    std::vector<std::vector<KBlk>> kblocks;
    std::vector<int64_t> val(labs.size());
    std::vector<char> fixed(labs.size());
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
                int64_t oa = A.off[a], ob = B.off[b], oc = C.off[c];
                for (size_t i = 0; i < labs.size(); i++)
                {
                    const Lab& x = labs[i];
                    int64_t v = val[i];
                    bool f = fixed[i];
                    if (!f && x.pc >= 0 && !C.dense(x.pc)) v = C.value(c, x.pc), f = true;
                    if (!f) continue;  // dense everywhere: a GEMM dimension
                    if (x.pa >= 0 && A.dense(x.pa)) oa += v * A.stride(x.pa);
                    if (x.pb >= 0 && B.dense(x.pb)) ob += v * B.stride(x.pb);
                    if (x.pc >= 0 && C.dense(x.pc)) oc += v * C.stride(x.pc);
                }
                auto w = where.find(oc);
                size_t gi;
                if (w == where.end())
                {
                    gi = kblocks.size();
                    where[oc] = gi;
                    cbase.push_back(oc);
                    kblocks.emplace_back();
                }
                else
                    gi = w->second;
                kblocks[gi].push_back(KBlk{oa, ob, coef, coef_im});
            }
        }
    }
    if (!sliced)
        for (int64_t e = 0; e < C.n; e++) covered[e] = where.count(C.off[e]) ? 1 : 0;
    ip->whole = !sliced;
    upload();
    if (kblocks.empty()) return ip;

    PlanBuilder pb(es, dtype, family);
    if (cplx) pb.set_complex(sa, sb);
    std::vector<int64_t> bi(blen.size(), 0);
    for (int64_t t = 0; t < NBATCH; t++)
    {
        int64_t da = 0, db = 0, dc = 0;
        for (size_t i = 0; i < blen.size(); i++)
        {
            da += bi[i] * bsa[i];
            db += bi[i] * bsb[i];
            dc += bi[i] * bsc[i];
        }
        for (size_t gi = 0; gi < kblocks.size(); gi++)
        {
            PanelGemm g;
            g.mlen = {M};
            g.nlen = {N};
            for (const KBlk& kbk : kblocks[gi])
            {
                g.klen.push_back(K);
                g.a.push_back(PBlock{kbk.oa + da, mA, kA});
                g.b.push_back(PBlock{kbk.ob + db, nB, kB});
                g.kscale.push_back(kbk.coef);
                if (cplx) g.kscale_im.push_back(kbk.coef_im);
            }
            CBlock cb;
            cb.base = cbase[gi] + dc;
            cb.rows = mC;
            cb.cols = nC;
            g.c.push_back(cb);
            pb.add_panel(g);
        }
        for (size_t i = 0; i < blen.size(); i++)
        {
            if (++bi[i] < blen[i]) break;
            bi[i] = 0;
        }
    }
    ip->gemm = pb.build();
    return ip;
}

// Offsets [lo, hi] of an entry's dense part.
void dense_extent(const Ix& x, int64_t& lo, int64_t& hi)
{
    lo = hi = 0;
    for (int i = 0; i < x.ndd; i++)
    {
        if (x.t->len[i] == 0)
        {
            lo = 0, hi = -1;
            return;
        }
        const int64_t s = (x.t->len[i] - 1) * x.stride(i);
        if (s < 0)
            lo += s;
        else
            hi += s;
    }
}

void check_entries(const Ix& x, const char* nm, int es)
{
    int64_t lo, hi;
    dense_extent(x, lo, hi);
    if (hi < lo) return;
    for (int64_t e = 0; e < x.n; e++) check_device_pointer(x.t->data[e], nm, lo, hi, es);
}

// A complex call: C_e <- g_e C_e before the GEMMs (g_e = f_e beta, or 1
// when beta == 1, as the real rule; divided by alpha when alpha is not
// real), the GEMMs accumulating sum f_a f_b f_c a.b, then C_e <- alpha C_e
// for a complex alpha (a real alpha goes to the GEMM epilogue).
void run_complex(IndexedPlan& ip, const Ix& a, const Ix& b, const Ix& c, const tblis_indexed_tensor* A,
                 const tblis_indexed_tensor* B, tblis_indexed_tensor* C, int es, int dt)
{
    const bool z = A->type == TYPE_DCOMPLEX;
    auto part = [&](const tblis_scalar& s, int i) { return z ? s.data.z[i] : (double)s.data.c[i]; };
    const double ar0 = part(A->scalar, 0), ai0 = part(A->scalar, 1), br0 = part(B->scalar, 0),
                 bi0 = part(B->scalar, 1);
    double ar = ar0 * br0 - ai0 * bi0, ai = ar0 * bi0 + ai0 * br0;
    if (!z) ar = (double)(float)ar, ai = (double)(float)ai;
    const double betar = part(C->scalar, 0), betai = part(C->scalar, 1);
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
    const void* a0 = a.n ? A->data[0] : nullptr;
    const void* b0 = b.n ? B->data[0] : nullptr;
    void* c0 = C->data[0];
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

}  // namespace

void indexed_tensor_mult(const tblis_config* cfg, const tblis_indexed_tensor* A, const label_type* idx_A,
                         const tblis_indexed_tensor* B, const label_type* idx_B,
                         tblis_indexed_tensor* C, const label_type* idx_C)
{
    if (!A || !B || !C) fail("indexed: null descriptor");
    const type_t type = A->type;
    // complex types (src/iface/3t/mult.cxx:405-409 instantiates all four)
    const bool cplx = type == TYPE_SCOMPLEX || type == TYPE_DCOMPLEX;
    const type_t rtype = type == TYPE_DOUBLE || type == TYPE_DCOMPLEX ? TYPE_DOUBLE : TYPE_SINGLE;
    const int es = elem_size(type);
    const double sa = A->conj ? -1.0 : 1.0, sb = B->conj ? -1.0 : 1.0;
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
            int64_t len = x->t->len[i];
            for (const Ix* y : {&a, &b, &c})
            {
                const int p = y->pos_of(l);
                if (p >= 0 && y->t->len[p] != len)
                    fail("indexed: label '" + std::string(1, l) + "' has unequal lengths");
            }
        }

    Algo algo = resolve_algo(cfg, rtype);
    if (algo == Algo::Reference) algo = rtype == TYPE_DOUBLE ? Algo::Fp64Mfma : Algo::Bf16x3;
    thread_state().last_algo = algo_name(algo);
    thread_state().last_tile = -1;

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

    ThreadState& ts = thread_state();
    if (cplx)
    {
        run_complex(*ip, a, b, c, A, B, C, es, rtype == TYPE_DOUBLE ? 0 : 1);
        if (type == TYPE_DCOMPLEX)
            C->scalar.data.z[0] = 1.0, C->scalar.data.z[1] = 0.0;
        else
            C->scalar.data.c[0] = 1.0f, C->scalar.data.c[1] = 0.0f;
        C->conj = 0;
        return;
    }
    const double alpha = product_scalar(A->scalar, B->scalar, type);
    const double beta = scalar_value(C->scalar, type);
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    hip_ok(hipStreamIsCapturing(ts.stream, &cap), "hipStreamIsCapturing");
    const bool capturing = cap != hipStreamCaptureStatusNone;
    if (capturing) ip->pinned = true;

    const bool gemm = alpha != 0.0 && ip->gemm && ip->gemm->ntasks > 0;
    // beta == 0 with GEMMs that write whole C entries: the GEMM epilogue
    // overwrites those entries, the pass zeroes only the others
    const bool epilogue_zero = gemm && beta == 0.0 && ip->whole;
    const int64_t npass = epilogue_zero ? ip->nunc : ip->nent;
    const bool scale = beta != 1.0 && npass > 0 && ip->cdesc.nb > 0;
    if (scale || gemm)
    {
        const void* a0 = a.n ? A->data[0] : nullptr;
        const void* b0 = b.n ? B->data[0] : nullptr;
        void* c0 = C->data[0];
        Plan* vp = ip->gemm ? ip->gemm.get() : nullptr;
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
            const int64_t* offs = (const int64_t*)ip->d_tab + (epilogue_zero ? 2 * ip->nent : 0);
            const double* fac = (const double*)(offs + npass);
            hip_ok(launch_level1_entries(type == TYPE_DOUBLE ? 0 : 1, ip->cdesc, offs, fac, npass, beta, c0,
                                         ts.stream),
                   "indexed beta pass");
        }
        if (gemm)
            run_plan(ip->gemm, a0, b0, c0, alpha, epilogue_zero ? 0.0 : 1.0, true);
        else if (!ts.async && !capturing)
            hip_ok(hipStreamSynchronize(ts.stream), "hipStreamSynchronize");
    }
    if (type == TYPE_SINGLE)
        C->scalar.data.s = 1.0f;
    else
        C->scalar.data.d = 1.0;
    C->conj = 0;
}

}  // namespace tblis_amd
