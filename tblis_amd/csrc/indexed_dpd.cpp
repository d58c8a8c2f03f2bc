// Indexed DPD contractions.  By default the blocked algorithm
// (indexed_dpd_blocked, indexed.cpp); with TBLIS_HIP_INDEXED_DPD=full the
// reference's FULL algorithm
// (src/internal/3t/indexed_dpd/mult.cxx:68-118), here: each operand's entries are
// expanded from DPD storage to dense arrays over the summed per-irrep
// extents (index values shifted by their irrep's offset), the indexed
// contraction (indexed.cpp) runs on those plain indexed tensors, and C's
// symmetry-allowed blocks are copied back.  The expansion and write-back are
// block copies on the device; the dense scratch lives in a grow-only buffer
// per (device, stream) that is never freed (a captured graph may refer to
// it).  The reference's front end tests irreps with A ^ (B != C)
// (src/iface/3t/mult.cxx:512); the FULL algorithm needs no such test, since
// products of mismatched irreps land outside C's allowed blocks.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>

#include "contract.hpp"
#include "dpd.hpp"
#include "indexed.hpp"
#include "level1.hpp"
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

struct Op
{
    const tblis_indexed_dpd_tensor* t = nullptr;
    const label_type* idx = nullptr;
    int nd = 0, ndd = 0, nix = 0, nir = 1;
    int64_t n = 0;
    unsigned dense_irrep = 0;
    DpdLayout L;                          // dense part
    std::vector<int64_t> full;            // summed extent per dimension
    std::vector<std::vector<int64_t>> io; // offset of each irrep per dimension
    std::vector<int64_t> off;             // entry offsets from data[0]
    int64_t dense_elems = 1;              // elements of one expanded entry
    int64_t len(int d, int r) const { return t->len[d * nir + r]; }
};

void load(Op& x, const tblis_indexed_dpd_tensor* t, const label_type* idx, const char* nm, type_t type, int es)
{
    const std::string name(nm);
    if (!t) fail(name + ": null descriptor");
    if (t->type != type) fail("A, B and C must have the same type");
    if (t->nirrep != 1 && t->nirrep != 2 && t->nirrep != 4 && t->nirrep != 8)
        fail(name + ": the number of irreps must be 1, 2, 4 or 8");
    if (t->irrep >= t->nirrep) fail(name + ": irrep out of range");
    if (t->dense_ndim > t->ndim) fail(name + ": more dense dimensions than dimensions");
    if (t->ndim > 0 && (!idx || !t->len)) fail(name + ": null label or length array");
    if (t->num_indices < 0) fail(name + ": negative number of entries");
    if (t->num_indices > 0 && !t->data) fail(name + ": null data array");
    x.t = t;
    x.idx = idx;
    x.nd = (int)t->ndim;
    x.ndd = (int)t->dense_ndim;
    x.nix = x.nd - x.ndd;
    x.nir = (int)t->nirrep;
    x.n = t->num_indices;
    if (x.ndd > BLOCK_COPY_MAXDIM) fail(name + ": more than 8 dense dimensions");
    if (x.nix > 0 && (!t->idx_irrep || (x.n > 0 && !t->indices))) fail(name + ": null index arrays");
    if (x.nix == 0 && x.n > 1) fail(name + ": several entries without indexed dimensions");
    x.dense_irrep = t->irrep;
    for (int j = 0; j < x.nix; j++)
    {
        if (t->idx_irrep[j] >= t->nirrep) fail(name + ": index irrep out of range");
        x.dense_irrep ^= t->idx_irrep[j];
    }
    x.full.assign(x.nd, 0);
    x.io.assign(x.nd, std::vector<int64_t>(x.nir, 0));
    for (int d = 0; d < x.nd; d++)
    {
        for (int j = 0; j < d; j++)
            if (idx[j] == idx[d]) fail(name + ": repeated label '" + std::string(1, idx[d]) + "'");
        for (int r = 0; r < x.nir; r++)
        {
            if (x.len(d, r) < 0) fail(name + ": negative tensor length");
            x.io[d][r] = x.full[d];
            x.full[d] += x.len(d, r);
        }
    }
    for (int d = 0; d < x.ndd; d++) x.dense_elems *= x.full[d];
    if (x.ndd > 0) x.L.init(x.dense_irrep, t->nirrep, (unsigned)x.ndd, t->len, t->layout);
    x.off.resize(x.n);
    for (int64_t e = 0; e < x.n; e++)
    {
        if (!t->data[e]) fail(name + ": null entry pointer");
        const intptr_t d = (const char*)t->data[e] - (const char*)t->data[0];
        if (d % es) fail(name + ": entry pointers are not aligned to the element size");
        x.off[e] = (int64_t)(d / es);
        for (int j = 0; j < x.nix; j++)
        {
            const int64_t v = t->indices[e * x.nix + j];
            if (v < 0 || v >= x.len(x.ndd + j, (int)t->idx_irrep[j])) fail(name + ": index value out of range");
        }
    }
}

// The allowed blocks of an operand's dense part: copies between each
// entry's DPD storage (offset from data[0]) and its dense scratch slot
// (offset base + e * dense_elems, column-major over the summed extents).
constexpr int64_t kPiece = 2048;  // elements per block-copy workgroup

void block_tasks(const Op& x, int64_t base, bool to_dense, std::vector<BlockCopy>& out)
{
    if (x.ndd == 0)
    {
        for (int64_t e = 0; e < x.n; e++)
        {
            BlockCopy b{};
            b.e0 = 0;
            b.n = 1;
            b.src = to_dense ? x.off[e] : base + e;
            b.dst = to_dense ? base + e : x.off[e];
            out.push_back(b);
        }
        return;
    }
    std::vector<int64_t> fs(x.ndd);
    int64_t acc = 1;
    for (int d = 0; d < x.ndd; d++)
    {
        fs[d] = acc;
        acc *= x.full[d];
    }
    std::vector<unsigned> irr(x.ndd, 0);
    const int64_t ncomb = (int64_t)1 << (x.ndd * (x.nir == 1 ? 0 : x.nir == 2 ? 1 : x.nir == 4 ? 2 : 3));
    for (int64_t c = 0; c < ncomb; c++)
    {
        unsigned xr = 0;
        int64_t rest = c;
        for (int d = 0; d < x.ndd; d++)
        {
            irr[d] = (unsigned)(rest % x.nir);
            rest /= x.nir;
            xr ^= irr[d];
        }
        if (xr != x.dense_irrep) continue;
        int64_t blen[DPD_MAXDIM], bst[DPD_MAXDIM], boff = 0, n = 1;
        x.L.block(irr.data(), blen, &boff, bst);
        int64_t doff = 0;
        for (int d = 0; d < x.ndd; d++)
        {
            n *= blen[d];
            doff += x.io[d][irr[d]] * fs[d];
        }
        if (n == 0) continue;
        for (int64_t e = 0; e < x.n; e++)
        {
            BlockCopy b{};
            b.nd = x.ndd;
            for (int d = 0; d < x.ndd; d++)
            {
                b.len[d] = blen[d];
                b.ss[d] = to_dense ? bst[d] : fs[d];
                b.ds[d] = to_dense ? fs[d] : bst[d];
            }
            b.src = to_dense ? x.off[e] + boff : base + e * x.dense_elems + doff;
            b.dst = to_dense ? base + e * x.dense_elems + doff : x.off[e] + boff;
            // pieces of kPiece elements, one workgroup each
            for (int64_t e0 = 0; e0 < n; e0 += kPiece)
            {
                b.e0 = e0;
                b.n = std::min<int64_t>(kPiece, n - e0);
                out.push_back(b);
            }
        }
    }
}

// Device copy tables of one shape (cached by the operands' structure).
struct Tables
{
    BlockCopy* d = nullptr;
    int64_t na = 0, nb = 0, nc = 0, nw = 0;  // expansion of A, B, C; write-back of C
    bool pinned = false;
    ~Tables()
    {
        if (d && !pinned) (void)hipFree(d);
    }
};

std::mutex g_mu;
auto* g_tables = new std::map<std::string, std::shared_ptr<Tables>>();

// Dense scratch per (device, stream), grow-only; an outgrown buffer is kept.
void* scratch(size_t bytes, hipStream_t s)
{
    static std::mutex mu;
    static auto* m = new std::map<std::pair<int, hipStream_t>, std::pair<void*, size_t>>();
    int dev = 0;
    hip_ok(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> g(mu);
    auto& e = (*m)[{dev, s}];
    if (e.second < bytes)
    {
        void* p = nullptr;
        hip_ok(device_alloc((void**)&p, bytes), "hipMalloc (indexed DPD scratch)");
        e = {p, bytes};
    }
    return e.first;
}

// TBLIS_HIP_INDEXED_DPD=full selects the FULL algorithm (read at each call).
bool full_algorithm()
{
    const char* e = std::getenv("TBLIS_HIP_INDEXED_DPD");
    return e && std::strcmp(e, "full") == 0;
}

}  // namespace

void indexed_dpd_tensor_mult(const tblis_config* cfg, const tblis_indexed_dpd_tensor* A, const label_type* idx_A,
                             const tblis_indexed_dpd_tensor* B, const label_type* idx_B,
                             tblis_indexed_dpd_tensor* C, const label_type* idx_C)
{
    if (!A || !B || !C) fail("indexed DPD: null descriptor");
    const type_t type = A->type;
    // all four types (src/iface/3t/mult.cxx:539-543); complex entries are
    // copied as (re, im) pairs and contracted by the complex indexed path
    const bool cplx = type == TYPE_SCOMPLEX || type == TYPE_DCOMPLEX;
    const int es = elem_size(type);
    Op a, b, c;
    load(a, A, idx_A, "A", type, es);
    load(b, B, idx_B, "B", type, es);
    load(c, C, idx_C, "C", type, es);
    if (a.nir != b.nir || a.nir != c.nir) fail("indexed DPD: A, B and C must have the same number of irreps");
    // labels shared by two tensors have equal per-irrep lengths
    for (const Op* x : {&a, &b, &c})
        for (const Op* y : {&a, &b, &c})
            for (int i = 0; i < x->nd; i++)
                for (int j = 0; j < y->nd; j++)
                    if (x->idx[i] == y->idx[j])
                        for (int r = 0; r < a.nir; r++)
                            if (x->len(i, r) != y->len(j, r))
                                fail("indexed DPD: label '" + std::string(1, x->idx[i]) + "' has unequal lengths");
    for (const Op* x : {&a, &b, &c})
        for (int i = 0; i < x->nd; i++)
        {
            int n = 0;
            for (const Op* y : {&a, &b, &c})
                for (int j = 0; j < y->nd; j++) n += y->idx[j] == x->idx[i];
            if (n < 2) fail("indexed DPD: label '" + std::string(1, x->idx[i]) + "' appears in only one tensor");
        }
    if (!full_algorithm())
    {
        indexed_dpd_blocked(cfg, A, idx_A, B, idx_B, C, idx_C);
        return;
    }

    // the plain indexed operands over the expanded entries
    const int64_t na = a.n * a.dense_elems, nb = b.n * b.dense_elems, nc = c.n * c.dense_elems;
    ThreadState& ts = thread_state();
    char* base = (char*)scratch((size_t)std::max<int64_t>(na + nb + nc, 1) * es, ts.stream);
    const int64_t oa = 0, ob = na, oc = na + nb;

    struct Plain
    {
        tblis_indexed_tensor d;
        std::vector<len_type> len, stride, indices;
        std::vector<void*> ptr;
    } pa, pb, pc;
    auto make = [&](Plain& p, const Op& x, int64_t at) {
        p.len = x.full;
        int64_t acc = 1;
        for (int d = 0; d < x.ndd; d++)
        {
            p.stride.push_back(acc);
            acc *= x.full[d];
        }
        for (int64_t e = 0; e < x.n; e++)
        {
            p.ptr.push_back(base + (at + e * x.dense_elems) * es);
            for (int j = 0; j < x.nix; j++)
                p.indices.push_back(x.io[x.ndd + j][x.t->idx_irrep[j]] + x.t->indices[e * x.nix + j]);
        }
        std::memset(&p.d, 0, sizeof(p.d));
        p.d.type = type;
        p.d.conj = x.t->conj;
        p.d.scalar = x.t->scalar;
        p.d.ndim = (unsigned)x.nd;
        p.d.dense_ndim = (unsigned)x.ndd;
        p.d.len = p.len.data();
        p.d.dense_stride = p.stride.data();
        p.d.num_indices = x.n;
        p.d.data = p.ptr.data();
        p.d.indices = p.indices.data();
        p.d.factor = x.t->factor;
    };
    make(pa, a, oa);
    make(pb, b, ob);
    make(pc, c, oc);

    std::string key;
    put(key, (int32_t)type);
    put(key, (int32_t)(A->conj * 2 + B->conj));
    for (const Op* x : {&a, &b, &c})
    {
        put(key, (int32_t)x->nd);
        put(key, (int32_t)x->ndd);
        put(key, x->t->irrep);
        put(key, x->t->nirrep);
        put(key, (int32_t)x->t->layout);
        put(key, x->n);
        key.append((const char*)x->t->len, sizeof(len_type) * x->nd * x->nir);
        if (x->nix) key.append((const char*)x->t->idx_irrep, sizeof(unsigned) * x->nix);
        key.append((const char*)x->off.data(), sizeof(int64_t) * x->n);
    }
    std::shared_ptr<Tables> tb;
    {
        std::lock_guard<std::mutex> g(g_mu);
        auto it = g_tables->find(key);
        if (it != g_tables->end()) tb = it->second;
    }
    if (!tb)
    {
        tb = std::make_shared<Tables>();
        std::vector<BlockCopy> all;
        block_tasks(a, oa, true, all);
        tb->na = (int64_t)all.size();
        block_tasks(b, ob, true, all);
        tb->nb = (int64_t)all.size() - tb->na;
        block_tasks(c, oc, true, all);
        tb->nc = (int64_t)all.size() - tb->na - tb->nb;
        block_tasks(c, oc, false, all);
        tb->nw = (int64_t)all.size() - tb->na - tb->nb - tb->nc;
        if (!all.empty())
        {
            hip_ok(device_alloc((void**)&tb->d, all.size() * sizeof(BlockCopy)), "hipMalloc (indexed DPD tables)");
            hip_ok(hipMemcpy(tb->d, all.data(), all.size() * sizeof(BlockCopy), hipMemcpyHostToDevice),
                   "hipMemcpy (indexed DPD tables)");
        }
        std::lock_guard<std::mutex> g(g_mu);
        if (g_tables->size() >= 128)
            for (auto it = g_tables->begin(); it != g_tables->end();)
                if (!it->second->pinned && it->second.use_count() == 1)
                    it = g_tables->erase(it);
                else
                    ++it;
        (*g_tables)[key] = tb;
    }
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    hip_ok(hipStreamIsCapturing(ts.stream, &cap), "hipStreamIsCapturing");
    const bool capturing = cap != hipStreamCaptureStatusNone;
    if (capturing) tb->pinned = true;

    // real type of the zero fill; block-copy element type
    const int dt = type == TYPE_DOUBLE || type == TYPE_DCOMPLEX ? 0 : 1;
    const int ct = type == TYPE_DCOMPLEX ? 2 : type == TYPE_SCOMPLEX ? 3 : dt;
    const int64_t rl = cplx ? 2 : 1;  // reals per element
    auto check = [&](const Op& x, const char* nm) {
        if (x.ndd == 0)
        {
            for (int64_t e = 0; e < x.n; e++) check_device_pointer(x.t->data[e], nm, 0, 0, es);
            return;
        }
        const int64_t hi = x.L.total() - 1;
        if (hi < 0) return;
        for (int64_t e = 0; e < x.n; e++) check_device_pointer(x.t->data[e], nm, 0, hi, es);
    };
    check(a, "A");
    check(b, "B");
    check(c, "C");
    if (na + nb + nc > 0)
    {
        // zero the scratch (positions outside the allowed blocks), expand
        Level1Desc z;
        z.ndb = 1;
        z.len_b[0] = rl * (na + nb + nc);
        z.sb[0] = 1;
        z.sa_b[0] = 0;
        z.nb = rl * (na + nb + nc);
        hip_ok(launch_level1(LEVEL1_SET, dt, z, nullptr, base, 0.0, 0.0, 0.0, nullptr, ts.stream),
               "indexed DPD zero fill");
        if (tb->na) hip_ok(launch_block_copy(ct, tb->d, tb->na, A->data[0], base, ts.stream), "indexed DPD expand A");
        if (tb->nb)
            hip_ok(launch_block_copy(ct, tb->d + tb->na, tb->nb, B->data[0], base, ts.stream), "indexed DPD expand B");
        if (tb->nc)
            hip_ok(launch_block_copy(ct, tb->d + tb->na + tb->nb, tb->nc, C->data[0], base, ts.stream),
                   "indexed DPD expand C");
    }
    {
        struct Restore
        {
            ThreadState& t;
            bool async;
            ~Restore() { t.async = async; }
        } restore{ts, ts.async};
        ts.async = true;
        indexed_tensor_mult(cfg, &pa.d, idx_A, &pb.d, idx_B, &pc.d, idx_C);
    }
    if (tb->nw)
        hip_ok(launch_block_copy(ct, tb->d + tb->na + tb->nb + tb->nc, tb->nw, base, C->data[0], ts.stream),
               "indexed DPD write-back");
    if (!ts.async && !capturing) hip_ok(hipStreamSynchronize(ts.stream), "hipStreamSynchronize");
    if (type == TYPE_SINGLE)
        C->scalar.data.s = 1.0f;
    else if (type == TYPE_DOUBLE)
        C->scalar.data.d = 1.0;
    else if (type == TYPE_SCOMPLEX)
        C->scalar.data.c[0] = 1.0f, C->scalar.data.c[1] = 0.0f;
    else
        C->scalar.data.z[0] = 1.0, C->scalar.data.z[1] = 0.0;
    C->conj = 0;
}

}  // namespace tblis_amd
