// Dense and batched contraction drivers, algorithm selection, introspection.
#include "contract.hpp"

#include <cstring>

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

void tensor_key(std::string& k, const tblis_tensor& t, const label_type* idx)
{
    put(k, (int32_t)t.type);
    put(k, (uint32_t)t.ndim);
    k.append((const char*)t.len, sizeof(len_type) * t.ndim);
    k.append((const char*)t.stride, sizeof(stride_type) * t.ndim);
    k.append(idx, t.ndim);
}

}  // namespace

int elem_size(type_t t)
{
    switch (t)
    {
        case TYPE_SINGLE: return 4;
        case TYPE_DOUBLE: return 8;
        case TYPE_SCOMPLEX: return 8;
        case TYPE_DCOMPLEX: return 16;
    }
    return 0;
}

void check_real(type_t t)
{
    if (t != TYPE_SINGLE && t != TYPE_DOUBLE)
        fail("complex tensor types are supported by the dense, DPD and batched contractions only");
}

double scalar_value(const tblis_scalar& s, type_t t)
{
    return t == TYPE_SINGLE ? (double)s.data.s : s.data.d;
}

// alpha = alpha_A * alpha_B in the tensor type (src/iface/3t/mult.cxx:123).
double product_scalar(const tblis_scalar& a, const tblis_scalar& b, type_t t)
{
    if (t == TYPE_SINGLE) return (double)(a.data.s * b.data.s);
    return a.data.d * b.data.d;
}

void reset_scalar(tblis_tensor* C)
{
    if (C->type == TYPE_SINGLE)
        C->scalar.data.s = 1.0f;
    else if (C->type == TYPE_DOUBLE)
        C->scalar.data.d = 1.0;
    else if (C->type == TYPE_SCOMPLEX)
        C->scalar.data.c[0] = 1.0f, C->scalar.data.c[1] = 0.0f;
    else
        C->scalar.data.z[0] = 1.0, C->scalar.data.z[1] = 0.0;
    C->conj = 0;
}

// Element offset range [lo, hi] spanned by a strided tensor (lo > hi if empty).
void tensor_extent(const tblis_tensor& t, int64_t& lo, int64_t& hi)
{
    lo = hi = 0;
    for (unsigned i = 0; i < t.ndim; i++)
    {
        if (t.len[i] == 0)
        {
            lo = 0;
            hi = -1;
            return;
        }
        int64_t span = (t.len[i] - 1) * t.stride[i];
        if (span < 0)
            lo += span;
        else
            hi += span;
    }
}

void check_data(const tblis_tensor* t, const char* name)
{
    int64_t lo, hi;
    tensor_extent(*t, lo, hi);
    check_device_pointer(t->data, name, lo, hi, elem_size(t->type));
}

void check_tensor(const tblis_tensor* t, const label_type* idx, const char* name)
{
    if (!t) fail(std::string(name) + ": null tensor descriptor");
    if (t->ndim > 0 && (!t->len || !t->stride || !idx))
        fail(std::string(name) + ": null length, stride or label array");
    if (t->ndim > 32) fail(std::string(name) + ": too many dimensions");
}

namespace
{

// Odometer offsets of every position of a (len, stride) list.
std::vector<int64_t> all_offsets(const Group& g, int slot)
{
    int64_t n = prod_len(g);
    std::vector<int64_t> out(std::max<int64_t>(n, 0));
    std::vector<int64_t> digit(g.size(), 0);
    int64_t off = 0;
    for (int64_t i = 0; i < n; i++)
    {
        out[i] = off;
        for (size_t d = 0; d < g.size(); d++)
        {
            if (digit[d] + 1 < g[d].len)
            {
                digit[d]++;
                off += g[d].s[slot];
                break;
            }
            off -= digit[d] * g[d].s[slot];
            digit[d] = 0;
        }
    }
    return out;
}

// The GEMM of one canonical dense contraction (m = AC, n = BC, k = AB,
// batch = ABC), with bases added for batched calls.
int dtype_code(type_t t) { return t == TYPE_DOUBLE ? 0 : 1; }

}  // namespace

bool dense_gemm(const Canonical& c, GemmDesc& g, int64_t base_a, int64_t base_b, int64_t base_c)
{
    MatrixViews v = blis_views(c);
    g.m = prod_len(v.m);
    g.n = prod_len(v.n);
    int64_t k = prod_len(v.k);
    int64_t L = prod_len(v.l);
    if (g.m == 0 || g.n == 0 || L == 0) return false;
    g.c_base = base_c;
    g.c_rows = side_of(v.m, 1, v.pack_m_3d);
    g.c_cols = side_of(v.n, 1, v.pack_n_3d);
    if (k > 0)
    {
        GemmSeg s;
        s.a_base = base_a;
        s.b_base = base_b;
        s.a_rows = side_of(v.m, 0, v.pack_m_3d);
        s.a_cols = side_of(v.k, 0, v.pack_k_3d);
        s.b_rows = side_of(v.k, 1, v.pack_k_3d);
        s.b_cols = side_of(v.n, 0, v.pack_n_3d);
        s.k = k;
        g.segs.push_back(s);
    }
    if (!v.l.empty())
    {
        g.l_a = all_offsets(v.l, 0);
        g.l_b = all_offsets(v.l, 1);
        g.l_c = all_offsets(v.l, 2);
    }
    return true;
}

const char* algo_name(Algo a)
{
    switch (a)
    {
        case Algo::Default: return "mi355x";
        case Algo::Fp64Mfma: return "fp64_mfma";
        case Algo::Bf16x3: return "bf16x3";
        case Algo::Bf16x3Fast: return "bf16x3_fast";
        case Algo::Reference: return "reference";
        case Algo::Ozaki: return "ozaki";
    }
    return "?";
}

Algo resolve_algo(const tblis_config* cfg, type_t type)
{
    // "reference" (the reference's portable configuration) selects the
    // same kernels as the default: every configuration runs on MFMA here
    Algo a = cfg ? cfg->algo : Algo::Default;
    if (a == Algo::Ozaki) return a;
    if (type == TYPE_DOUBLE) return Algo::Fp64Mfma;
    if (a == Algo::Default || a == Algo::Reference) return Algo::Bf16x3;
    return a;
}

namespace
{

int family_of(Algo a)
{
    if (a == Algo::Bf16x3) return FAMILY_BF16X3;
    if (a == Algo::Bf16x3Fast) return FAMILY_BF16X3_FAST;
    if (a == Algo::Ozaki) return FAMILY_OZAKI;
    return FAMILY_F64;
}

std::shared_ptr<Plan> empty_plan() { return std::make_shared<Plan>(); }

bool gemm_reads(const GemmDesc& g)
{
    for (const GemmSeg& s : g.segs)
        if (s.k > 0) return true;
    return false;
}

}  // namespace

int algo_family(Algo a) { return family_of(a); }

void tensor_mult(const tblis_config* cfg, const tblis_tensor* A, const label_type* idx_A,
                 const tblis_tensor* B, const label_type* idx_B, tblis_tensor* C,
                 const label_type* idx_C)
{
    check_tensor(A, idx_A, "A");
    check_tensor(B, idx_B, "B");
    check_tensor(C, idx_C, "C");
    if (A->type != B->type || A->type != C->type) fail("A, B and C must have the same type");
    if (A->type == TYPE_SCOMPLEX || A->type == TYPE_DCOMPLEX)
        return complex_tensor_mult(cfg, A, idx_A, B, idx_B, C, idx_C);
    check_real(A->type);
    const type_t type = A->type;
    Algo algo = resolve_algo(cfg, type);
    thread_state().last_algo = algo_name(algo);

    std::string key;
    put(key, (int32_t)algo);
    tensor_key(key, *A, idx_A);
    tensor_key(key, *B, idx_B);
    tensor_key(key, *C, idx_C);

    const double alpha = product_scalar(A->scalar, B->scalar, type);
    const double beta = scalar_value(C->scalar, type);

    std::shared_ptr<Plan> plan = cache_get(key);
    if (!plan)
    {
        Canonical c;
        std::string why;
        if (!canonicalize(*A, idx_A, *B, idx_B, *C, idx_C, c, why)) fail("invalid contraction: " + why);
        GemmDesc g;
        if (dense_gemm(c, g, 0, 0, 0))
        {
            // the autotuner may time its candidates on A and B (read only):
            // check them first
            const bool reads = alpha != 0.0 && gemm_reads(g);
            if (reads)
            {
                check_data(A, "A");
                check_data(B, "B");
            }
            TuneOperandsScope ops(reads ? A->data : nullptr, reads ? B->data : nullptr);
            PlanBuilder pb(elem_size(type), dtype_code(type), family_of(algo));
            pb.add(g);
            plan = pb.build();
        }
        if (!plan) plan = empty_plan();
        cache_put(key, plan);
    }

    if (plan->ntasks > 0)
    {
        const bool use_k = alpha != 0.0;
        if (use_k || beta != 1.0)
        {
            const bool k = use_k && plan->reads_ab;
            if (!plan->already_verified(A->data, B->data, C->data, k))
            {
                check_data(C, "C");
                if (k)
                {
                    check_data(A, "A");
                    check_data(B, "B");
                }
                plan->mark_verified(A->data, B->data, C->data, k);
            }
            run_plan(plan, A->data, B->data, C->data, alpha, beta, use_k);
        }
    }
    reset_scalar(C);
}

// The same contraction with its plan tables and workspace in the device
// scratch of noalloc.hip: a plan built for this call alone (no cache, the
// cost model's tile, no autotuning), uploaded in stream order into the
// scratch, then the panel launches.  Sizes whose tables and workspace
// exceed the scratch fail with an error; so do complex types.
void tensor_mult_noalloc(const tblis_config* cfg, const tblis_tensor* A, const label_type* idx_A,
                         const tblis_tensor* B, const label_type* idx_B, tblis_tensor* C,
                         const label_type* idx_C)
{
    check_tensor(A, idx_A, "A");
    check_tensor(B, idx_B, "B");
    check_tensor(C, idx_C, "C");
    if (A->type != B->type || A->type != C->type) fail("A, B and C must have the same type");
    if (A->type != TYPE_SINGLE && A->type != TYPE_DOUBLE)
        fail("the no-allocation entry point takes real (float, double) tensors only");
    const type_t type = A->type;
    const Algo algo = resolve_algo(cfg, type);
    thread_state().last_algo = algo_name(algo);
    const double alpha = product_scalar(A->scalar, B->scalar, type);
    const double beta = scalar_value(C->scalar, type);

    Canonical c;
    std::string why;
    if (!canonicalize(*A, idx_A, *B, idx_B, *C, idx_C, c, why)) fail("invalid contraction: " + why);
    GemmDesc g;
    if (dense_gemm(c, g, 0, 0, 0))
    {
        size_t bytes = 0;
        char* base = noalloc_scratch(bytes);
        AutotuneOffScope no_tune;
        StaticStorageScope storage(base, bytes);
        PlanBuilder pb(elem_size(type), dtype_code(type), family_of(algo));
        pb.add(g);
        std::shared_ptr<Plan> plan = pb.build();
        if (plan && plan->ntasks > 0)
        {
            const bool use_k = alpha != 0.0;
            if (use_k || beta != 1.0)
            {
                check_data(C, "C");
                if (use_k && plan->reads_ab)
                {
                    check_data(A, "A");
                    check_data(B, "B");
                }
                run_plan(plan, A->data, B->data, C->data, alpha, beta, use_k);
            }
        }
    }
    reset_scalar(C);
}

// Matrix multiplication (src/iface/3m/mult.cxx:19-191) is the contraction
// C[ab] = A[ak] B[kb] with each matrix's (rs, cs) as its two strides; the
// tensor front end then covers the reference's special cases (empty C,
// alpha == 0 or k == 0 scaling C by beta, length-1 dimensions).
void matrix_mult(const tblis_config* cfg, const tblis_matrix* A, const tblis_matrix* B,
                 tblis_matrix* C)
{
    if (!A || !B || !C) fail("matrix mult: null descriptor");
    if (A->m != C->m || B->n != C->n || A->n != B->m)
        fail("matrix mult: shapes do not agree (need A.m == C.m, B.n == C.n, A.n == B.m)");
    if (A->type != B->type || A->type != C->type) fail("A, B and C must have the same type");
    if (A->conj || B->conj || C->conj) fail("matrix mult: conjugation is not supported");
    if (A->m < 0 || A->n < 0 || B->n < 0) fail("matrix mult: negative length");
    len_type la[2] = {A->m, A->n}, lb[2] = {B->m, B->n}, lc[2] = {C->m, C->n};
    stride_type sa[2] = {A->rs, A->cs}, sb[2] = {B->rs, B->cs}, sc[2] = {C->rs, C->cs};
    tblis_tensor ta, tb, tc;
    auto make = [](tblis_tensor& t, const tblis_matrix* m, len_type* len, stride_type* str) {
        t.type = m->type;
        t.conj = 0;
        t.scalar = m->scalar;
        t.data = m->data;
        t.ndim = 2;
        t.len = len;
        t.stride = str;
    };
    make(ta, A, la, sa);
    make(tb, B, lb, sb);
    make(tc, C, lc, sc);
    tensor_mult(cfg, &ta, "ak", &tb, "kb", &tc, "ab");
    C->scalar = tc.scalar;
    C->conj = 0;
}

void tensor_mult_batch(const tblis_config* cfg, int64_t count, const tblis_tensor* const* A,
                       const label_type* const* idx_A, const tblis_tensor* const* B,
                       const label_type* const* idx_B, tblis_tensor* const* C,
                       const label_type* const* idx_C)
{
    if (count < 0) fail("batch: negative count");
    if (count == 0) return;
    if (!A || !B || !C || !idx_A || !idx_B || !idx_C) fail("batch: null array argument");
    check_tensor(A[0], idx_A[0], "A");
    check_tensor(B[0], idx_B[0], "B");
    check_tensor(C[0], idx_C[0], "C");
    const type_t type = A[0]->type;
    // complex types: each instance's GEMM as two real GEMMs over the
    // interleaved parts (the same split as the dense complex call)
    const bool cplx = type == TYPE_SCOMPLEX || type == TYPE_DCOMPLEX;
    const type_t rtype = type == TYPE_DOUBLE || type == TYPE_DCOMPLEX ? TYPE_DOUBLE : TYPE_SINGLE;
    const int es = elem_size(type);
    double alpha = 0, beta = 0, alpha_im = 0, beta_im = 0;
    const double sa = A[0]->conj ? -1.0 : 1.0, sb = B[0]->conj ? -1.0 : 1.0;
    for (int64_t i = 0; i < count; i++)
    {
        check_tensor(A[i], idx_A[i], "A");
        check_tensor(B[i], idx_B[i], "B");
        check_tensor(C[i], idx_C[i], "C");
        if (A[i]->type != type || B[i]->type != type || C[i]->type != type)
            fail("batch: every tensor must have the same type");
        if ((A[i]->conj ? -1.0 : 1.0) != sa || (B[i]->conj ? -1.0 : 1.0) != sb)
            fail("batch: every instance must conjugate the same operands");
        double al, ali = 0, be, bei = 0;
        if (cplx)
        {
            const bool z = type == TYPE_DCOMPLEX;
            const double ar = z ? A[i]->scalar.data.z[0] : A[i]->scalar.data.c[0];
            const double ai = z ? A[i]->scalar.data.z[1] : A[i]->scalar.data.c[1];
            const double br = z ? B[i]->scalar.data.z[0] : B[i]->scalar.data.c[0];
            const double bi = z ? B[i]->scalar.data.z[1] : B[i]->scalar.data.c[1];
            al = ar * br - ai * bi;
            ali = ar * bi + ai * br;
            if (!z) al = (double)(float)al, ali = (double)(float)ali;
            be = z ? C[i]->scalar.data.z[0] : C[i]->scalar.data.c[0];
            bei = z ? C[i]->scalar.data.z[1] : C[i]->scalar.data.c[1];
        }
        else
        {
            al = product_scalar(A[i]->scalar, B[i]->scalar, type);
            be = scalar_value(C[i]->scalar, type);
        }
        if (i == 0)
        {
            alpha = al;
            beta = be;
            alpha_im = ali;
            beta_im = bei;
        }
        else if (al != alpha || be != beta || ali != alpha_im || bei != beta_im)
            fail("batch: every instance must use the same alpha and beta");
    }
    Algo algo = resolve_algo(cfg, rtype);
    thread_state().last_algo = algo_name(algo);

    auto delta = [&](const void* p, const void* p0) -> int64_t {
        intptr_t d = (const char*)p - (const char*)p0;
        if (d % es) fail("batch: tensor data pointers are not aligned to the element size");
        return (int64_t)(d / es);
    };

    std::string key;
    put(key, (int32_t)algo);
    put(key, count);
    put(key, sa);
    put(key, sb);
    for (int64_t i = 0; i < count; i++)
    {
        tensor_key(key, *A[i], idx_A[i]);
        tensor_key(key, *B[i], idx_B[i]);
        tensor_key(key, *C[i], idx_C[i]);
        put(key, delta(A[i]->data, A[0]->data));
        put(key, delta(B[i]->data, B[0]->data));
        put(key, delta(C[i]->data, C[0]->data));
    }

    std::shared_ptr<Plan> plan = cache_get(key);
    if (!plan)
    {
        PlanBuilder pb(cplx ? es / 2 : es, dtype_code(rtype), family_of(algo));
        if (cplx) pb.set_complex(sa, sb);
        bool reads = false;
        for (int64_t i = 0; i < count; i++)
        {
            Canonical c;
            std::string why;
            if (!canonicalize(*A[i], idx_A[i], *B[i], idx_B[i], *C[i], idx_C[i], c, why))
                fail("batch: invalid contraction " + std::to_string(i) + ": " + why);
            GemmDesc g;
            if (dense_gemm(c, g, delta(A[i]->data, A[0]->data), delta(B[i]->data, B[0]->data),
                           delta(C[i]->data, C[0]->data)))
            {
                reads |= gemm_reads(g);
                pb.add(g);
            }
        }
        // the autotuner may time its candidates on the A and B instances
        // (read only): check them first
        reads &= alpha != 0.0 || alpha_im != 0.0;
        if (reads)
            for (int64_t i = 0; i < count; i++)
            {
                check_data(A[i], "A");
                check_data(B[i], "B");
            }
        TuneOperandsScope ops(reads ? A[0]->data : nullptr, reads ? B[0]->data : nullptr);
        DirectAllScope direct;
        plan = pb.build();
        if (!plan) plan = empty_plan();
        cache_put(key, plan);
    }
    if (plan->ntasks > 0)
    {
        const bool use_k = alpha != 0.0 || alpha_im != 0.0;
        if (use_k || beta != 1.0 || beta_im != 0.0)
        {
            const bool k = use_k && plan->reads_ab;
            if (!plan->already_verified(A[0]->data, B[0]->data, C[0]->data, k))
            {
                for (int64_t i = 0; i < count; i++)
                {
                    check_data(C[i], "C");
                    if (k)
                    {
                        check_data(A[i], "A");
                        check_data(B[i], "B");
                    }
                }
                plan->mark_verified(A[0]->data, B[0]->data, C[0]->data, k);
            }
            if (cplx)
            {
                // complex alpha or beta: one level-1 pass per instance's C
                std::vector<CPass> cps;
                if (alpha_im != 0.0 || beta_im != 0.0 || !use_k)
                    for (int64_t i = 0; i < count; i++) cps.push_back(CPass{C[i]->data, c_desc(*C[i], idx_C[i])});
                run_complex_plan(plan, A[0]->data, B[0]->data, C[0]->data, alpha, alpha_im, beta, beta_im, cps,
                                 rtype == TYPE_DOUBLE ? 0 : 1);
            }
            else
                run_plan(plan, A[0]->data, B[0]->data, C[0]->data, alpha, beta, use_k);
        }
    }
    for (int64_t i = 0; i < count; i++) reset_scalar(C[i]);
}

int64_t debug_canonical(const tblis_tensor* A, const label_type* idx_A, const tblis_tensor* B,
                        const label_type* idx_B, const tblis_tensor* C, const label_type* idx_C,
                        int64_t* out, int64_t max_out)
{
    check_tensor(A, idx_A, "A");
    check_tensor(B, idx_B, "B");
    check_tensor(C, idx_C, "C");
    Canonical c;
    std::string why;
    if (!canonicalize(*A, idx_A, *B, idx_B, *C, idx_C, c, why)) return -1;
    MatrixViews v = blis_views(c);
    std::vector<int64_t> o;
    auto emit = [&](const Group& g, int nslot) {
        o.push_back((int64_t)g.size());
        for (auto& d : g) o.push_back(d.len);
        for (int s = 0; s < nslot; s++)
            for (auto& d : g) o.push_back(d.s[s]);
    };
    emit(c.ab, 2);
    emit(c.ac, 2);
    emit(c.bc, 2);
    emit(c.abc, 3);
    emit(v.m, 2);
    emit(v.n, 2);
    emit(v.k, 2);
    emit(v.l, 3);
    o.push_back(v.pack_m_3d);
    o.push_back(v.pack_n_3d);
    o.push_back(v.pack_k_3d);
    if ((int64_t)o.size() > max_out) return -2;
    std::memcpy(out, o.data(), o.size() * sizeof(int64_t));
    return (int64_t)o.size();
}

}  // namespace tblis_amd
