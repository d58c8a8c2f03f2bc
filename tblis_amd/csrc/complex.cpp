// Complex contractions (scomplex / dcomplex) on the real panel path.
//
// An interleaved complex tensor with strides s (complex elements) is two
// real tensors: the real parts at real offset 0 with strides 2s, the
// imaginary parts at offset 1.  With P = op(A) op(B) (op: conjugation when
// the descriptor's conj flag is set, sa / sb = -1),
//     Re P = Ar Br - sa sb Ai Bi,    Im P = sb Ar Bi + sa Ai Br,
// and C <- alpha P + beta C splits into two real panel GEMMs, one writing
// Re C and one Im C, whose k blocks are the terms above, each scaled by its
// sign while packed.  A real alpha and beta go to the GEMM epilogue.  A
// complex beta is applied to C by a pass before the GEMMs, which then
// accumulate with beta = 1; a complex alpha by C <- (beta / alpha) C before
// and C <- alpha C after GEMMs that accumulate P (two passes over C instead
// of twice the MFMA work of folding alpha into the terms).  The
// canonical form, the reference's special cases (alpha == 0, k == 0, empty
// C) and the C <- scalar = 1 side effect are those of the real call
// (src/iface/3t/mult.cxx:24-165); the reference's mult_blis asserts that no
// operand is conjugated (src/internal/3t/dense/mult.cxx:212), which this
// library relaxes.
#include <algorithm>
#include <cstdlib>
#include <cstring>

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

struct Cplx
{
    double re = 0, im = 0;
};

Cplx scalar_of(const tblis_scalar& s, type_t t)
{
    if (t == TYPE_SCOMPLEX) return Cplx{s.data.c[0], s.data.c[1]};
    return Cplx{s.data.z[0], s.data.z[1]};
}

// alpha = alpha_A * alpha_B, rounded to the tensor's precision
Cplx product(const tblis_scalar& a, const tblis_scalar& b, type_t t)
{
    if (t == TYPE_SCOMPLEX)
    {
        const float ar = a.data.c[0], ai = a.data.c[1], br = b.data.c[0], bi = b.data.c[1];
        return Cplx{(double)(ar * br - ai * bi), (double)(ar * bi + ai * br)};
    }
    const double ar = a.data.z[0], ai = a.data.z[1], br = b.data.z[0], bi = b.data.z[1];
    return Cplx{ar * br - ai * bi, ar * bi + ai * br};
}

Side twice(Side s)
{
    for (auto& x : s.stride) x *= 2;
    for (auto& x : s.offs) x *= 2;
    return s;
}

PBlock part(const PBlock& b, int im)
{
    PBlock r;
    r.base = b.base == PANEL_MISSING ? PANEL_MISSING : 2 * b.base + im;
    r.rows = twice(b.rows);
    r.k = twice(b.k);
    return r;
}

// One real GEMM's k blocks: the parts of A and B they multiply (0 real,
// 1 imaginary) and the real scale of each original k block.
struct Term
{
    int a_im, b_im;
    double scale;
};

// One complex GEMM (k block q scaled by s_q = kscale[q] + i kscale_im[q])
// as the two real GEMMs of Re C and Im C.  With P_q = op(A_q) op(B_q),
//   Re(s P) = sr (Ar Br - sa sb Ai Bi) - si (sb Ar Bi + sa Ai Br)
//   Im(s P) = sr (sb Ar Bi + sa Ai Br) + si (Ar Br - sa sb Ai Bi);
// terms whose scale is zero are left out.
void split_gemm(const PanelGemm& g, double sa, double sb, std::vector<PanelGemm>& out)
{
    const size_t NM = g.mlen.size(), NN = g.nlen.size(), NK = g.klen.size();
    for (int ci = 0; ci < 2; ci++)
    {
        PanelGemm r;
        r.mlen = g.mlen;
        r.nlen = g.nlen;
        // (k block, term) pairs in term-major order, as the real GEMM's k blocks
        std::vector<std::pair<size_t, Term>> ks;
        for (int pass = 0; pass < 4; pass++)
            for (size_t q = 0; q < NK; q++)
            {
                const double sr = g.kscale.empty() ? 1.0 : g.kscale[q];
                const double si = g.kscale_im.empty() ? 0.0 : g.kscale_im[q];
                Term t{};
                if (ci == 0)
                {
                    const Term all[4] = {{0, 0, sr}, {1, 1, -sa * sb * sr}, {0, 1, -sb * si}, {1, 0, -sa * si}};
                    t = all[pass];
                }
                else
                {
                    const Term all[4] = {{0, 1, sb * sr}, {1, 0, sa * sr}, {0, 0, si}, {1, 1, -sa * sb * si}};
                    t = all[pass];
                }
                if (t.scale != 0.0) ks.push_back({q, t});
            }
        const size_t NK2 = ks.size();
        for (const auto& kt : ks)
        {
            r.klen.push_back(g.klen[kt.first]);
            r.kscale.push_back(kt.second.scale);
        }
        r.a.resize(NM * NK2);
        r.b.resize(NN * NK2);
        for (size_t x = 0; x < NK2; x++)
        {
            const size_t q = ks[x].first;
            const Term& t = ks[x].second;
            for (size_t p = 0; p < NM; p++) r.a[p * NK2 + x] = part(g.a[p * NK + q], t.a_im);
            for (size_t s = 0; s < NN; s++) r.b[s * NK2 + x] = part(g.b[s * NK + q], t.b_im);
        }
        for (const CBlock& c : g.c)
        {
            CBlock d;
            d.base = 2 * c.base + ci;
            d.rows = twice(c.rows);
            d.cols = twice(c.cols);
            r.c.push_back(d);
        }
        out.push_back(std::move(r));
    }
}

// Complex scale of the elements of C the contraction writes, as a level-1
// descriptor in complex elements: C's diagonal (a repeated label addresses
// one element per value, src/iface/3t/mult.cxx:44-48), length-1 dimensions
// dropped, smallest stride first, contiguous dimensions folded.
}  // namespace

Level1Desc c_desc(const tblis_tensor& C, const label_type* idx_C)
{
    Group g;
    std::string why;
    if (!canonicalize_one(C, idx_C, g, why)) fail("complex: invalid C: " + why);
    std::stable_sort(g.begin(), g.end(),
                     [](const Dim& a, const Dim& b) { return std::llabs(a.s[1]) < std::llabs(b.s[1]); });
    Level1Desc d;
    for (const Dim& x : g)
    {
        d.nb *= x.len;
        if (x.len == 1) continue;
        if (d.ndb > 0 && x.s[1] == d.sb[d.ndb - 1] * d.len_b[d.ndb - 1])
        {
            d.len_b[d.ndb - 1] *= x.len;
            continue;
        }
        if (d.ndb >= LEVEL1_MAXDIM) fail("complex: more than 16 dimensions of C after folding");
        d.len_b[d.ndb] = x.len;
        d.sb[d.ndb] = x.s[1];
        d.sa_b[d.ndb] = 0;
        d.ndb++;
    }
    return d;
}

void complex_split_gemms(const std::vector<PanelGemm>& in, double sa, double sb, std::vector<PanelGemm>& out)
{
    for (const PanelGemm& p : in) split_gemm(p, sa, sb, out);
}

void run_complex_plan(const std::shared_ptr<Plan>& plan, const void* A, const void* B, void* C, double alpha_re,
                      double alpha_im, double beta_re, double beta_im, const std::vector<CPass>& cps, int dt)
{
    const Cplx alpha{alpha_re, alpha_im}, beta{beta_re, beta_im};
    const bool use_k = alpha.re != 0.0 || alpha.im != 0.0;
    const bool one = beta.re == 1.0 && beta.im == 0.0;
    const bool bzero = beta.re == 0.0 && beta.im == 0.0;
    const bool calpha = alpha.im != 0.0;
    // C <- pre C before the GEMMs (pre: beta, or beta / alpha for a complex
    // alpha), C <- alpha C after them for a complex alpha
    Cplx pre = beta;
    if (calpha && !bzero)
    {
        const double d = alpha.re * alpha.re + alpha.im * alpha.im;
        pre = Cplx{(beta.re * alpha.re + beta.im * alpha.im) / d, (beta.im * alpha.re - beta.re * alpha.im) / d};
    }
    const bool pass = use_k ? (calpha ? !bzero : beta.im != 0.0) : !one;
    if (plan->ntasks == 0 || !(use_k || !one)) return;
    ThreadState& ts = thread_state();
    auto scale_c = [&](Cplx f, const char* what) {
        for (const CPass& cp : cps)
        {
            const hipError_t e = launch_level1_cscale(dt, cp.desc, f.re, f.im, cp.data, ts.stream);
            if (e != hipSuccess) fail(std::string(what) + ": " + hipGetErrorString(e));
        }
    };
    if (pass) scale_c(pre, "complex beta pass");
    auto finish = [&]() {
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(ts.stream, &cap) != hipSuccess) fail("hipStreamIsCapturing failed");
        if (!ts.async && cap == hipStreamCaptureStatusNone && hipStreamSynchronize(ts.stream) != hipSuccess)
            fail("hipStreamSynchronize failed");
    };
    if (use_k && calpha)
    {
        // GEMMs accumulate P (over C scaled by beta / alpha, or over
        // nothing when beta == 0), then C <- alpha C; the synchronising
        // variant of run_plan must not return before the last pass
        {
            struct Restore
            {
                ThreadState& t;
                bool async;
                ~Restore() { t.async = async; }
            } restore{ts, ts.async};
            ts.async = true;
            run_plan(plan, A, B, C, 1.0, bzero ? 0.0 : 1.0, true);
        }
        scale_c(alpha, "complex alpha pass");
        finish();
    }
    else if (use_k)
        run_plan(plan, A, B, C, alpha.re, pass ? 1.0 : beta.re, true);
    else
        finish();
}

void complex_tensor_mult(const tblis_config* cfg, const tblis_tensor* A, const label_type* idx_A,
                         const tblis_tensor* B, const label_type* idx_B, tblis_tensor* C,
                         const label_type* idx_C)
{
    const type_t type = A->type;
    const type_t rtype = type == TYPE_SCOMPLEX ? TYPE_SINGLE : TYPE_DOUBLE;
    Algo algo = resolve_algo(cfg, rtype);
    if (algo == Algo::Reference) algo = rtype == TYPE_DOUBLE ? Algo::Fp64Mfma : Algo::Bf16x3;
    thread_state().last_algo = algo_name(algo);

    const Cplx alpha = product(A->scalar, B->scalar, type);
    const Cplx beta = scalar_of(C->scalar, type);
    const double sa = A->conj ? -1.0 : 1.0, sb = B->conj ? -1.0 : 1.0;

    std::string key;
    put(key, (int32_t)algo);
    put(key, (int32_t)type);
    for (auto [t, idx] : {std::pair<const tblis_tensor*, const label_type*>{A, idx_A}, {B, idx_B}, {C, idx_C}})
    {
        put(key, (uint32_t)t->ndim);
        key.append((const char*)t->len, sizeof(len_type) * t->ndim);
        key.append((const char*)t->stride, sizeof(stride_type) * t->ndim);
        key.append(idx, t->ndim);
    }
    put(key, sa);
    put(key, sb);

    std::shared_ptr<Plan> plan = cache_get(key);
    if (!plan)
    {
        Canonical c;
        std::string why;
        if (!canonicalize(*A, idx_A, *B, idx_B, *C, idx_C, c, why)) fail("invalid contraction: " + why);
        GemmDesc g;
        if (dense_gemm(c, g, 0, 0, 0))
        {
            PlanBuilder pb(rtype == TYPE_DOUBLE ? 8 : 4, rtype == TYPE_DOUBLE ? 0 : 1, algo_family(algo));
            pb.set_complex(sa, sb);
            pb.add(g);
            plan = pb.build();
        }
        if (!plan) plan = std::make_shared<Plan>();
        cache_put(key, plan);
    }

    const bool use_k = alpha.re != 0.0 || alpha.im != 0.0;
    const bool one = beta.re == 1.0 && beta.im == 0.0;
    if (plan->ntasks > 0 && (use_k || !one))
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
        run_complex_plan(plan, A->data, B->data, C->data, alpha.re, alpha.im, beta.re, beta.im,
                         {CPass{C->data, c_desc(*C, idx_C)}}, rtype == TYPE_DOUBLE ? 0 : 1);
    }
    reset_scalar(C);
}

}  // namespace tblis_amd
