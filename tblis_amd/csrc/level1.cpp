// Host side of the level-1 operations: canonical form, special cases of the
// reference's front ends, kernel descriptor, launch.
#include "level1.hpp"

#include "contract.hpp"
#include "plan.hpp"

namespace tblis_amd
{

namespace
{

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) fail(std::string(what) + ": " + hipGetErrorString(e));
}

// Iteration order: smallest stride of the written tensor first.
void fill_dims(Group g, int slot_b, int slot_a, int64_t* len, int64_t* sb, int64_t* sa, int& n,
               int64_t& count)
{
    std::stable_sort(g.begin(), g.end(), [&](const Dim& x, const Dim& y) {
        const int64_t a = x.s[slot_b] < 0 ? -x.s[slot_b] : x.s[slot_b];
        const int64_t b = y.s[slot_b] < 0 ? -y.s[slot_b] : y.s[slot_b];
        return a < b;
    });
    for (const Dim& d : g)
    {
        if (n >= LEVEL1_MAXDIM) fail("level-1 operation: more than 16 dimensions after folding");
        len[n] = d.len;
        sb[n] = d.s[slot_b];
        if (sa) sa[n] = slot_a >= 0 ? d.s[slot_a] : 0;
        count *= d.len;
        n++;
    }
}

void run(int mode, type_t type, const Level1Desc& d, const void* A, void* B, double alpha,
         double beta, double value)
{
    ThreadState& ts = thread_state();
    const size_t wsb = level1_workspace_bytes(mode, d);
    void* ws = wsb ? workspace(wsb, ts.stream) : nullptr;
    hip_check(launch_level1(mode, type == TYPE_DOUBLE ? 0 : 1, d, A, B, alpha, beta, value, ws, ts.stream),
              "level-1 kernel launch");
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    hip_check(hipStreamIsCapturing(ts.stream, &cap), "hipStreamIsCapturing");
    if (!ts.async && cap == hipStreamCaptureStatusNone)
        hip_check(hipStreamSynchronize(ts.stream), "hipStreamSynchronize");
}

}  // namespace

// src/iface/1t/add.cxx:24-110, src/internal/1t/dense/add.cxx
void tensor_add(const tblis_config* cfg, const tblis_tensor* A, const label_type* idx_A,
                tblis_tensor* B, const label_type* idx_B)
{
    (void)cfg;
    check_tensor(A, idx_A, "A");
    check_tensor(B, idx_B, "B");
    if (A->type != B->type) fail("A and B must have the same type");
    check_real(A->type);
    thread_state().last_tile = -1;
    thread_state().last_direct = 0;
    Level1Form f;
    std::string why;
    if (!canonicalize_add(*A, idx_A, *B, idx_B, f, why)) fail("invalid add: " + why);
    const double alpha = scalar_value(A->scalar, A->type);
    const double beta = scalar_value(B->scalar, B->type);

    Level1Desc d;
    // B's dimensions: B-only (A stride 0) and AB
    Group bdims = f.b_only;
    for (auto& x : f.ab) bdims.push_back(x);
    fill_dims(bdims, 1, 0, d.len_b, d.sb, d.sa_b, d.ndb, d.nb);
    fill_dims(f.a_only, 0, -1, d.len_a, d.sa_a, nullptr, d.nda, d.na);
    if (d.nb == 0) return reset_scalar(B);  // empty B: nothing to write

    if (alpha == 0.0 || d.na == 0)
    {
        // A is not read (alpha == 0, src/iface/1t/add.cxx:89-104) or its
        // sums are empty (src/internal/1t/dense/add.cxx:45-49): B <- beta B
        if (beta != 1.0)
        {
            check_data(B, "B");
            run(beta == 0.0 ? LEVEL1_SET : LEVEL1_SCALE, B->type, d, nullptr, B->data, 0.0, beta, 0.0);
        }
        return reset_scalar(B);
    }
    check_data(A, "A");
    check_data(B, "B");
    run(LEVEL1_ADD, B->type, d, A->data, B->data, alpha, beta, 0.0);
    reset_scalar(B);
}

// src/iface/1t/scale.cxx:20-55
void tensor_scale(const tblis_config* cfg, tblis_tensor* A, const label_type* idx_A)
{
    (void)cfg;
    check_tensor(A, idx_A, "A");
    check_real(A->type);
    Group g;
    std::string why;
    if (!canonicalize_one(*A, idx_A, g, why)) fail("invalid scale: " + why);
    const double alpha = scalar_value(A->scalar, A->type);
    Level1Desc d;
    fill_dims(g, 1, -1, d.len_b, d.sb, d.sa_b, d.ndb, d.nb);
    if (d.nb > 0 && alpha != 1.0)
    {
        check_data(A, "A");
        run(alpha == 0.0 ? LEVEL1_SET : LEVEL1_SCALE, A->type, d, nullptr, A->data, 0.0, alpha, 0.0);
    }
    reset_scalar(A);
}

// src/iface/1t/set.cxx:20-51
void tensor_set(const tblis_config* cfg, const tblis_scalar* alpha, tblis_tensor* A,
                const label_type* idx_A)
{
    (void)cfg;
    check_tensor(A, idx_A, "A");
    if (!alpha) fail("set: null scalar");
    if (alpha->type != A->type) fail("set: the scalar's type differs from the tensor's");
    check_real(A->type);
    Group g;
    std::string why;
    if (!canonicalize_one(*A, idx_A, g, why)) fail("invalid set: " + why);
    Level1Desc d;
    fill_dims(g, 1, -1, d.len_b, d.sb, d.sa_b, d.ndb, d.nb);
    if (d.nb > 0)
    {
        check_data(A, "A");
        run(LEVEL1_SET, A->type, d, nullptr, A->data, 0.0, 0.0, scalar_value(*alpha, A->type));
    }
    reset_scalar(A);
}

}  // namespace tblis_amd
