// A C++ caller of the drop-in boundary through include/tblis.hpp's
// tblis::mult<T> (the reference's templates, src/iface/3t/mult.h:28-38, and
// its install check test/test_install.cxx:5-12: C[ead] = A[abc] B[cbde] with
// A 10x4x3, B 3x4x6x7, C 7x10x6, row-major as MArray's default layout), and
// the permuted-layout contraction C[aibj] += 0.5 A[aebf] B[fjei] at
// a=b=e=f=24, i=j=12, column-major.  Device memory from hipMalloc; every
// element of C is compared with the oracle's contraction (oracle/oracle.h).
// Exit status 0 when both are within 1e-12 relative Frobenius error.
#include <cmath>
#include <cstdio>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "tblis.hpp"
#include "../../oracle/oracle.h"

namespace
{

std::vector<stride_type> row_major(const std::vector<len_type>& l)
{
    std::vector<stride_type> s(l.size());
    stride_type a = 1;
    for (size_t i = l.size(); i-- > 0;) { s[i] = a; a *= l[i]; }
    return s;
}

std::vector<stride_type> col_major(const std::vector<len_type>& l)
{
    std::vector<stride_type> s(l.size());
    stride_type a = 1;
    for (size_t i = 0; i < l.size(); i++) { s[i] = a; a *= l[i]; }
    return s;
}

int64_t count(const std::vector<len_type>& l)
{
    int64_t n = 1;
    for (auto x : l) n *= x;
    return n;
}

double* to_device(const std::vector<double>& h)
{
    void* p = nullptr;
    if (hipMalloc(&p, h.size() * 8) != hipSuccess) return nullptr;
    if (hipMemcpy(p, h.data(), h.size() * 8, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    return (double*)p;
}

// C <- alpha A B + beta C through tblis::mult<double>, checked against the
// oracle on host copies
double run(double alpha, const std::vector<len_type>& lA, const std::vector<stride_type>& sA, const char* iA,
           const std::vector<len_type>& lB, const std::vector<stride_type>& sB, const char* iB, double beta,
           const std::vector<len_type>& lC, const std::vector<stride_type>& sC, const char* iC, uint64_t seed)
{
    std::vector<double> A(count(lA)), B(count(lB)), C0(count(lC)), got(count(lC));
    orc_fill_uniform_d(A.data(), (int64_t)A.size(), seed);
    orc_fill_uniform_d(B.data(), (int64_t)B.size(), seed + 100);
    orc_fill_uniform_d(C0.data(), (int64_t)C0.size(), seed + 200);
    double *dA = to_device(A), *dB = to_device(B), *dC = to_device(C0);
    if (!dA || !dB || !dC) return INFINITY;
    tblis::varray_view<const double> a{dA, lA, sA};
    tblis::varray_view<const double> b{dB, lB, sB};
    tblis::varray_view<double> c{dC, lC, sC};
    tblis::mult<double>(alpha, a, iA, b, iB, beta, c, iC);
    if (hipMemcpy(got.data(), dC, got.size() * 8, hipMemcpyDeviceToHost) != hipSuccess) return INFINITY;
    hipFree(dA);
    hipFree(dB);
    hipFree(dC);
    std::vector<double> want = C0;
    std::vector<int64_t> la(lA.begin(), lA.end()), lb(lB.begin(), lB.end()), lc(lC.begin(), lC.end());
    std::vector<int64_t> xa(sA.begin(), sA.end()), xb(sB.begin(), sB.end()), xc(sC.begin(), sC.end());
    orc_tensor_mult_d(1, alpha, (int)la.size(), la.data(), xa.data(), iA, A.data(), (int)lb.size(), lb.data(),
                      xb.data(), iB, B.data(), beta, (int)lc.size(), lc.data(), xc.data(), iC, want.data());
    double num = 0, den = 0;
    for (size_t i = 0; i < want.size(); i++)
    {
        num += (got[i] - want[i]) * (got[i] - want[i]);
        den += want[i] * want[i];
    }
    return std::sqrt(num / den);
}

}  // namespace

int main()
{
    const double e1 = run(1.0, {10, 4, 3}, row_major({10, 4, 3}), "abc", {3, 4, 6, 7}, row_major({3, 4, 6, 7}),
                          "cbde", 0.0, {7, 10, 6}, row_major({7, 10, 6}), "ead", 7);
    std::printf("C++ tblis::mult: C[ead] = A[abc] B[cbde] (row major): rel. Frobenius error %.3e\n", e1);
    const std::vector<len_type> lA{24, 24, 24, 24}, lB{24, 12, 24, 12}, lC{24, 12, 24, 12};
    const double e2 = run(0.5, lA, col_major(lA), "aebf", lB, col_major(lB), "fjei", 1.0, lC, col_major(lC),
                          "aibj", 42);
    std::printf("C++ tblis::mult: C[aibj] += 0.5 A[aebf] B[fjei] (24/12): rel. Frobenius error %.3e\n", e2);
    return e1 <= 1e-12 && e2 <= 1e-12 ? 0 : 1;
}
