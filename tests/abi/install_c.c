/*
 * A C caller of the drop-in boundary, built against include/tblis.h only
 * (the reference's own install check is test/test_install.cxx:5-12): device
 * memory from hipMalloc, the permuted-layout contraction
 *     C[aibj] += 0.5 A[aebf] B[fjei],  a=b=e=f=64, i=j=32, column-major,
 * through tblis_tensor_mult, then 4096 entries of C against the oracle's
 * extended-precision dot products (oracle/oracle.h).  Exit status 0 when
 * the relative Frobenius error of the samples is <= 1e-12 and C's scalar
 * was reset to 1 (src/iface/3t/mult.cxx:162-163).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <hip/hip_runtime_api.h>

#include "tblis.h"
#include "../../oracle/oracle.h"

#define CHECK(x)                                                        \
    do {                                                                \
        hipError_t e_ = (x);                                            \
        if (e_ != hipSuccess) {                                         \
            fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));     \
            return 2;                                                   \
        }                                                               \
    } while (0)

static void col_major(int n, const len_type* len, stride_type* st)
{
    stride_type s = 1;
    for (int i = 0; i < n; i++) { st[i] = s; s *= len[i]; }
}

int main(void)
{
    len_type lA[4] = {64, 64, 64, 64}, lB[4] = {64, 32, 64, 32}, lC[4] = {64, 32, 64, 32};
    stride_type sA[4], sB[4], sC[4];
    col_major(4, lA, sA);
    col_major(4, lB, sB);
    col_major(4, lC, sC);
    const int64_t nA = 64LL * 64 * 64 * 64, nB = 64LL * 32 * 64 * 32, nC = nB;
    double *hA = malloc(8 * nA), *hB = malloc(8 * nB), *hC0 = malloc(8 * nC), *hC = malloc(8 * nC);
    orc_fill_uniform_d(hA, nA, 42);
    orc_fill_uniform_d(hB, nB, 142);
    orc_fill_uniform_d(hC0, nC, 242);
    double *dA, *dB, *dC;
    CHECK(hipMalloc((void**)&dA, 8 * nA));
    CHECK(hipMalloc((void**)&dB, 8 * nB));
    CHECK(hipMalloc((void**)&dC, 8 * nC));
    CHECK(hipMemcpy(dA, hA, 8 * nA, hipMemcpyHostToDevice));
    CHECK(hipMemcpy(dB, hB, 8 * nB, hipMemcpyHostToDevice));
    CHECK(hipMemcpy(dC, hC0, 8 * nC, hipMemcpyHostToDevice));

    tblis_tensor A, B, C;
    tblis_init_tensor_scaled_d(&A, 0.5, 4, lA, dA, sA);
    tblis_init_tensor_d(&B, 4, lB, dB, sB);
    tblis_init_tensor_scaled_d(&C, 1.0, 4, lC, dC, sC);
    tblis_tensor_mult(NULL, NULL, &A, "aebf", &B, "fjei", &C, "aibj");
    CHECK(hipMemcpy(hC, dC, 8 * nC, hipMemcpyDeviceToHost));

    int64_t olA[4], olB[4], olC[4], osA[4], osB[4], osC[4];
    for (int i = 0; i < 4; i++)
    {
        olA[i] = lA[i]; olB[i] = lB[i]; olC[i] = lC[i];
        osA[i] = sA[i]; osB[i] = sB[i]; osC[i] = sC[i];
    }
    orc_groups g;
    if (!orc_canonicalize(4, olA, osA, "aebf", 4, olB, osB, "fjei", 4, olC, osC, "aibj", &g)) return 3;
    enum { NS = 4096 };
    int64_t el[NS], em[NS], en[NS], off[NS];
    double want[NS];
    int64_t M = 1, N = 1;
    for (int i = 0; i < g.n_ac; i++) M *= g.len_ac[i];
    for (int i = 0; i < g.n_bc; i++) N *= g.len_bc[i];
    uint64_t x = 88172645463325252ULL;
    for (int s = 0; s < NS; s++)
    {
        x ^= x << 13; x ^= x >> 7; x ^= x << 17;
        el[s] = 0;
        em[s] = (int64_t)(x % (uint64_t)M);
        en[s] = (int64_t)((x >> 32) % (uint64_t)N);
        int64_t o = 0, f = em[s];
        for (int i = 0; i < g.n_ac; i++) { o += (f % g.len_ac[i]) * g.sC_ac[i]; f /= g.len_ac[i]; }
        f = en[s];
        for (int i = 0; i < g.n_bc; i++) { o += (f % g.len_bc[i]) * g.sC_bc[i]; f /= g.len_bc[i]; }
        off[s] = o;
    }
    orc_mult_entries_d(&g, 0.5, hA, hB, 1.0, hC0, 0, NS, el, em, en, want);
    double num = 0, den = 0;
    for (int s = 0; s < NS; s++)
    {
        double d = hC[off[s]] - want[s];
        num += d * d;
        den += want[s] * want[s];
    }
    double err = sqrt(num / den);
    printf("C ABI: C[aibj] += 0.5 A[aebf] B[fjei] (64/32): rel. Frobenius error %.3e over %d samples, "
           "C scalar after the call %g\n", err, NS, C.scalar.data.d);
    hipFree(dA);
    hipFree(dB);
    hipFree(dC);
    return (err <= 1e-12 && C.scalar.data.d == 1.0) ? 0 : 1;
}
