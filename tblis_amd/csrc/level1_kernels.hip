// Level-1 tensor operations around a contraction (gfx950):
//     add    B <- alpha * sum_{A-only} A + beta * B     (src/iface/1t/add.cxx)
//     scale  B <- beta * B                              (src/iface/1t/scale.cxx)
//     set    B <- value                                 (src/iface/1t/set.cxx)
// One HBM-bound kernel: a grid-stride loop over B's elements in the order of
// B's strides (smallest first, so that consecutive lanes store to
// consecutive addresses), each element's A offset from the same multi-index
// (A stride 0 for dimensions only B has: replication), and a sum over the
// dimensions only A has (the reference's reduction, src/internal/1t/dense/
// add.cxx:95-122, accumulated in FP64).  beta == 0 writes B without reading
// it.
#include <hip/hip_runtime.h>

#include "level1.hpp"

namespace tblis_amd
{

namespace
{

template <typename T, int MODE>
__global__ __launch_bounds__(256) void level1_kernel(Level1Desc d, const T* __restrict__ A,
                                                     T* __restrict__ B, double alpha, double beta,
                                                     double value)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < d.nb; i += stride)
    {
        int64_t offb = 0, offa = 0, r = i;
        for (int k = 0; k < d.ndb; k++)
        {
            const int64_t q = r % d.len_b[k];
            r /= d.len_b[k];
            offb += q * d.sb[k];
            offa += q * d.sa_b[k];
        }
        T* p = B + offb;
        if (MODE == LEVEL1_SET)
        {
            *p = (T)value;
            continue;
        }
        if (MODE == LEVEL1_SCALE)
        {
            *p = (T)(beta * (double)*p);
            continue;
        }
        // LEVEL1_ADD: sum over the A-only dimensions in odometer order
        double sum = 0.0;
        for (int64_t j = 0; j < d.na; j++)
        {
            int64_t off = offa, rr = j;
            for (int k = 0; k < d.nda; k++)
            {
                const int64_t q = rr % d.len_a[k];
                rr /= d.len_a[k];
                off += q * d.sa_a[k];
            }
            sum += (double)A[off];
        }
        double v = alpha * sum;
        if (beta != 0.0) v += beta * (double)*p;
        *p = (T)v;
    }
}

template <typename T, int MODE>
hipError_t launch_t(const Level1Desc& d, const void* A, void* B, double alpha, double beta,
                    double value, hipStream_t s)
{
    if (d.nb == 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((d.nb + 255) / 256, 8192);
    hipLaunchKernelGGL((level1_kernel<T, MODE>), dim3((unsigned)blocks), dim3(256), 0, s, d,
                       (const T*)A, (T*)B, alpha, beta, value);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_level1(int mode, int dtype, const Level1Desc& d, const void* A, void* B,
                         double alpha, double beta, double value, hipStream_t s)
{
    if (dtype == 0)
    {
        if (mode == LEVEL1_ADD) return launch_t<double, LEVEL1_ADD>(d, A, B, alpha, beta, value, s);
        if (mode == LEVEL1_SCALE) return launch_t<double, LEVEL1_SCALE>(d, A, B, alpha, beta, value, s);
        return launch_t<double, LEVEL1_SET>(d, A, B, alpha, beta, value, s);
    }
    if (mode == LEVEL1_ADD) return launch_t<float, LEVEL1_ADD>(d, A, B, alpha, beta, value, s);
    if (mode == LEVEL1_SCALE) return launch_t<float, LEVEL1_SCALE>(d, A, B, alpha, beta, value, s);
    return launch_t<float, LEVEL1_SET>(d, A, B, alpha, beta, value, s);
}

}  // namespace tblis_amd
