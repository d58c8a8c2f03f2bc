// Device scratch of the no-allocation contraction entry point
// (tblis_hip_tensor_mult_noalloc): a file-scope device array that holds the
// plan tables and the pack / C panel workspace of that entry point's calls,
// so that they call no hipMalloc.  Each GPU holds its own copy (64 MiB)
// from the library's first use on it until the process ends; every call on
// one GPU reuses it.
#include <mutex>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "contract.hpp"
#include "plan.hpp"

namespace tblis_amd
{

namespace
{

constexpr size_t kScratchBytes = (size_t)64 << 20;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

}  // namespace

// The current device's address of scratch, looked up once per device with
// that device current (the host-side name is not a device address).
char* noalloc_scratch(size_t& bytes)
{
    static std::mutex mu;
    static std::vector<void*> ptr;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) fail(std::string("hipGetDevice: ") + hipGetErrorString(e));
    std::lock_guard<std::mutex> g(mu);
    if ((size_t)dev >= ptr.size()) ptr.resize((size_t)dev + 1, nullptr);
    if (!ptr[dev])
    {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
        {
            (void)hipGetLastError();
            fail(std::string("hipGetSymbolAddress (no-allocation scratch): ") + hipGetErrorString(e));
        }
        ptr[dev] = p;
    }
    bytes = kScratchBytes;
    return (char*)ptr[dev];
}

}  // namespace tblis_amd
