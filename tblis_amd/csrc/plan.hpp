// Planner: turns one or many (sub-)GEMMs over scatter-described matrices
// into the task/segment lists the kernels walk, and owns their device copy.
#pragma once

#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "kernels.hpp"
#include "plan_types.hpp"

namespace tblis_amd
{

// One k segment of a GEMM: A (m x k) and B (k x n) patches.
struct GemmSeg
{
    int64_t a_base = 0, b_base = 0;
    Side a_rows, a_cols, b_rows, b_cols;
    int64_t k = 0;
};

// C (m x n) patch, optionally repeated over a batch of base offsets, with
// its k segments (summed into the same C patch).
struct GemmDesc
{
    int64_t c_base = 0;
    Side c_rows, c_cols;
    int64_t m = 0, n = 0;
    std::vector<GemmSeg> segs;
    // batch: extra offsets added to the bases (empty = one instance)
    std::vector<int64_t> l_a, l_b, l_c;
};

struct Plan
{
    KernelKey key{};
    bool swap = false;  // kernel computes C^T = B^T A^T
    int ntasks = 0;
    int64_t useful_flops = 0;
    bool reads_ab = false;   // some task has k segments
    int elem = 8;            // element size of A, B, C
    // Panel plan (panel.hpp): pack jobs -> panel GEMM tasks -> C unpack, in
    // a workspace of ws_bytes
    size_t ws_bytes = 0;
    int panel_tile = 0;
    size_t ws_c_offset = 0;  // byte offset of the C panels in the workspace
    size_t ws_rmax_offset = 0;  // Ozaki: byte offset of the row-maximum array
    int64_t oz_rows = 0;        // Ozaki: its entries
    bool direct[2] = {false, false};  // FP64 operands read from the tensors (every task)
    int64_t group_n[4] = {0, 0, 0, 0};  // panel tasks reading none / A / B / both from the tensors
    PanelSchedule sched;                // stages of chunks and task groups
    int p_nchunks = 0, p_nuchunks = 0;
    PackChunk* d_pchunks = nullptr;
    PanelTask* d_ptasks = nullptr;
    UnpackJob* d_ujobs = nullptr;
    UnpackChunk* d_uchunks = nullptr;
    // Data pointers (A, B, C of the first instance) whose memory checks
    // passed on the last call; a call with the same pointers skips them.
    // The key of a batched plan fixes every other instance's pointer
    // relative to these.
    std::mutex verified_mu;
    const void* verified[3] = {nullptr, nullptr, nullptr};
    bool verified_k = false;
    bool already_verified(const void* a, const void* b, const void* c, bool k)
    {
        std::lock_guard<std::mutex> g(verified_mu);
        return verified[2] == c && (!k || (verified_k && verified[0] == a && verified[1] == b));
    }
    void mark_verified(const void* a, const void* b, const void* c, bool k)
    {
        std::lock_guard<std::mutex> g(verified_mu);
        verified[0] = a;
        verified[1] = b;
        verified[2] = c;
        verified_k = k;
    }
    int64_t mfma_flops = 0;  // incl. tile padding at 16-row MFMA granularity
    void* dev = nullptr;     // one allocation: chunks | tasks | unpack jobs | pool
    bool owns_dev = true;    // false: dev is caller storage (StaticStorageScope)
    char* ws_static = nullptr;  // workspace in caller storage (else the stream's pool)
    int64_t* d_pool = nullptr;
    size_t bytes = 0;
    int device = 0;
    ~Plan();
};

struct PanelGemm;

// While one lives on a thread, panel plans built there place their device
// tables and their workspace in [base, base + bytes) of device memory
// instead of allocating: the tables at base, the workspace after them.  The
// upload is ordered on the thread's stream and waited for.  A plan that
// does not fit fails with an error naming the limit.
struct StaticStorageScope
{
    StaticStorageScope(void* base, size_t bytes);
    ~StaticStorageScope();
};
// The active scope's region (base nullptr when none).
void static_storage(char*& base, size_t& bytes);

class PlanBuilder
{
  public:
    PlanBuilder(int elem_size, int dtype, int family = FAMILY_F64);
    ~PlanBuilder();
    void add(const GemmDesc& g) { gemms_.push_back(g); }
    // A GEMM assembled from blocks.
    void add_panel(const PanelGemm& g);
    // The GEMMs added are complex (offsets in complex elements, element
    // size that of the real part): build() runs each as two real GEMMs.
    void set_complex(double sa, double sb)
    {
        complex_ = true;
        sa_ = sa;
        sb_ = sb;
    }
    // Orientation, tile, pack jobs, tasks; upload.  Returns nullptr if
    // there is nothing to launch.
    std::shared_ptr<Plan> build();

  private:
    int elem_size_, dtype_, family_;
    bool complex_ = false;
    double sa_ = 1.0, sb_ = 1.0;
    std::vector<GemmDesc> gemms_;
    std::unique_ptr<std::vector<PanelGemm>> panel_gemms_;
};

// Per-thread execution settings.
struct ThreadState
{
    hipStream_t stream = nullptr;
    bool async = false;
    bool return_on_error = false;
    std::string last_error;
    std::string last_algo;
    int last_tile = -1;  // panel tile of the last launched plan (-1: other kernels)
    int last_direct = 0;  // its operands read from the tensors: 1 A, 2 B
    // side streams for staged packs, by (device, main stream), and the
    // events that order them (reused call after call)
    std::map<std::pair<int, hipStream_t>, hipStream_t> side;
    std::vector<hipEvent_t> stage_ev;
};
ThreadState& thread_state();

// Plan cache (process-wide, thread-safe).
std::shared_ptr<Plan> cache_get(const std::string& key);
void cache_put(const std::string& key, std::shared_ptr<Plan> p);
void cache_clear();

// Launch a plan on the current thread's stream and wait unless async.
// During stream capture (hipGraph) the call never waits, and the plan is
// pinned for the life of the process: a captured graph keeps referring to
// its device tables after the cache would have evicted them.
// Grow-only device workspace per (device, stream); calls on one stream share it.
void* workspace(size_t bytes, hipStream_t s);

void run_plan(const std::shared_ptr<Plan>& p, const void* A, const void* B, void* C, double alpha,
              double beta, bool use_k);

// The launches of a non-Ozaki panel plan: pack (when k is used), one GEMM
// launch per group of tasks with the same directly read operands
// (group_n[g] tasks from P.tasks on, g = direct A + 2 direct B), unpack.
// The pack, GEMM and unpack launches of a panel plan.  With overlap and
// several stages the packs go to a side stream of the calling thread (one
// per device and main stream, created on first use) and each stage's GEMMs
// wait for its pack; otherwise one pack launch precedes the GEMMs.
void launch_panel_calls(const PanelLaunch& P, const PanelSchedule& S, bool overlap);

// Every device allocation the library makes goes through this (counted, so
// that a test can show that a call made none).
hipError_t device_alloc(void** p, size_t bytes);
int64_t device_alloc_count();

// Device pointer check for a tensor that will be accessed: the memory must
// be device-accessible on the current device and, for device allocations,
// element offsets [lo, hi] (in elements of size elem) must lie inside it.
void check_device_pointer(const void* p, const char* what, int64_t lo, int64_t hi, int elem);

int device_cu_count();

}  // namespace tblis_amd
