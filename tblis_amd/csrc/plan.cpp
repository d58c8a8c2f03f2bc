// Planner, plan cache and launch.
#include <atomic>
#include "plan.hpp"

#include "contract.hpp"
#include "panel.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <list>
#include <map>
#include <mutex>
#include <unordered_map>

namespace tblis_amd
{

namespace
{

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) fail(std::string(what) + ": " + hipGetErrorString(e));
}

std::atomic<int64_t> g_device_allocs{0};

}  // namespace

hipError_t device_alloc(void** p, size_t bytes)
{
    g_device_allocs.fetch_add(1);
    return hipMalloc(p, bytes);
}

int64_t device_alloc_count() { return g_device_allocs.load(); }

namespace
{
struct StaticRegion
{
    char* base = nullptr;
    size_t bytes = 0;
};
thread_local StaticRegion t_static;
}  // namespace

StaticStorageScope::StaticStorageScope(void* base, size_t bytes) { t_static = StaticRegion{(char*)base, bytes}; }
StaticStorageScope::~StaticStorageScope() { t_static = StaticRegion{}; }

void static_storage(char*& base, size_t& bytes)
{
    base = t_static.base;
    bytes = t_static.bytes;
}

Plan::~Plan()
{
    if (dev && owns_dev)
    {
        int cur = 0;
        (void)hipGetDevice(&cur);
        (void)hipSetDevice(device);
        (void)hipFree(dev);
        (void)hipSetDevice(cur);
    }
}

PlanBuilder::PlanBuilder(int elem_size, int dtype, int family)
    : elem_size_(elem_size),
      dtype_(dtype),
      family_(family),
      panel_gemms_(new std::vector<PanelGemm>())
{
}

PlanBuilder::~PlanBuilder() = default;

void PlanBuilder::add_panel(const PanelGemm& g) { panel_gemms_->push_back(g); }


std::shared_ptr<Plan> PlanBuilder::build()
{
    if (gemms_.empty() && panel_gemms_->empty()) return nullptr;
    auto plan = std::make_shared<Plan>();
    std::vector<PanelGemm> pg(*panel_gemms_);
    for (auto& g : gemms_) panel_from_desc(g, pg);
    if (complex_)
    {
        std::vector<PanelGemm> cg;
        complex_split_gemms(pg, sa_, sb_, cg);
        pg.swap(cg);
    }
    if (!build_panel_plan(*plan, std::move(pg), elem_size_, dtype_, family_)) return nullptr;
    return plan;
}

// ------------------------------------------------------------------ state

ThreadState& thread_state()
{
    static thread_local ThreadState ts;
    return ts;
}

namespace
{

struct Cache
{
    std::mutex mu;
    std::list<std::pair<std::string, std::shared_ptr<Plan>>> lru;
    std::unordered_map<std::string, decltype(lru)::iterator> index;
    size_t max_entries = 256;
};

// Never destroyed: plans hold device memory that must not be freed after
// the HIP runtime has shut down at process exit.
Cache& cache()
{
    static Cache* c = new Cache();
    return *c;
}

}  // namespace

std::shared_ptr<Plan> cache_get(const std::string& key)
{
    Cache& c = cache();
    std::lock_guard<std::mutex> g(c.mu);
    auto it = c.index.find(key);
    if (it == c.index.end()) return nullptr;
    c.lru.splice(c.lru.begin(), c.lru, it->second);
    return it->second->second;
}

void cache_put(const std::string& key, std::shared_ptr<Plan> p)
{
    Cache& c = cache();
    std::lock_guard<std::mutex> g(c.mu);
    auto it = c.index.find(key);
    if (it != c.index.end())
    {
        c.lru.erase(it->second);
        c.index.erase(it);
    }
    c.lru.emplace_front(key, std::move(p));
    c.index[key] = c.lru.begin();
    while (c.lru.size() > c.max_entries)
    {
        c.index.erase(c.lru.back().first);
        c.lru.pop_back();
    }
}

void cache_clear()
{
    Cache& c = cache();
    std::lock_guard<std::mutex> g(c.mu);
    c.index.clear();
    c.lru.clear();
}

int device_cu_count()
{
    static int cus[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (!cus[dev])
    {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
            n = 256;
        cus[dev] = n;
    }
    return cus[dev];
}

void check_device_pointer(const void* p, const char* what, int64_t lo, int64_t hi, int elem)
{
    if (!p) fail(std::string(what) + ": null data pointer");
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess)
    {
        (void)hipGetLastError();
        fail(std::string(what) + ": data is not HIP-visible memory");
    }
    int cur = 0;
    (void)hipGetDevice(&cur);
    switch (attr.type)
    {
        case hipMemoryTypeDevice:
            if (attr.device != cur)
                fail(std::string(what) + ": data lives on HIP device " + std::to_string(attr.device) +
                     ", not on the current device " + std::to_string(cur));
            break;
        case hipMemoryTypeManaged:
        case hipMemoryTypeUnified:
            break;
        case hipMemoryTypeHost:
            if (attr.devicePointer) break;
            fail(std::string(what) + ": host memory that is not mapped for the device");
        default:
            fail(std::string(what) +
                 ": data is not device-accessible memory (host pointers are not supported; "
                 "allocate with hipMalloc, hipHostMalloc or a torch CUDA tensor)");
    }
    // Every element the plan touches must lie inside the allocation.
    if (attr.type == hipMemoryTypeDevice && lo <= hi)
    {
        hipDeviceptr_t base = nullptr;
        size_t size = 0;
        if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p) == hipSuccess && base)
        {
            const char* b = (const char*)base;
            const char* q = (const char*)p;
            if (q + lo * elem < b || q + (hi + 1) * elem > b + size)
                fail(std::string(what) + ": the tensor's lengths and strides reach outside its "
                     "allocation");
        }
        else
        {
            (void)hipGetLastError();
        }
    }
}

// Grow-only device workspace per (device, stream), for packed operands.
// Calls on one stream are ordered, so they can share it.  A buffer that is
// outgrown is retired, not freed: a captured graph may still refer to it.
void* workspace(size_t bytes, hipStream_t s)
{
    struct Entry
    {
        void* ptr = nullptr;
        size_t size = 0;
    };
    static std::mutex mu;
    static auto* ws = new std::map<std::pair<int, hipStream_t>, Entry>();
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> g(mu);
    Entry& e = (*ws)[{dev, s}];
    if (e.size < bytes)
    {
        size_t grow = e.size + e.size / 2;
        size_t want = bytes > grow ? bytes : grow;
        void* p = nullptr;
        hip_check(device_alloc(&p, want), "hipMalloc (workspace)");
        e.ptr = p;
        e.size = want;
    }
    return e.ptr;
}

namespace
{

// Plans used while a stream was being captured (never released).
void pin_plan(const std::shared_ptr<Plan>& p)
{
    static std::mutex mu;
    static auto* pinned = new std::vector<std::shared_ptr<Plan>>();
    std::lock_guard<std::mutex> g(mu);
    for (auto& q : *pinned)
        if (q == p) return;
    pinned->push_back(p);
}

}  // namespace

namespace
{

// The side stream and at least n events for staged packs, or false when
// they do not exist yet and the stream is being captured (no creation
// inside a capture: the call then runs its stages unstaged).
bool stage_resources(ThreadState& ts, hipStream_t main, bool capturing, size_t n, hipStream_t& side)
{
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    auto it = ts.side.find({dev, main});
    if (it == ts.side.end() || ts.stage_ev.size() < n)
    {
        if (capturing) return false;
        if (it == ts.side.end())
        {
            hipStream_t s = nullptr;
            hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags");
            it = ts.side.emplace(std::make_pair(dev, main), s).first;
        }
        while (ts.stage_ev.size() < n)
        {
            hipEvent_t e = nullptr;
            hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreateWithFlags");
            ts.stage_ev.push_back(e);
        }
    }
    side = it->second;
    return true;
}

}  // namespace

void launch_panel_calls(const PanelLaunch& P, const PanelSchedule& S, bool overlap)
{
    ThreadState& ts = thread_state();
    hipStream_t side = nullptr;
    bool staged = false;
    if (overlap && S.nstages > 1 && P.use_k && P.nchunks > 0)
    {
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        hip_check(hipStreamIsCapturing(P.stream, &cap), "hipStreamIsCapturing");
        staged = stage_resources(ts, P.stream, cap != hipStreamCaptureStatusNone, (size_t)S.nstages + 1, side);
    }
    if (staged)
    {
        // fork: the side stream starts after the work queued on P.stream;
        // stage s's pack signals stage_ev[s + 1]
        hip_check(hipEventRecord(ts.stage_ev[0], P.stream), "hipEventRecord");
        hip_check(hipStreamWaitEvent(side, ts.stage_ev[0], 0), "hipStreamWaitEvent");
        int64_t c0 = 0;
        for (int s = 0; s < S.nstages; s++)
        {
            PanelLaunch K = P;
            K.chunks = P.chunks + c0;
            K.nchunks = (int)S.chunks[s];
            K.stream = side;
            hip_check(launch_panel_pack(K), "pack launch");
            hip_check(hipEventRecord(ts.stage_ev[s + 1], side), "hipEventRecord");
            c0 += S.chunks[s];
        }
    }
    else if (P.use_k)
        hip_check(launch_panel_pack(P), "pack launch");
    // per stage, one GEMM launch per group of tasks with the same direct
    // operands; the last stage's wait joins the side stream back
    int64_t at = 0;
    for (int s = 0; s < S.nstages; s++)
    {
        if (staged) hip_check(hipStreamWaitEvent(P.stream, ts.stage_ev[s + 1], 0), "hipStreamWaitEvent");
        for (int g = 0; g < 4; g++)
        {
            if (S.group_n[s][g] == 0) continue;
            PanelLaunch G = P;
            G.tasks = P.tasks + at;
            G.ntasks = (int)S.group_n[s][g];
            G.direct_a = (g & 1) != 0;
            G.direct_b = (g & 2) != 0;
            hip_check(launch_panel_gemm(G), "kernel launch");
            at += S.group_n[s][g];
        }
    }
    hip_check(launch_panel_unpack(P), "unpack launch");
}

void run_plan(const std::shared_ptr<Plan>& pp, const void* A, const void* B, void* C, double alpha,
              double beta, bool use_k)
{
    const Plan& p = *pp;
    ThreadState& ts = thread_state();
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    hip_check(hipStreamIsCapturing(ts.stream, &cap), "hipStreamIsCapturing");
    const bool capturing = cap != hipStreamCaptureStatusNone;
    if (capturing) pin_plan(pp);
    ts.last_tile = p.panel_tile;
    // p.direct is in the GEMM's operand order, swapped when it computes C^T
    ts.last_direct = (p.direct[p.swap ? 1 : 0] ? 1 : 0) | (p.direct[p.swap ? 0 : 1] ? 2 : 0);
    PanelLaunch P{};
    P.chunks = p.d_pchunks;
    P.nchunks = p.p_nchunks;
    P.tasks = p.d_ptasks;
    P.ntasks = p.ntasks;
    P.ujobs = p.d_ujobs;
    P.uchunks = p.d_uchunks;
    P.nuchunks = p.p_nuchunks;
    P.pool = p.d_pool;
    P.A = p.swap ? B : A;
    P.B = p.swap ? A : B;
    P.C = C;
    char* ws = p.ws_static ? p.ws_static : (char*)workspace(p.ws_bytes, ts.stream);
    P.ws = ws;
    P.cws = (double*)(ws + p.ws_c_offset);
    P.family = p.key.family;
    P.direct_a = p.direct[0];
    P.direct_b = p.direct[1];
    P.alpha = alpha;
    P.beta = beta;
    P.use_k = use_k && p.reads_ab ? 1 : 0;
    P.dtype = p.key.dtype;
    P.tile = p.panel_tile;
    P.stream = ts.stream;
    if (P.family == FAMILY_OZAKI)
    {
        // row maxima, digit planes, int8 GEMM; then C panels if any
        hip_check(launch_panel_ozaki(P, ozaki_slices(), (unsigned long long*)(ws + p.ws_rmax_offset),
                                     P.use_k ? p.oz_rows : 0),
                  "kernel launch");
        hip_check(launch_panel_unpack(P), "unpack launch");
    }
    else
        launch_panel_calls(P, p.sched, p.ws_static == nullptr);
    if (!ts.async && !capturing) hip_check(hipStreamSynchronize(ts.stream), "hipStreamSynchronize");
}

}  // namespace tblis_amd
