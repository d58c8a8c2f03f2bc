// Planner, plan cache and launch.
#include <atomic>
#include "plan.hpp"

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

bool unit_first(const Side& s) { return !s.len.empty() && s.stride[0] == 1 && s.len[0] > 1; }

int64_t first_stride(const Side& s)
{
    if (s.len.empty()) return 0;
    return s.stride[0] < 0 ? -s.stride[0] : s.stride[0];
}

// Coalesce loads along the rows of a (rows x cols) operand tile?
bool rows_fast(const Side& rows, const Side& cols)
{
    if (unit_first(rows)) return true;
    if (unit_first(cols)) return false;
    if (rows.len.empty()) return false;
    if (cols.len.empty()) return true;
    return first_stride(rows) <= first_stride(cols);
}

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

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

PlanBuilder::PlanBuilder(int elem_size, int dtype, int use_mfma, int family)
    : elem_size_(elem_size),
      dtype_(dtype),
      use_mfma_(use_mfma),
      family_(family),
      panel_gemms_(new std::vector<PanelGemm>())
{
}

PlanBuilder::~PlanBuilder() = default;

void PlanBuilder::add_panel(const PanelGemm& g) { panel_gemms_->push_back(g); }

bool PlanBuilder::panel_route() const { return panel_enabled(family_, use_mfma_); }

SideRef PlanBuilder::side_ref(const Side& s, int64_t BS, int64_t start)
{
    if (s.len.size() <= 1) return SideRef{start, s.len.empty() ? 0 : s.stride[0]};
    std::vector<int64_t> key;
    key.push_back(BS);
    key.push_back(s.pack_3d);
    key.insert(key.end(), s.len.begin(), s.len.end());
    key.insert(key.end(), s.stride.begin(), s.stride.end());
    auto it = dedupe_.find(key);
    int64_t at;
    if (it != dedupe_.end())
    {
        at = it->second;
    }
    else
    {
        int64_t total = 1;
        for (auto l : s.len) total *= l;
        at = (int64_t)pool_.size();
        pool_.resize(pool_.size() + std::max<int64_t>(total, 1));
        fill_scatter(s, BS, 0, total, pool_.data() + at, elem_size_);
        dedupe_[key] = at;
    }
    return SideRef{at + start, SCATTERED};
}

int64_t PlanBuilder::kblock_table(const Side& s, int64_t BK)
{
    if (s.len.size() <= 1) return -1;
    std::vector<int64_t> key;
    key.push_back(-BK);  // distinct from scatter keys
    key.push_back(s.pack_3d);
    key.insert(key.end(), s.len.begin(), s.len.end());
    key.insert(key.end(), s.stride.begin(), s.stride.end());
    auto it = dedupe_.find(key);
    if (it != dedupe_.end()) return it->second;
    SideRef r = side_ref(s, BK, 0);
    int64_t total = 1;
    for (auto l : s.len) total *= l;
    int64_t nblk = (total + BK - 1) / BK;
    int64_t at = (int64_t)pool_.size();
    pool_.resize(pool_.size() + 2 * std::max<int64_t>(nblk, 1));
    const int64_t* scat = pool_.data() + r.start;
    for (int64_t b = 0; b < nblk; b++)
    {
        int64_t k0 = b * BK, n = std::min(BK, total - k0);
        int64_t stride = n > 1 ? scat[k0 + 1] - scat[k0] : 0;
        for (int64_t j = k0 + 1; j + 1 < k0 + n; j++)
            if (scat[j + 1] - scat[j] != stride) stride = SCATTERED;
        pool_[at + 2 * b] = scat[k0];
        pool_[at + 2 * b + 1] = stride;
    }
    dedupe_[key] = at;
    return at;
}

// A single dense GEMM in a split-BF16 family with enough tiles that every
// A' element would otherwise be re-split by many workgroups: pack both
// operands once, then run the panel GEMM.
bool PlanBuilder::try_packed(Plan& plan, std::vector<TaskDesc>& tasks, std::vector<SegDesc>& segs)
{
    if (family_ == FAMILY_F64 || !use_mfma_ || gemms_.size() != 1) return false;
    if (const char* e = std::getenv("TBLIS_HIP_NO_PACK"))
        if (e[0] == '1') return false;
    const GemmDesc& g = gemms_[0];
    if (g.segs.size() != 1 || !g.l_c.empty()) return false;
    const bool sw = plan.swap;
    const int64_t m = sw ? g.n : g.m, n = sw ? g.m : g.n, k = g.segs[0].k;
    const int64_t PR = pack_bf16_rows(), PK = pack_bf16_k();
    if (m < 2 * PR || n < 2 * PR || k < PK) return false;

    const GemmSeg& s = g.segs[0];
    const Side& ar = sw ? s.b_cols : s.a_rows;
    const Side& ac = sw ? s.b_rows : s.a_cols;
    const Side& br = sw ? s.a_cols : s.b_rows;
    const Side& bc = sw ? s.a_rows : s.b_cols;
    const int64_t nkt = ceil_div(k, PK);
    const int64_t chunk = 3 * PR * PK;
    const Side* rows[2] = {&ar, &bc};
    const Side* ks[2] = {&ac, &br};
    const int64_t nr[2] = {m, n};
    const int64_t base[2] = {sw ? s.b_base : s.a_base, sw ? s.a_base : s.b_base};
    int64_t out = 0;
    for (int o = 0; o < 2; o++)
    {
        PackDesc& d = plan.pack[o];
        d.base = base[o];
        d.rows = side_ref(*rows[o], PR, 0);
        d.kside = side_ref(*ks[o], PK, 0);
        d.ktbl = kblock_table(*ks[o], PK);
        d.nrows = nr[o];
        d.klen = k;
        d.nrt = (int32_t)ceil_div(nr[o], PR);
        d.nkt = (int32_t)nkt;
        d.out = out;
        plan.pack_kfast[o] = !rows_fast(*rows[o], *ks[o]);
        out += (int64_t)d.nrt * nkt * chunk;
    }
    plan.ws_bytes = (size_t)out * 2;

    const Side& cr = sw ? g.c_cols : g.c_rows;
    const Side& cc = sw ? g.c_rows : g.c_cols;
    for (int64_t tm = 0; tm < plan.pack[0].nrt; tm++)
        for (int64_t tn = 0; tn < plan.pack[1].nrt; tn++)
        {
            TaskDesc t{};
            t.c_base = g.c_base;
            t.c_row = side_ref(cr, PR, tm * PR);
            t.c_col = side_ref(cc, PR, tn * PR);
            t.mlen = (int32_t)std::min<int64_t>(PR, m - tm * PR);
            t.nlen = (int32_t)std::min<int64_t>(PR, n - tn * PR);
            t.seg0 = (int32_t)segs.size();
            t.nseg = 1;
            SegDesc d{};
            d.a_base = tm * nkt * chunk;
            d.b_base = tn * nkt * chunk;
            d.klen = k;
            d.a_kblk = d.b_kblk = -1;
            segs.push_back(d);
            tasks.push_back(t);
        }
    plan.packed = true;
    plan.useful_flops = 2 * g.m * g.n * k;
    plan.mfma_flops = 2 * PR * PR * nkt * PK * (int64_t)tasks.size() * 6;
    return true;
}

std::shared_ptr<Plan> PlanBuilder::build()
{
    if (gemms_.empty() && panel_gemms_->empty()) return nullptr;
    auto plan = std::make_shared<Plan>();
    if (!panel_gemms_->empty() || panel_route())
    {
        if (!panel_route()) fail("internal: block GEMMs need the panel path");
        std::vector<PanelGemm> pg(*panel_gemms_);
        for (auto& g : gemms_) panel_from_desc(g, pg);
        if (!build_panel_plan(*plan, std::move(pg), elem_size_, dtype_, family_)) return nullptr;
        return plan;
    }

    // Orientation: the MFMA output puts consecutive lanes along n, so make
    // n the unit-stride side of C when C is contiguous along m.
    const GemmDesc& g0 = gemms_[0];
    plan->swap = unit_first(g0.c_rows) && !unit_first(g0.c_cols);
    const bool sw = plan->swap;

    // Thread maps from the first segment that exists.
    plan->key.a_mfast = 1;
    plan->key.b_nfast = 1;
    for (auto& g : gemms_)
    {
        if (g.segs.empty()) continue;
        const GemmSeg& s = g.segs[0];
        const Side& ar = sw ? s.b_cols : s.a_rows;
        const Side& ac = sw ? s.b_rows : s.a_cols;
        const Side& br = sw ? s.a_cols : s.b_rows;
        const Side& bc = sw ? s.a_rows : s.b_cols;
        plan->key.a_mfast = rows_fast(ar, ac);
        plan->key.b_nfast = rows_fast(bc, br);
        break;
    }
    plan->key.dtype = dtype_;
    plan->key.use_mfma = use_mfma_;
    plan->key.family = family_;

    // Tile choice: a cost model over the candidate tiles.  Per resident
    // workgroup each tile sustains a rate measured on MI355X (FP64 MFMA,
    // GFLOP/s per workgroup at full occupancy); the estimated time is the
    // padded work spread over min(tasks, resident slots) workgroups, but no
    // less than the largest single task.
    struct Cand
    {
        int tile;
        double gflops_per_wg;
    };
    static const Cand f64_cands[] = {{1, 33.8}, {5, 11.1}, {7, 66.0}, {4, 113.0}};
    static const Cand bf16_cands[] = {{0, 600.0}, {1, 180.0}, {2, 350.0}};
    const Cand* cands = family_ == FAMILY_F64 ? f64_cands : bf16_cands;
    const int ncand = family_ == FAMILY_F64 ? 4 : 3;
    plan->key.tile = family_ == FAMILY_F64 ? 1 : 0;
    if (use_mfma_)
    {
        double best = 0;
        for (int ci = 0; ci < ncand; ci++)
        {
            const Cand& c = cands[ci];
            KernelKey k = plan->key;
            k.tile = c.tile;
            const int64_t bm = tile_m(k), bn = tile_n(k), bk = tile_k(k);
            double work = 0, maxwork = 0;
            int64_t ntask = 0;
            for (auto& g : gemms_)
            {
                int64_t m = sw ? g.n : g.m, n = sw ? g.m : g.n;
                int64_t L = g.l_c.empty() ? 1 : (int64_t)g.l_c.size();
                int64_t kpad = 0;
                for (auto& s : g.segs) kpad += ceil_div(s.k, bk) * bk;
                double tw = 2.0 * bm * bn * std::max<int64_t>(kpad, bk);
                int64_t nt = ceil_div(m, bm) * ceil_div(n, bn) * L;
                ntask += nt;
                work += tw * nt;
                maxwork = std::max(maxwork, tw);
            }
            int64_t slots = (int64_t)device_cu_count() * std::max(1, occupancy(k));
            double rate = c.gflops_per_wg * 1e9;
            double t = std::max(work / (std::min(ntask, slots) * rate), maxwork / rate);
            if (best == 0 || t < best)
            {
                best = t;
                plan->key.tile = c.tile;
            }
        }
    }
    // TBLIS_HIP_TILE=<id> forces the tile (tuning runs)
    if (const char* e = std::getenv("TBLIS_HIP_TILE"))
    {
        int t = std::atoi(e);
        if (use_mfma_ && e[0] >= '0' && e[0] <= '9' && t < num_tiles(family_)) plan->key.tile = t;
    }
    const int BM = tile_m(plan->key), BN = tile_n(plan->key);
    const int BK = tile_k(plan->key);

    std::vector<TaskDesc> tasks;
    std::vector<SegDesc> segs;
    std::vector<int64_t> cost;
    const bool packed = try_packed(*plan, tasks, segs);
    for (auto& g : gemms_)
    {
        if (packed) break;
        const int64_t m = sw ? g.n : g.m, n = sw ? g.m : g.n;
        if (m == 0 || n == 0) continue;
        const Side& cr = sw ? g.c_cols : g.c_rows;
        const Side& cc = sw ? g.c_rows : g.c_cols;
        const int64_t L = g.l_c.empty() ? 1 : (int64_t)g.l_c.size();
        int64_t ksum = 0;
        for (auto& s : g.segs) ksum += s.k;
        for (int64_t l = 0; l < L; l++)
        {
            int64_t oa = g.l_a.empty() ? 0 : g.l_a[l];
            int64_t ob = g.l_b.empty() ? 0 : g.l_b[l];
            int64_t oc = g.l_c.empty() ? 0 : g.l_c[l];
            if (sw) std::swap(oa, ob);
            plan->useful_flops += 2 * g.m * g.n * ksum;
            for (int64_t tm = 0; tm < ceil_div(m, BM); tm++)
                for (int64_t tn = 0; tn < ceil_div(n, BN); tn++)
                {
                    TaskDesc t{};
                    t.c_base = g.c_base + oc;
                    t.c_row = side_ref(cr, BM, tm * BM);
                    t.c_col = side_ref(cc, BN, tn * BN);
                    t.mlen = (int32_t)std::min<int64_t>(BM, m - tm * BM);
                    t.nlen = (int32_t)std::min<int64_t>(BN, n - tn * BN);
                    t.seg0 = (int32_t)segs.size();
                    t.nseg = 0;
                    int64_t kpad = 0;
                    for (auto& s : g.segs)
                    {
                        if (s.k == 0) continue;
                        SegDesc d{};
                        const Side& ar = sw ? s.b_cols : s.a_rows;
                        const Side& ac = sw ? s.b_rows : s.a_cols;
                        const Side& br = sw ? s.a_cols : s.b_rows;
                        const Side& bc = sw ? s.a_rows : s.b_cols;
                        d.a_base = (sw ? s.b_base : s.a_base) + oa;
                        d.b_base = (sw ? s.a_base : s.b_base) + ob;
                        d.a_row = side_ref(ar, BM, tm * BM);
                        d.a_col = side_ref(ac, BK, 0);
                        d.b_row = side_ref(br, BK, 0);
                        d.b_col = side_ref(bc, BN, tn * BN);
                        d.klen = s.k;
                        d.a_kblk = kblock_table(ac, BK);
                        d.b_kblk = kblock_table(br, BK);
                        segs.push_back(d);
                        t.nseg++;
                        kpad += ceil_div(s.k, BK) * BK;
                    }
                    plan->mfma_flops += 2 * (int64_t)BM * BN * kpad;
                    tasks.push_back(t);
                    cost.push_back((int64_t)t.mlen * t.nlen * std::max<int64_t>(kpad, 1));
                }
        }
    }
    if (tasks.empty()) return nullptr;
    plan->reads_ab = !segs.empty();
    plan->elem = elem_size_;

    // Uneven work (DPD, batches): longest tasks first so the static
    // round-robin over the persistent grid balances.
    bool uniform = cost.empty() ||
                   std::all_of(cost.begin(), cost.end(), [&](int64_t c) { return c == cost[0]; });
    if (!uniform)
    {
        std::vector<int> order(tasks.size());
        for (size_t i = 0; i < order.size(); i++) order[i] = (int)i;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
        std::vector<TaskDesc> t2;
        t2.reserve(tasks.size());
        for (int i : order) t2.push_back(tasks[i]);
        tasks.swap(t2);
    }
    if (tasks.size() > (size_t)INT32_MAX || segs.size() > (size_t)INT32_MAX)
        fail("contraction too large: task list exceeds 2^31 entries");

    plan->ntasks = (int)tasks.size();
    int occ = packed ? gemm_bf16x3_packed_occupancy(family_ == FAMILY_BF16X3) : occupancy(plan->key);
    plan->grid = (int)std::min<int64_t>(plan->ntasks, (int64_t)device_cu_count() * std::max(occ, 1));

    size_t tb = tasks.size() * sizeof(TaskDesc);
    size_t sb = segs.size() * sizeof(SegDesc);
    size_t pb = std::max<size_t>(pool_.size(), 1) * sizeof(int64_t);
    size_t off_s = (tb + 255) / 256 * 256;
    size_t off_p = off_s + (sb + 255) / 256 * 256;
    plan->bytes = off_p + pb;
    hip_check(hipGetDevice(&plan->device), "hipGetDevice");
    hip_check(device_alloc(&plan->dev, plan->bytes), "hipMalloc (plan)");
    char* base = (char*)plan->dev;
    plan->d_tasks = (TaskDesc*)base;
    plan->d_segs = (SegDesc*)(base + off_s);
    plan->d_pool = (int64_t*)(base + off_p);
    hip_check(hipMemcpy(plan->d_tasks, tasks.data(), tb, hipMemcpyHostToDevice), "hipMemcpy");
    if (sb) hip_check(hipMemcpy(plan->d_segs, segs.data(), sb, hipMemcpyHostToDevice), "hipMemcpy");
    if (!pool_.empty())
        hip_check(hipMemcpy(plan->d_pool, pool_.data(), pool_.size() * sizeof(int64_t),
                            hipMemcpyHostToDevice),
                  "hipMemcpy");
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

void launch_panel_calls(const PanelLaunch& P, const int64_t group_n[4])
{
    if (P.use_k) hip_check(launch_panel_pack(P), "pack launch");
    // one GEMM launch per group of tasks with the same direct operands
    int64_t at = 0;
    for (int g = 0; g < 4; g++)
    {
        if (group_n[g] == 0) continue;
        PanelLaunch G = P;
        G.tasks = P.tasks + at;
        G.ntasks = (int)group_n[g];
        G.direct_a = (g & 1) != 0;
        G.direct_b = (g & 2) != 0;
        hip_check(launch_panel_gemm(G), "kernel launch");
        at += group_n[g];
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
    ts.last_tile = p.panel ? p.panel_tile : -1;
    GemmLaunch L;
    L.tasks = p.d_tasks;
    L.ntasks = p.ntasks;
    L.segs = p.d_segs;
    L.pool = p.d_pool;
    L.A = p.swap ? B : A;
    L.B = p.swap ? A : B;
    L.C = C;
    L.alpha = alpha;
    L.beta = beta;
    L.use_k = use_k && p.reads_ab ? 1 : 0;
    L.grid = p.grid;
    L.stream = ts.stream;
    if (p.panel)
    {
        PanelLaunch P{};
        P.chunks = p.d_pchunks;
        P.nchunks = p.p_nchunks;
        P.tasks = p.d_ptasks;
        P.ntasks = p.ntasks;
        P.ujobs = p.d_ujobs;
        P.uchunks = p.d_uchunks;
        P.nuchunks = p.p_nuchunks;
        P.pool = p.d_pool;
        P.A = L.A;
        P.B = L.B;
        P.C = C;
        char* ws = p.ws_static ? p.ws_static : (char*)workspace(p.ws_bytes, ts.stream);
        P.ws = ws;
        P.cws = (double*)(ws + p.ws_c_offset);
        P.family = p.key.family;
        P.direct_a = p.direct[0];
        P.direct_b = p.direct[1];
        P.alpha = alpha;
        P.beta = beta;
        P.use_k = L.use_k;
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
            if (!ts.async && !capturing) hip_check(hipStreamSynchronize(ts.stream), "hipStreamSynchronize");
            return;
        }
        launch_panel_calls(P, p.group_n);
    }
    else if (p.packed && L.use_k)
    {
        __bf16* ws = (__bf16*)workspace(p.ws_bytes, ts.stream);
        const void* src[2] = {L.A, L.B};
        for (int o = 0; o < 2; o++)
            hip_check(launch_pack_bf16x3(p.pack[o], p.pack_kfast[o], p.d_pool, (const float*)src[o], ws,
                                         ts.stream),
                      "pack launch");
        L.A = ws + p.pack[0].out;
        L.B = ws + p.pack[1].out;
        hip_check(launch_gemm_bf16x3_packed(L, p.key.family == FAMILY_BF16X3), "kernel launch");
    }
    else if (p.packed)
    {
        // alpha == 0: only C <- beta*C; the panel kernel skips its k loop
        hip_check(launch_gemm_bf16x3_packed(L, p.key.family == FAMILY_BF16X3), "kernel launch");
    }
    else
    {
        hip_check(launch_gemm(L, p.key), "kernel launch");
    }
    if (!ts.async && !capturing) hip_check(hipStreamSynchronize(ts.stream), "hipStreamSynchronize");
}

}  // namespace tblis_amd
