// FP64 panel path: GEMMs assembled from operand blocks, packed once into
// fragment-ordered panels (plan_types.hpp) and multiplied by the panel GEMM.
#pragma once

#include <vector>

#include "plan.hpp"

namespace tblis_amd
{

// Rows x k of one operand block; base == PANEL_MISSING means zeros.
struct PBlock
{
    int64_t base = PANEL_MISSING;
    Side rows, k;
};

struct CBlock
{
    int64_t base = 0;
    Side rows, cols;
};

// One GEMM over blocks: m = sum(mlen), n = sum(nlen), k = sum(klen).  A block
// (p, q) is a[p * nk + q] (rows along m), B block (s, q) is b[s * nk + q]
// (rows along n), C block (p, s) is c[p * nn + s].  A dense contraction is
// one block of each; a DPD irrep GEMM has one block per irrep combination.
struct PanelGemm
{
    std::vector<int64_t> mlen, nlen, klen;
    std::vector<PBlock> a, b;
    std::vector<CBlock> c;
    std::vector<double> kscale;  // per k block, applied to A while packing (empty: all 1)
    // imaginary parts of the k-block scales of a complex GEMM (empty: all
    // 0); complex_split_gemms turns them into real term scales
    std::vector<double> kscale_im;
};

// C^T = B^T A^T.
PanelGemm transpose_gemm(const PanelGemm& g);

// The planner's offset vector of one side (odometer order, first dimension
// fastest) and its (base, stride) table at fragment-group size G for the
// side placed at panel row r0 (introspection for the parity tests).
void debug_side_tables(const Side& s, int64_t G, int64_t r0, std::vector<int64_t>& offs,
                       std::vector<int64_t>& tab);

// One PanelGemm per batch element of a GemmDesc (one C patch with its k
// segments).
void panel_from_desc(const GemmDesc& g, std::vector<PanelGemm>& out);


// FP64 operands whose layout the GEMM can read directly skip packing;
// TBLIS_HIP_DIRECT=0 packs them anyway.
bool direct_enabled();
// GEMMs over several C blocks write them from the GEMM epilogue through
// per-row/column block codes; TBLIS_HIP_C_UNPACK=1 stores a C panel and
// scatters it with the unpack kernel instead.
bool c_unpack_forced();
// Re-read limit (tiles) under which a k-fast operand of at most 4 MiB is
// read directly (TBLIS_HIP_DIRECT_KFAST, default 32).
int direct_kfast_rereads();

// Host tables of a panel plan.  Stages (plans of many GEMMs): consecutive
// ranges of GEMMs of about equal flops whose pack launches can run on a
// second stream ahead of the GEMM launches.
struct PanelTables
{
    std::vector<PanelJob> jobs;
    std::vector<PanelChunk> chunks;
    std::vector<PanelTask> tasks;  // longest first
    std::vector<UnpackJob> ujobs;
    std::vector<UnpackChunk> uchunks;
    std::vector<int64_t> pool;
    int tile = 0;
    int family = 0;
    bool swap = false;  // GEMMs were transposed: operand 0 is the B tensor
    bool reads_ab = false;
    bool direct[2] = {false, false};  // operand read from the tensor by every task
    // tasks come in four consecutive groups by which operands they read from
    // the tensors (none, A, B, both); one GEMM launch per non-empty group
    int64_t group_n[4] = {0, 0, 0, 0};
    // chunks and tasks ordered by stage: sched.chunks[s] pack chunks and
    // sched.group_n[s][g] tasks of group g in stage s (group-major within
    // the stage); task_stage is each task's stage (host bookkeeping)
    PanelSchedule sched;
    std::vector<int8_t> task_stage;
    int64_t ws_panel = 0;  // panel values (double for FP64, BF16 for split-BF16, bytes for Ozaki)
    int64_t ws_rows = 0;   // Ozaki: row-maximum entries (one per panel row)
    int64_t ws_c = 0;      // C panel doubles
    int64_t useful_flops = 0, mfma_flops = 0;
};

// Can the LDS-staged FP64 tiles run these GEMMs (single blocks, operands in
// reach of a buffer resource)?  src_elem: bytes per tensor element.
bool lds_applicable(const std::vector<PanelGemm>& gemms, int src_elem);

// Orientation, tile, pack jobs, tasks and C unpack jobs (host only).
// Returns false when there is nothing to launch.  dtype: 0 double, 1 float.
bool plan_panel_tables(std::vector<PanelGemm> gemms, PanelTables& T, int family, int dtype);

// The pack kernels' chunk records (job fields and group tables per chunk).
std::vector<PackChunk> pack_records(const PanelTables& T, int family);

// Tensors the autotuner may read, never write, while it times candidate
// tiles: those of the call being planned, already checked by the caller
// (set for the life of a TuneOperandsScope on this thread).  With them an
// FP64 candidate is timed as it will run -- pack, then one GEMM launch per
// group of tasks with the same operands read straight from the tensors --
// since which operands can be read directly depends on the tile.  Without
// them every candidate is timed on scratch panels.
struct TuneOperandsScope
{
    TuneOperandsScope(const void* A, const void* B);
    ~TuneOperandsScope();
};

// While one of these lives on a thread, plans built there keep the cost
// model's tile (the autotuner allocates device memory for its timing runs).
struct AutotuneOffScope
{
    AutotuneOffScope();
    ~AutotuneOffScope();

  private:
    bool was_;
};

// While one of these lives on a thread, FP64 plans built there read every
// operand that the GEMM can read straight from its tensor (re-read by at
// most 64 tiles) instead of weighing it against a pack pass: a batch of
// small instances is then one GEMM launch with no pack (the 4096-instance
// batch: 23.15 against 23.37 ms with the pack; the merged DPD irrep GEMMs
// are slower that way and do not use it).
struct DirectAllScope
{
    DirectAllScope();
    ~DirectAllScope();

  private:
    bool was_;
};

// Autotuner switch (initially TBLIS_HIP_AUTOTUNE) and the number of tuning
// measurements made by this process.
void set_autotune(int on);
int autotune_runs();

// plan_panel_tables + upload into the plan.
bool build_panel_plan(Plan& plan, std::vector<PanelGemm> gemms, int elem_size, int dtype, int family);

}  // namespace tblis_amd
