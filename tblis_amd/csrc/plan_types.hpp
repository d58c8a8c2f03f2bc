// Device-side description of a contraction on the panel path: pack
// chunks, GEMM tasks (C tiles) and C unpack jobs, with the offset tables
// they index in one int64 pool.  Shared by the host planner and the
// kernels.  A dense contraction has one block per operand; a DPD
// contraction one per irrep block; a batch one per instance.
#pragma once

#include <cstdint>

namespace tblis_amd
{

constexpr int64_t SCATTERED = INT64_MIN;

// ---------------------------------------------------------------- panels
// FP64 panel path.  Each GEMM operand (A': rows = m, B': rows = n, both
// against k) is gathered once into a fragment-ordered panel: 16-row by
// 8-k fragment groups of 128 doubles, group (rb, kg) at
//     out + (rb * pnkg + kg) * 128,
// lane l of a wave holding X[16 rb + l % 16][8 kg + l / 16] and
// X[16 rb + l % 16][8 kg + 4 + l / 16] as one 16-byte pair -- exactly the
// operand of two v_mfma_f64_16x16x4 steps.  Rows and k are padded with
// zeros to the tile and to 16.
constexpr int64_t PANEL_MISSING = INT64_MIN;

// One block of one operand (a dense tensor has one block per operand, a
// DPD irrep GEMM one per (row irreps, k irreps) pair): panel rows
// [r0, r0 + nr) x k [k0, k0 + nk), element (i, kk) read from
//     X[src + pool[rowmap + i] + pool[kmap + kk]],
// or zeros when src == PANEL_MISSING (padding, absent blocks).
// Group tables (rowtab, ktab) hold one (base, stride) pair per fragment
// group the job touches, counted from group r0 / rows-per-group (k0 /
// k-per-group): element i of the group is at base + i * stride, i counted
// from the group's first panel row (k), or stride == SCATTERED and the
// per-element maps apply -- the reference's block strides (rbs / cbs,
// src/matrix/block_scatter_matrix.hpp:232-247) at fragment-group size.
struct PanelJob
{
    int64_t src;
    int64_t rowmap, kmap;  // pool indices
    int64_t rowtab, ktab;  // pool indices of the group tables
    int64_t out;           // panel offset (values) in the workspace
    int32_t r0, nr, k0, nk;
    int32_t pnkg;          // k groups of the whole panel
    int32_t operand;       // 0: A pointer, 1: B pointer (after orientation)
    int32_t kfast;         // k groups unit-stride, rows not: read rows, transpose in LDS
    int32_t pad_;
    double scale = 1.0;    // every value is multiplied by this while packed
    int64_t eoff = 0;      // Ozaki family: row-maximum array index of panel row 0
};

// Rectangle of fragment groups of one job: rows blocks [rb0, rb0 + nrb),
// k groups [kg0, kg0 + nkg).
struct PanelChunk
{
    int32_t job, rb0, nrb, kg0, nkg, pad;
};

// A pack chunk as the pack kernels read it: its job's fields and the
// (base, stride) pairs of the chunk's own row and k groups (row group
// rb0 + i at rt[i], k group kg0 + i at kt[i]; unset for zero chunks), so
// that every data address of the chunk follows from this one record.
struct PackChunk
{
    int64_t src;           // element offset of the block, or PANEL_MISSING (zeros)
    int64_t out;           // panel offset (values) in the workspace
    int64_t rowmap, kmap;  // pool indices of the per-element maps (scattered groups)
    double scale;          // applied to every value
    int32_t r0, nr, k0, nk;
    int32_t pnkg, operand, kfast, pad;
    int32_t rb0, nrb, kg0, nkg;
    int64_t rt[8][2];
    int64_t kt[8][2];
    int64_t eoff;  // Ozaki family: row-maximum array index of the panel's row 0
};

// One C tile of one panel GEMM.  An operand is either read from its panel
// (a, b) or straight from the tensor ("direct"): through a buffer resource
// at tensor + x_src spanning x_bytes bytes; each lane takes its row offsets
// from the explicit row offsets once per task, and per 8-k group adds the
// group's (base + lane k * stride) or, for a group that is not affine
// (stride SCATTERED), its own k offset; lanes past the valid rows or k read
// offset 0xffffffff, which the buffer bounds check returns as zero.
struct PanelTask
{
    int64_t a, b;        // workspace offsets (values) of the tile's first row block of A', B'
    int64_t c_src;       // C: element offset of the C block; C panel: workspace offset
    int64_t c_rowmap;    // pool index of the tile's first row offset (direct epilogue)
    int64_t c_colmap;    // pool index of the tile's first col offset
    int32_t nkg;         // k groups (even)
    int32_t mlen, nlen;  // valid rows / cols of the tile
    // 0: write C directly through c_rowmap / c_colmap offsets; > 0: leading
    // dimension of a C panel at c_src; < 0: C spans -c_ld column patches
    // (per-row and per-column codes (patch << 32 | index) at c_rowmap /
    // c_colmap, a (base, row offsets, column offsets) triple per block at
    // pool[c_src], row-patch major)
    int32_t c_ld;
    // direct operands (FP64 family)
    int64_t a_src, b_src;  // element offset of the block in its tensor
    // pool index of the offset of the tile's first row; in the Ozaki family
    // the row-maximum array index of the tile's first row of A' / B'
    int64_t a_rt, b_rt;
    int64_t a_kt, b_kt;    // pool index of k group 0's (base, stride)
    int64_t a_km, b_km;    // pool index of the per-element k offsets
    uint32_t a_bytes, b_bytes;  // extent of the block from x_src
    int32_t klen;          // valid k
    int32_t pad_;
};

// C panel -> C block: C[dst + rowmap[i] + colmap[j]] = alpha * P[i * ld + j] + beta * C[...]
struct UnpackJob
{
    int64_t dst, rowmap, colmap, panel;
    int32_t nr, nc, ld, rows_fast;
};

struct UnpackChunk
{
    int32_t job, r0, c0, pad;  // 64 x 64 element sub-block
};

// Kernel families.
constexpr int FAMILY_F64 = 0;          // FP64 MFMA (FP32 data widened)
constexpr int FAMILY_BF16X3 = 1;       // FP32 as 3-term split BF16, chunked accumulation
constexpr int FAMILY_BF16X3_FAST = 2;  // same, single accumulator
constexpr int FAMILY_OZAKI = 3;        // FP64-accurate: int8 digit planes (gemm_ozaki.hip)

// Launch schedule of a panel plan: chunks and tasks ordered by stage;
// stage s has chunks[s] pack chunks and group_n[s][g] GEMM tasks reading
// operands directly by group g (0 none, 1 A, 2 B, 3 both), in that order.
constexpr int kMaxStages = 8;
struct PanelSchedule
{
    int nstages = 1;
    int64_t chunks[kMaxStages] = {};
    int64_t group_n[kMaxStages][4] = {};
};

// Kernel variant selector.
struct KernelKey
{
    int family;  // FAMILY_*
    int tile;    // index into the family's tile table
    int dtype;   // 0: double in/out, 1: float in/out
};

}  // namespace tblis_amd
