// Internal types shared by the host side of libtblis_amd.
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tblis.h"

namespace tblis_amd
{

// Argument or environment error; the C ABI turns it into an abort with a
// message (default) or a stored error string.
struct Error : std::runtime_error
{
    explicit Error(const std::string& s) : std::runtime_error(s) {}
};

[[noreturn]] inline void fail(const std::string& s) { throw Error(s); }

// One dimension of a contraction group: its label, length, and the stride
// it has in each tensor of the group (tensor order fixed per group, see
// Group::who).
struct Dim
{
    char label = 0;
    int64_t len = 1;
    int64_t s[3] = {0, 0, 0};
};

using Group = std::vector<Dim>;

// Canonical form after diagonal / group split / fold
// (src/iface/3t/mult.cxx:24-116).  Stride slots per group:
//   ab : s[0]=A  s[1]=B
//   ac : s[0]=A  s[1]=C
//   bc : s[0]=B  s[1]=C
//   abc: s[0]=A  s[1]=B  s[2]=C
struct Canonical
{
    Group ab, ac, bc, abc;
};

// Matrix views handed to the GEMM after the stride sort and pack_3d rotation
// of mult_blis (src/internal/3t/dense/mult.cxx:212-263, :479-556).  Slots:
//   m: s[0]=A s[1]=C     n: s[0]=B s[1]=C
//   k: s[0]=A s[1]=B     l: s[0]=A s[1]=B s[2]=C
struct MatrixViews
{
    Group m, n, k, l;
    bool pack_m_3d = false, pack_n_3d = false, pack_k_3d = false;
};

// (len, stride) list of one matrix side for one tensor.
// A side may instead list its element offsets explicitly (offs non-empty;
// len and stride then unused): a DPD irrep GEMM's merged rows or k, whose
// offsets are per-block bases plus the blocks' own odometer offsets.
struct Side
{
    std::vector<int64_t> len, stride;
    bool pack_3d = false;
    std::vector<int64_t> offs;
};

inline int64_t side_count(const Side& s)
{
    if (!s.offs.empty()) return (int64_t)s.offs.size();
    int64_t n = 1;
    for (auto l : s.len) n *= l;
    return n;
}

// Stride between the first two elements (0 when the side has one element).
inline int64_t first_step(const Side& s)
{
    if (!s.offs.empty()) return s.offs.size() > 1 ? s.offs[1] - s.offs[0] : 0;
    return !s.len.empty() && s.len[0] > 1 ? s.stride[0] : 0;
}

inline int64_t prod_len(const Group& g)
{
    int64_t p = 1;
    for (auto& d : g) p *= d.len;
    return p;
}

inline Side side_of(const Group& g, int slot, bool pack_3d)
{
    Side s;
    for (auto& d : g)
    {
        s.len.push_back(d.len);
        s.stride.push_back(d.s[slot]);
    }
    s.pack_3d = pack_3d;
    return s;
}

// canon.cpp
// Level-1 form of B <- alpha A + beta B (src/iface/1t/add.cxx:24-110):
// groups of labels in both tensors (s[0] = A stride, s[1] = B stride),
// only in A (summed, s[0]) and only in B (replicated, s[1]), each folded.
struct Level1Form
{
    Group ab, a_only, b_only;
};
bool canonicalize_add(const tblis_tensor& A, const label_type* idx_A, const tblis_tensor& B,
                      const label_type* idx_B, Level1Form& out, std::string& why);
// One tensor (scale, set): diagonal taken, length-1 dims dropped, folded (s[1]).
bool canonicalize_one(const tblis_tensor& A, const label_type* idx_A, Group& out, std::string& why);

bool canonicalize(const tblis_tensor& A, const label_type* idx_A, const tblis_tensor& B,
                  const label_type* idx_B, const tblis_tensor& C, const label_type* idx_C,
                  Canonical& out, std::string& why);
MatrixViews blis_views(const Canonical& c);

// scatter.cpp
void fill_scatter(const Side& side, int64_t BS, int64_t off, int64_t size, int64_t* scat,
                  int elem_size);
void fill_block_stride(int64_t BS, int64_t size, const int64_t* scat, int64_t* bs);

// Algorithm selection (tblis_config objects are instances of this).
enum class Algo : int
{
    Default = 0,
    Fp64Mfma = 1,
    Bf16x3 = 2,
    Bf16x3Fast = 3,
    Reference = 4,
    Ozaki = 5,  // FP64-accurate products from int8 MFMA digit planes
};

const char* algo_name(Algo a);

}  // namespace tblis_amd

struct tblis_config_s
{
    tblis_amd::Algo algo;
    const char* name;
};
