// Block-scatter vectors: the element offset of every logical row (or
// column) of a matrix view, and the per-block common stride.
//
// Same sequence as the reference's block_scatter_matrix
// (src/matrix/block_scatter_matrix.hpp:68-247):
//   - plain order: logical index i -> offset of the multi-index of i with the
//     first dimension fastest (viterator odometer order), for i in
//     [off, off+size);
//   - pack_3d order: the first two dimensions are walked in
//     cache-line-sized 2-D blocks (BS0 x BS1, CL = 64 bytes of elements),
//     remainder columns after, then the partial first/last rows of the range;
//   - block stride: for each run of BS entries, the common difference if the
//     run is an arithmetic progression, else 0 (1 for a run of length 1).
// The GPU kernels consume these vectors directly (tile loads and C stores
// go through them), so the tested vector is the one used.
#include <algorithm>
#include <cstring>

#include "common.hpp"

namespace tblis_amd
{
namespace
{

// Binary gcd exactly as the reference computes it (one shift per
// subtraction; block_scatter_matrix.hpp:18-42), so lcm(BS, CL) and the
// pack_3d block sizes match it bit for bit.
int64_t binary_gcd(int64_t a, int64_t b)
{
    a = a < 0 ? -a : a;
    b = b < 0 ? -b : b;
    if (a == 0) return b;
    if (b == 0) return a;
    int d = __builtin_ctzll((unsigned long long)(a | b));
    a >>= __builtin_ctzll((unsigned long long)a);
    b >>= __builtin_ctzll((unsigned long long)b);
    while (a != b)
    {
        if (a > b)
            a = (a - b) >> 1;
        else
            b = (b - a) >> 1;
    }
    return a << d;
}

// Odometer over the outer dimensions (first fastest).
struct Odometer
{
    const int64_t* len;
    const int64_t* stride;
    int n;
    int64_t digit[32];
    int64_t offset = 0;

    Odometer(const int64_t* l, const int64_t* s, int nd, int64_t pos) : len(l), stride(s), n(nd)
    {
        for (int i = 0; i < n; i++)
        {
            digit[i] = pos % len[i];
            pos /= len[i];
            offset += digit[i] * stride[i];
        }
    }

    // Advance; false once the last position has been passed.
    bool advance()
    {
        for (int i = 0; i < n; i++)
        {
            if (digit[i] + 1 < len[i])
            {
                digit[i]++;
                offset += stride[i];
                return true;
            }
            offset -= digit[i] * stride[i];
            digit[i] = 0;
        }
        return false;
    }
};

}  // namespace

void fill_scatter(const Side& side, int64_t BS, int64_t off, int64_t size, int64_t* scat,
                  int elem_size)
{
    if (size == 0) return;
    int nd = (int)side.len.size();
    if (nd == 0)
    {
        scat[0] = 0;
        return;
    }
    const int64_t CL = 64 / elem_size;
    const int64_t m0 = side.len[0], s0 = side.stride[0];

    if (!side.pack_3d)
    {
        const int64_t* ol = side.len.data() + 1;
        const int64_t* os = side.stride.data() + 1;
        for (int i = 0; i < nd - 1; i++)
            if (ol[i] == 0) return;
        Odometer outer(ol, os, nd - 1, off / m0);
        int64_t i0 = off % m0, idx = 0;
        while (idx < size)
        {
            int64_t run = std::min(m0 - i0, size - idx);
            int64_t p = outer.offset + i0 * s0;
            for (int64_t i = 0; i < run; i++, p += s0) scat[idx++] = p;
            i0 = 0;
            if (idx < size && !outer.advance()) break;
        }
        return;
    }

    // pack_3d: needs at least two dimensions.
    const int64_t m1 = side.len[1], s1 = side.stride[1];
    const int64_t* ol = side.len.data() + 2;
    const int64_t* os = side.stride.data() + 2;
    int no = nd - 2;
    for (int i = 0; i < no; i++)
        if (ol[i] == 0) return;

    // The reference's binary gcd returns 0 for some block sizes (e.g. 5 with
    // a power-of-two CL) and its pack_3d fill then divides by zero.
    const int64_t g = binary_gcd(BS, CL);
    if (g == 0) fail("pack_3d scatter: block size " + std::to_string(BS) +
                     " has no cache-line order in the reference (its gcd yields 0)");
    const int64_t lcm_bs_cl = BS * CL / g;
    const int64_t BS0 = std::min(lcm_bs_cl, std::max(BS, m0 - m0 % BS));
    const int64_t BS1 = std::min(CL, std::max(BS, m1 - m1 % BS));

    const int64_t plane = m0 * m1;
    const int64_t first = off % plane, last = (off + size - 1) % plane;
    const int64_t p0 = first % m0, q0 = last % m0;
    int64_t p1 = first / m0, q1 = last / m0;
    const int64_t r0 = q0 - q0 % BS0;
    const int64_t b_min = off / plane, b_max = (off + size - 1) / plane + 1;
    if (p0 == 0) p1--;
    if (q0 == m0 - 1) q1++;

    int64_t idx = 0;
    auto rows_of = [&](int64_t b, int64_t& lo, int64_t& hi) {
        lo = b == b_min ? p1 + 1 : 0;
        hi = b == b_max - 1 ? q1 : m1;
    };

    // Full BS0 x BS1 tiles of the first r0 entries of dimension 0.
    if (r0 > 0)
    {
        Odometer outer(ol, os, no, b_min);
        for (int64_t b = b_min; b < b_max; b++)
        {
            if (b > b_min && !outer.advance()) break;
            int64_t lo, hi;
            rows_of(b, lo, hi);
            for (int64_t i1 = lo; i1 < hi; i1 += BS1)
                for (int64_t i0 = 0; i0 < r0; i0 += BS0)
                    for (int64_t j1 = 0; j1 < std::min(BS1, hi - i1); j1++)
                        for (int64_t j0 = 0; j0 < BS0; j0++)
                            scat[idx++] = outer.offset + (i0 + j0) * s0 + (i1 + j1) * s1;
        }
    }

    // Remainder of dimension 0 for the full rows, then the partial first and
    // last rows of the range.
    Odometer outer(ol, os, no, b_min);
    for (int64_t b = b_min; b < b_max; b++)
    {
        if (b > b_min && !outer.advance()) break;
        int64_t lo, hi;
        rows_of(b, lo, hi);
        for (int64_t j1 = lo; j1 < hi; j1++)
            for (int64_t j0 = r0; j0 < m0; j0++) scat[idx++] = outer.offset + j0 * s0 + j1 * s1;

        bool head = b == b_min && p0 > 0;
        bool tail = b == b_max - 1 && q0 < m0 - 1;
        if (head && tail && p1 == q1)
        {
            for (int64_t j0 = p0; j0 <= q0; j0++) scat[idx++] = outer.offset + j0 * s0 + p1 * s1;
        }
        else
        {
            if (head)
                for (int64_t j0 = p0; j0 < m0; j0++)
                    scat[idx++] = outer.offset + j0 * s0 + p1 * s1;
            if (tail)
                for (int64_t j0 = 0; j0 <= q0; j0++)
                    scat[idx++] = outer.offset + j0 * s0 + q1 * s1;
        }
    }
    if (idx != size) fail("internal: pack_3d scatter produced a wrong number of entries");
}

void fill_block_stride(int64_t BS, int64_t size, const int64_t* scat, int64_t* bs)
{
    for (int64_t i = 0; i < size; i += BS)
    {
        int64_t n = std::min(size - i, BS);
        int64_t s = n > 1 ? scat[i + 1] - scat[i] : 1;
        for (int64_t j = i + 1; j + 1 < i + n; j++)
            if (scat[j + 1] - scat[j] != s) s = 0;
        bs[i] = s;
    }
}

}  // namespace tblis_amd
