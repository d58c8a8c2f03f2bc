/*
 * tblis.hpp -- C++ front end of libtblis_amd, mirroring the reference's
 * tblis::mult<T> templates (src/iface/3t/mult.h:28-86).
 *
 * The reference's templates take MArray views (varray_view / dpd_varray_view)
 * and build tblis_tensor descriptors for the C call; these do the same over
 * two small view types that carry exactly what the descriptors need, so a
 * caller holding MArray views passes view.data(), view.lengths() and
 * view.strides() into varray_view, and for a DPD view its data(), irrep(),
 * num_irreps() and per-dimension lengths into dpd_varray_view.
 */
#ifndef TBLIS_AMD_TBLIS_HPP
#define TBLIS_AMD_TBLIS_HPP

#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "tblis.h"

namespace tblis
{

// Strided dense tensor view (device memory), like MArray's varray_view.
template <typename T>
struct varray_view
{
    T* data;
    std::vector<len_type> len;
    std::vector<stride_type> stride;
};

// Strided matrix view (device memory), like MArray's matrix_view.
template <typename T>
struct matrix_view
{
    T* data;
    len_type m, n;
    stride_type rs, cs;
};

// DPD tensor view: irrep, nirrep, per-irrep lengths [dim][irrep], layout.
template <typename T>
struct dpd_varray_view
{
    T* data;
    unsigned irrep;
    unsigned nirrep;
    std::vector<std::vector<len_type>> len;
    int layout = TBLIS_DPD_PREFIX_ROW_MAJOR;
};

// Indexed tensor view (device memory), like MArray's indexed_varray_view:
// entries with one dense shape, placed at index values, with factors.
template <typename T>
struct indexed_varray_view
{
    std::vector<T*> data;                  // one pointer per entry
    std::vector<len_type> dense_len;       // dense dimensions (first)
    std::vector<stride_type> dense_stride;
    std::vector<len_type> idx_len;         // indexed dimensions (last)
    std::vector<len_type> indices;         // entries x idx_len.size(), row major
    std::vector<T> factor;                 // empty: all 1
};

// Indexed DPD tensor view (device memory), like MArray's
// indexed_dpd_varray_view: the dense part is DPD storage per entry.
template <typename T>
struct indexed_dpd_varray_view
{
    std::vector<T*> data;
    unsigned irrep;
    unsigned nirrep;
    std::vector<std::vector<len_type>> len;  // [dim][irrep], dense dims first
    unsigned dense_ndim;
    std::vector<unsigned> idx_irrep;         // one per indexed dimension
    std::vector<len_type> indices;           // entries x idx_irrep.size(), irrep-local
    std::vector<T> factor;                   // empty: all 1
    int layout = TBLIS_DPD_PREFIX_ROW_MAJOR;
};

namespace detail
{

template <typename T>
struct type_tag;
template <>
struct type_tag<float>
{
    static constexpr type_t value = TYPE_SINGLE;
};
template <>
struct type_tag<double>
{
    static constexpr type_t value = TYPE_DOUBLE;
};

// The descriptor refers to the view's own length and stride arrays, so the
// view must outlive the call.
template <typename T>
tblis_tensor make_tensor_of(T scalar, const void* data, const std::vector<len_type>& len,
                            const std::vector<stride_type>& stride)
{
    tblis_tensor t;
    t.type = type_tag<T>::value;
    t.conj = 0;
    t.scalar.type = t.type;
    if (std::is_same<T, float>::value)
        t.scalar.data.s = (float)scalar;
    else
        t.scalar.data.d = (double)scalar;
    t.data = const_cast<void*>(data);
    t.ndim = (unsigned)len.size();
    t.len = const_cast<len_type*>(len.data());
    t.stride = const_cast<stride_type*>(stride.data());
    return t;
}

template <typename T>
tblis_tensor make_tensor(T scalar, const varray_view<T>& v)
{
    return make_tensor_of<T>(scalar, v.data, v.len, v.stride);
}

template <typename T>
tblis_tensor make_tensor(T scalar, const varray_view<const T>& v)
{
    return make_tensor_of<T>(scalar, v.data, v.len, v.stride);
}

template <typename T>
tblis_dpd_tensor make_dpd(T scalar, const dpd_varray_view<T>& v, std::vector<len_type>& flat)
{
    flat.clear();
    for (auto& row : v.len) flat.insert(flat.end(), row.begin(), row.end());
    tblis_dpd_tensor t;
    if (std::is_same<T, float>::value)
        tblis_init_dpd_tensor_s(&t, (float)scalar, v.irrep, v.nirrep, (unsigned)v.len.size(), flat.data(),
                                (float*)v.data, v.layout);
    else
        tblis_init_dpd_tensor_d(&t, (double)scalar, v.irrep, v.nirrep, (unsigned)v.len.size(), flat.data(),
                                (double*)v.data, v.layout);
    return t;
}

template <typename T>
tblis_indexed_tensor make_indexed(T scalar, const indexed_varray_view<T>& v, std::vector<len_type>& len)
{
    len = v.dense_len;
    len.insert(len.end(), v.idx_len.begin(), v.idx_len.end());
    tblis_indexed_tensor t;
    const len_type n = (len_type)v.data.size();
    if (std::is_same<T, float>::value)
        tblis_init_indexed_tensor_s(&t, (float)scalar, (unsigned)len.size(), (unsigned)v.dense_len.size(),
                                    len.data(), v.dense_stride.data(), n, (float* const*)v.data.data(),
                                    v.indices.data(), v.factor.empty() ? nullptr : (const float*)v.factor.data());
    else
        tblis_init_indexed_tensor_d(&t, (double)scalar, (unsigned)len.size(), (unsigned)v.dense_len.size(),
                                    len.data(), v.dense_stride.data(), n, (double* const*)v.data.data(),
                                    v.indices.data(), v.factor.empty() ? nullptr : (const double*)v.factor.data());
    return t;
}

template <typename T>
tblis_indexed_dpd_tensor make_indexed_dpd(T scalar, const indexed_dpd_varray_view<T>& v, std::vector<len_type>& flat)
{
    flat.clear();
    for (auto& row : v.len) flat.insert(flat.end(), row.begin(), row.end());
    tblis_indexed_dpd_tensor t;
    t.type = type_tag<T>::value;
    t.conj = 0;
    t.scalar.type = t.type;
    if (std::is_same<T, float>::value)
        t.scalar.data.s = (float)scalar;
    else
        t.scalar.data.d = (double)scalar;
    t.ndim = (unsigned)v.len.size();
    t.dense_ndim = v.dense_ndim;
    t.irrep = v.irrep;
    t.nirrep = v.nirrep;
    t.layout = v.layout;
    t.len = flat.data();
    t.idx_irrep = v.idx_irrep.data();
    t.num_indices = (len_type)v.data.size();
    t.data = (void* const*)v.data.data();
    t.indices = v.indices.data();
    t.factor = v.factor.empty() ? nullptr : (const void*)v.factor.data();
    return t;
}

}  // namespace detail

// C <- alpha * A.B + beta * C (src/iface/3t/mult.h:28-37).
template <typename T>
void mult(T alpha, const varray_view<T>& A, const label_type* idx_A, const varray_view<T>& B,
          const label_type* idx_B, T beta, const varray_view<T>& C, const label_type* idx_C,
          const tblis_config* cfg = nullptr)
{
    tblis_tensor a = detail::make_tensor(alpha, A);
    tblis_tensor b = detail::make_tensor(T(1), B);
    tblis_tensor c = detail::make_tensor(beta, C);
    tblis_tensor_mult(nullptr, cfg, &a, idx_A, &b, idx_B, &c, idx_C);
}

// The same with read-only A and B views, the reference's signature
// (varray_view<const T>, src/iface/3t/mult.h:28-37).
template <typename T>
void mult(T alpha, const varray_view<const T>& A, const label_type* idx_A, const varray_view<const T>& B,
          const label_type* idx_B, T beta, const varray_view<T>& C, const label_type* idx_C,
          const tblis_config* cfg = nullptr)
{
    tblis_tensor a = detail::make_tensor(alpha, A);
    tblis_tensor b = detail::make_tensor(T(1), B);
    tblis_tensor c = detail::make_tensor(beta, C);
    tblis_tensor_mult(nullptr, cfg, &a, idx_A, &b, idx_B, &c, idx_C);
}

// C <- alpha * A B + beta * C (src/iface/3m/mult.h:27-38).
template <typename T>
void mult(T alpha, const matrix_view<T>& A, const matrix_view<T>& B, T beta, const matrix_view<T>& C,
          const tblis_config* cfg = nullptr)
{
    tblis_matrix a, b, c;
    auto init = [](tblis_matrix& x, T s, const matrix_view<T>& v) {
        if (std::is_same<T, float>::value)
            tblis_init_matrix_scaled_s(&x, (float)s, v.m, v.n, (float*)v.data, v.rs, v.cs);
        else
            tblis_init_matrix_scaled_d(&x, (double)s, v.m, v.n, (double*)v.data, v.rs, v.cs);
    };
    init(a, alpha, A);
    init(b, T(1), B);
    init(c, beta, C);
    tblis_matrix_mult(nullptr, cfg, &a, &b, &c);
}

// DPD contraction (src/iface/3t/mult.h:53-72).
template <typename T>
void mult(T alpha, const dpd_varray_view<T>& A, const label_type* idx_A, const dpd_varray_view<T>& B,
          const label_type* idx_B, T beta, const dpd_varray_view<T>& C, const label_type* idx_C,
          const tblis_config* cfg = nullptr)
{
    std::vector<len_type> fa, fb, fc;
    tblis_dpd_tensor a = detail::make_dpd(alpha, A, fa);
    tblis_dpd_tensor b = detail::make_dpd(T(1), B, fb);
    tblis_dpd_tensor c = detail::make_dpd(beta, C, fc);
    tblis_dpd_tensor_mult(nullptr, cfg, &a, idx_A, &b, idx_B, &c, idx_C);
}

// Indexed contraction (src/iface/3t/mult.h:74-86).
template <typename T>
void mult(T alpha, const indexed_varray_view<T>& A, const label_type* idx_A, const indexed_varray_view<T>& B,
          const label_type* idx_B, T beta, const indexed_varray_view<T>& C, const label_type* idx_C,
          const tblis_config* cfg = nullptr)
{
    std::vector<len_type> la, lb, lc;
    tblis_indexed_tensor a = detail::make_indexed(alpha, A, la);
    tblis_indexed_tensor b = detail::make_indexed(T(1), B, lb);
    tblis_indexed_tensor c = detail::make_indexed(beta, C, lc);
    tblis_indexed_tensor_mult(nullptr, cfg, &a, idx_A, &b, idx_B, &c, idx_C);
}

// Indexed DPD contraction (src/iface/3t/mult.h:95-114).
template <typename T>
void mult(T alpha, const indexed_dpd_varray_view<T>& A, const label_type* idx_A,
          const indexed_dpd_varray_view<T>& B, const label_type* idx_B, T beta,
          const indexed_dpd_varray_view<T>& C, const label_type* idx_C, const tblis_config* cfg = nullptr)
{
    std::vector<len_type> la, lb, lc;
    tblis_indexed_dpd_tensor a = detail::make_indexed_dpd(alpha, A, la);
    tblis_indexed_dpd_tensor b = detail::make_indexed_dpd(T(1), B, lb);
    tblis_indexed_dpd_tensor c = detail::make_indexed_dpd(beta, C, lc);
    tblis_indexed_dpd_tensor_mult(nullptr, cfg, &a, idx_A, &b, idx_B, &c, idx_C);
}

}  // namespace tblis

#endif
