// C ABI of libtblis_amd (declarations and citations in include/tblis.h).
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "contract.hpp"
#include "dpd.hpp"
#include "kernels.hpp"
#include "indexed.hpp"
#include "level1.hpp"
#include "panel.hpp"
#include "plan.hpp"

using namespace tblis_amd;

namespace
{

const tblis_comm single_comm = {nullptr, 1, 0, 1, 0};

unsigned g_num_threads = 0;

const tblis_config_s g_configs[] = {
    {Algo::Default, "mi355x"},       {Algo::Fp64Mfma, "fp64_mfma"},
    {Algo::Bf16x3, "bf16x3"},        {Algo::Bf16x3Fast, "bf16x3_fast"},
    {Algo::Reference, "reference"},  {Algo::Ozaki, "ozaki"},
};

// Run f; on an argument/runtime error abort with a message (the
// reference's debug-build behaviour) or record it, per the error mode.
template <typename F>
auto guarded(F&& f, decltype(f()) on_error) -> decltype(f())
{
    ThreadState& ts = thread_state();
    ts.last_error.clear();
    try
    {
        return f();
    }
    catch (const std::exception& e)
    {
        ts.last_error = e.what();
    }
    if (!ts.return_on_error)
    {
        std::fprintf(stderr, "tblis: %s\n", ts.last_error.c_str());
        std::abort();
    }
    return on_error;
}

template <typename F>
void guarded_void(F&& f)
{
    guarded(
        [&]() {
            f();
            return 0;
        },
        0);
}

template <typename T>
void init_tensor(tblis_tensor* t, type_t type, T scalar, unsigned ndim, len_type* len, void* data,
                 stride_type* stride)
{
    std::memset(&t->scalar, 0, sizeof(t->scalar));
    t->type = type;
    t->conj = 0;
    std::memcpy(&t->scalar.data, &scalar, sizeof(T));
    t->scalar.type = type;
    t->data = data;
    t->ndim = ndim;
    t->len = len;
    t->stride = stride;
}

template <typename T>
void init_matrix(tblis_matrix* mat, type_t type, T scalar, len_type m, len_type n, void* data,
                 stride_type rs, stride_type cs)
{
    std::memset(&mat->scalar, 0, sizeof(mat->scalar));
    mat->type = type;
    mat->conj = 0;
    std::memcpy(&mat->scalar.data, &scalar, sizeof(T));
    mat->scalar.type = type;
    mat->data = data;
    mat->m = m;
    mat->n = n;
    mat->rs = rs;
    mat->cs = cs;
}

}  // namespace

extern "C"
{

const tblis_comm* const tblis_single = &single_comm;

void tblis_init_scalar_s(tblis_scalar* s, float v)
{
    std::memset(s, 0, sizeof(*s));
    s->data.s = v;
    s->type = TYPE_SINGLE;
}
void tblis_init_scalar_d(tblis_scalar* s, double v)
{
    std::memset(s, 0, sizeof(*s));
    s->data.d = v;
    s->type = TYPE_DOUBLE;
}
void tblis_init_scalar_c(tblis_scalar* s, scomplex v)
{
    std::memset(s, 0, sizeof(*s));
    std::memcpy(&s->data, &v, sizeof(v));
    s->type = TYPE_SCOMPLEX;
}
void tblis_init_scalar_z(tblis_scalar* s, dcomplex v)
{
    std::memset(s, 0, sizeof(*s));
    std::memcpy(&s->data, &v, sizeof(v));
    s->type = TYPE_DCOMPLEX;
}

void tblis_init_tensor_scaled_s(tblis_tensor* t, float scalar, unsigned ndim, len_type* len,
                                float* data, stride_type* stride)
{
    init_tensor(t, TYPE_SINGLE, scalar, ndim, len, data, stride);
}
void tblis_init_tensor_scaled_d(tblis_tensor* t, double scalar, unsigned ndim, len_type* len,
                                double* data, stride_type* stride)
{
    init_tensor(t, TYPE_DOUBLE, scalar, ndim, len, data, stride);
}
void tblis_init_tensor_scaled_c(tblis_tensor* t, scomplex scalar, unsigned ndim, len_type* len,
                                scomplex* data, stride_type* stride)
{
    init_tensor(t, TYPE_SCOMPLEX, scalar, ndim, len, data, stride);
}
void tblis_init_tensor_scaled_z(tblis_tensor* t, dcomplex scalar, unsigned ndim, len_type* len,
                                dcomplex* data, stride_type* stride)
{
    init_tensor(t, TYPE_DCOMPLEX, scalar, ndim, len, data, stride);
}
void tblis_init_tensor_s(tblis_tensor* t, unsigned ndim, len_type* len, float* data,
                         stride_type* stride)
{
    init_tensor(t, TYPE_SINGLE, 1.0f, ndim, len, data, stride);
}
void tblis_init_tensor_d(tblis_tensor* t, unsigned ndim, len_type* len, double* data,
                         stride_type* stride)
{
    init_tensor(t, TYPE_DOUBLE, 1.0, ndim, len, data, stride);
}
void tblis_init_tensor_c(tblis_tensor* t, unsigned ndim, len_type* len, scomplex* data,
                         stride_type* stride)
{
    init_tensor(t, TYPE_SCOMPLEX, scomplex(1.0f, 0.0f), ndim, len, data, stride);
}
void tblis_init_tensor_z(tblis_tensor* t, unsigned ndim, len_type* len, dcomplex* data,
                         stride_type* stride)
{
    init_tensor(t, TYPE_DCOMPLEX, dcomplex(1.0, 0.0), ndim, len, data, stride);
}

void tblis_init_matrix_scaled_s(tblis_matrix* mat, float scalar, len_type m, len_type n, float* data,
                                stride_type rs, stride_type cs)
{
    init_matrix(mat, TYPE_SINGLE, scalar, m, n, data, rs, cs);
}
void tblis_init_matrix_scaled_d(tblis_matrix* mat, double scalar, len_type m, len_type n, double* data,
                                stride_type rs, stride_type cs)
{
    init_matrix(mat, TYPE_DOUBLE, scalar, m, n, data, rs, cs);
}
void tblis_init_matrix_scaled_c(tblis_matrix* mat, scomplex scalar, len_type m, len_type n,
                                scomplex* data, stride_type rs, stride_type cs)
{
    init_matrix(mat, TYPE_SCOMPLEX, scalar, m, n, data, rs, cs);
}
void tblis_init_matrix_scaled_z(tblis_matrix* mat, dcomplex scalar, len_type m, len_type n,
                                dcomplex* data, stride_type rs, stride_type cs)
{
    init_matrix(mat, TYPE_DCOMPLEX, scalar, m, n, data, rs, cs);
}
void tblis_init_matrix_s(tblis_matrix* mat, len_type m, len_type n, float* data, stride_type rs,
                         stride_type cs)
{
    init_matrix(mat, TYPE_SINGLE, 1.0f, m, n, data, rs, cs);
}
void tblis_init_matrix_d(tblis_matrix* mat, len_type m, len_type n, double* data, stride_type rs,
                         stride_type cs)
{
    init_matrix(mat, TYPE_DOUBLE, 1.0, m, n, data, rs, cs);
}
void tblis_init_matrix_c(tblis_matrix* mat, len_type m, len_type n, scomplex* data, stride_type rs,
                         stride_type cs)
{
    init_matrix(mat, TYPE_SCOMPLEX, scomplex(1.0f, 0.0f), m, n, data, rs, cs);
}
void tblis_init_matrix_z(tblis_matrix* mat, len_type m, len_type n, dcomplex* data, stride_type rs,
                         stride_type cs)
{
    init_matrix(mat, TYPE_DCOMPLEX, dcomplex(1.0, 0.0), m, n, data, rs, cs);
}

void tblis_matrix_mult(const tblis_comm* comm, const tblis_config* cfg, const tblis_matrix* A,
                       const tblis_matrix* B, tblis_matrix* C)
{
    (void)comm;
    guarded_void([&] { matrix_mult(cfg, A, B, C); });
}

void tblis_tensor_add(const tblis_comm* comm, const tblis_config* cfg, const tblis_tensor* A,
                      const label_type* idx_A, tblis_tensor* B, const label_type* idx_B)
{
    (void)comm;
    guarded_void([&] { tensor_add(cfg, A, idx_A, B, idx_B); });
}

void tblis_tensor_scale(const tblis_comm* comm, const tblis_config* cfg, tblis_tensor* A,
                        const label_type* idx_A)
{
    (void)comm;
    guarded_void([&] { tensor_scale(cfg, A, idx_A); });
}

void tblis_tensor_set(const tblis_comm* comm, const tblis_config* cfg, const tblis_scalar* alpha,
                      tblis_tensor* A, const label_type* idx_A)
{
    (void)comm;
    guarded_void([&] { tensor_set(cfg, alpha, A, idx_A); });
}

void tblis_tensor_mult(const tblis_comm* comm, const tblis_config* cfg, const tblis_tensor* A,
                       const label_type* idx_A, const tblis_tensor* B, const label_type* idx_B,
                       tblis_tensor* C, const label_type* idx_C)
{
    (void)comm;
    guarded_void([&] { tensor_mult(cfg, A, idx_A, B, idx_B, C, idx_C); });
}

void tblis_hip_tensor_mult_noalloc(const tblis_comm* comm, const tblis_config* cfg, const tblis_tensor* A,
                                   const label_type* idx_A, const tblis_tensor* B, const label_type* idx_B,
                                   tblis_tensor* C, const label_type* idx_C)
{
    (void)comm;
    guarded_void([&] { tensor_mult_noalloc(cfg, A, idx_A, B, idx_B, C, idx_C); });
}

unsigned tblis_get_num_threads(void)
{
    if (g_num_threads) return g_num_threads;
    const char* e = std::getenv("TBLIS_NUM_THREADS");
    if (!e) e = std::getenv("OMP_NUM_THREADS");
    if (e && std::atoi(e) > 0) return (unsigned)std::atoi(e);
    return 1;
}

void tblis_set_num_threads(unsigned n) { g_num_threads = n; }

const tblis_config* tblis_get_config(const char* name)
{
    for (auto& c : g_configs)
        if (name && std::strcmp(c.name, name) == 0) return &c;
    std::fprintf(stderr, "tblis: No configuration named %s!\n", name ? name : "(null)");
    std::abort();
}

void tblis_init_dpd_tensor_d(tblis_dpd_tensor* t, double scalar, unsigned irrep, unsigned nirrep,
                             unsigned ndim, const len_type* len, double* data, int layout)
{
    std::memset(t, 0, sizeof(*t));
    t->type = TYPE_DOUBLE;
    t->scalar.data.d = scalar;
    t->scalar.type = TYPE_DOUBLE;
    t->data = data;
    t->ndim = ndim;
    t->irrep = irrep;
    t->nirrep = nirrep;
    t->layout = layout;
    t->len = len;
}

void tblis_init_dpd_tensor_s(tblis_dpd_tensor* t, float scalar, unsigned irrep, unsigned nirrep,
                             unsigned ndim, const len_type* len, float* data, int layout)
{
    std::memset(t, 0, sizeof(*t));
    t->type = TYPE_SINGLE;
    t->scalar.data.s = scalar;
    t->scalar.type = TYPE_SINGLE;
    t->data = data;
    t->ndim = ndim;
    t->irrep = irrep;
    t->nirrep = nirrep;
    t->layout = layout;
    t->len = len;
}

void tblis_init_dpd_tensor_c(tblis_dpd_tensor* t, scomplex scalar, unsigned irrep, unsigned nirrep,
                             unsigned ndim, const len_type* len, scomplex* data, int layout)
{
    std::memset(t, 0, sizeof(*t));
    t->type = TYPE_SCOMPLEX;
    std::memcpy(&t->scalar.data, &scalar, sizeof(scalar));
    t->scalar.type = TYPE_SCOMPLEX;
    t->data = data;
    t->ndim = ndim;
    t->irrep = irrep;
    t->nirrep = nirrep;
    t->layout = layout;
    t->len = len;
}

void tblis_init_dpd_tensor_z(tblis_dpd_tensor* t, dcomplex scalar, unsigned irrep, unsigned nirrep,
                             unsigned ndim, const len_type* len, dcomplex* data, int layout)
{
    std::memset(t, 0, sizeof(*t));
    t->type = TYPE_DCOMPLEX;
    std::memcpy(&t->scalar.data, &scalar, sizeof(scalar));
    t->scalar.type = TYPE_DCOMPLEX;
    t->data = data;
    t->ndim = ndim;
    t->irrep = irrep;
    t->nirrep = nirrep;
    t->layout = layout;
    t->len = len;
}

stride_type tblis_dpd_tensor_size(const tblis_dpd_tensor* t)
{
    return guarded(
        [&]() -> stride_type {
            if (!t) fail("null DPD descriptor");
            DpdLayout L;
            L.init(t->irrep, t->nirrep, t->ndim, t->len, t->layout);
            return L.total();
        },
        -1);
}

void tblis_dpd_tensor_mult(const tblis_comm* comm, const tblis_config* cfg,
                           const tblis_dpd_tensor* A, const label_type* idx_A,
                           const tblis_dpd_tensor* B, const label_type* idx_B, tblis_dpd_tensor* C,
                           const label_type* idx_C)
{
    (void)comm;
    guarded_void([&] { dpd_tensor_mult(cfg, A, idx_A, B, idx_B, C, idx_C); });
}

static void init_indexed(tblis_indexed_tensor* t, type_t type, unsigned ndim, unsigned dense_ndim,
                         const len_type* len, const stride_type* dense_stride, len_type num_indices,
                         void* const* data, const len_type* indices, const void* factor)
{
    std::memset(t, 0, sizeof(*t));
    t->type = type;
    t->scalar.type = type;
    t->ndim = ndim;
    t->dense_ndim = dense_ndim;
    t->len = len;
    t->dense_stride = dense_stride;
    t->num_indices = num_indices;
    t->data = data;
    t->indices = indices;
    t->factor = factor;
}

void tblis_init_indexed_tensor_d(tblis_indexed_tensor* t, double scalar, unsigned ndim,
                                 unsigned dense_ndim, const len_type* len,
                                 const stride_type* dense_stride, len_type num_indices,
                                 double* const* data, const len_type* indices, const double* factor)
{
    init_indexed(t, TYPE_DOUBLE, ndim, dense_ndim, len, dense_stride, num_indices, (void* const*)data,
                 indices, factor);
    t->scalar.data.d = scalar;
}

void tblis_init_indexed_tensor_s(tblis_indexed_tensor* t, float scalar, unsigned ndim,
                                 unsigned dense_ndim, const len_type* len,
                                 const stride_type* dense_stride, len_type num_indices,
                                 float* const* data, const len_type* indices, const float* factor)
{
    init_indexed(t, TYPE_SINGLE, ndim, dense_ndim, len, dense_stride, num_indices, (void* const*)data,
                 indices, factor);
    t->scalar.data.s = scalar;
}

void tblis_init_indexed_tensor_c(tblis_indexed_tensor* t, scomplex scalar, unsigned ndim,
                                 unsigned dense_ndim, const len_type* len,
                                 const stride_type* dense_stride, len_type num_indices,
                                 scomplex* const* data, const len_type* indices, const scomplex* factor)
{
    init_indexed(t, TYPE_SCOMPLEX, ndim, dense_ndim, len, dense_stride, num_indices, (void* const*)data,
                 indices, factor);
    std::memcpy(&t->scalar.data, &scalar, sizeof(scalar));
}

void tblis_init_indexed_tensor_z(tblis_indexed_tensor* t, dcomplex scalar, unsigned ndim,
                                 unsigned dense_ndim, const len_type* len,
                                 const stride_type* dense_stride, len_type num_indices,
                                 dcomplex* const* data, const len_type* indices, const dcomplex* factor)
{
    init_indexed(t, TYPE_DCOMPLEX, ndim, dense_ndim, len, dense_stride, num_indices, (void* const*)data,
                 indices, factor);
    std::memcpy(&t->scalar.data, &scalar, sizeof(scalar));
}

void tblis_indexed_tensor_mult(const tblis_comm* comm, const tblis_config* cfg,
                               const tblis_indexed_tensor* A, const label_type* idx_A,
                               const tblis_indexed_tensor* B, const label_type* idx_B,
                               tblis_indexed_tensor* C, const label_type* idx_C)
{
    (void)comm;
    guarded_void([&] { indexed_tensor_mult(cfg, A, idx_A, B, idx_B, C, idx_C); });
}

void tblis_indexed_dpd_tensor_mult(const tblis_comm* comm, const tblis_config* cfg,
                                   const tblis_indexed_dpd_tensor* A, const label_type* idx_A,
                                   const tblis_indexed_dpd_tensor* B, const label_type* idx_B,
                                   tblis_indexed_dpd_tensor* C, const label_type* idx_C)
{
    (void)comm;
    guarded_void([&] { indexed_dpd_tensor_mult(cfg, A, idx_A, B, idx_B, C, idx_C); });
}

void tblis_tensor_mult_batch(const tblis_comm* comm, const tblis_config* cfg, len_type count,
                             const tblis_tensor* const* A, const label_type* const* idx_A,
                             const tblis_tensor* const* B, const label_type* const* idx_B,
                             tblis_tensor* const* C, const label_type* const* idx_C)
{
    (void)comm;
    guarded_void([&] { tensor_mult_batch(cfg, count, A, idx_A, B, idx_B, C, idx_C); });
}

void tblis_hip_set_stream(void* s) { thread_state().stream = (hipStream_t)s; }
void* tblis_hip_get_stream(void) { return (void*)thread_state().stream; }
void tblis_hip_set_async(int a) { thread_state().async = a != 0; }
void tblis_hip_set_error_mode(int r) { thread_state().return_on_error = r != 0; }
const char* tblis_hip_get_last_error(void) { return thread_state().last_error.c_str(); }
void tblis_hip_clear_plan_cache(void) { cache_clear(); }
void tblis_hip_set_autotune(int on) { set_autotune(on); }
int tblis_hip_last_tile(void) { return thread_state().last_tile; }
int tblis_hip_last_direct(void) { return thread_state().last_direct; }
int tblis_hip_autotune_runs(void) { return autotune_runs(); }

int64_t tblis_hip_device_alloc_count(void) { return device_alloc_count(); }
const char* tblis_hip_last_algorithm(void) { return thread_state().last_algo.c_str(); }

int64_t tblis_hip_debug_canonical(const tblis_tensor* A, const label_type* idx_A,
                                  const tblis_tensor* B, const label_type* idx_B,
                                  const tblis_tensor* C, const label_type* idx_C, int64_t* out,
                                  int64_t max_out)
{
    return guarded([&] { return debug_canonical(A, idx_A, B, idx_B, C, idx_C, out, max_out); },
                   (int64_t)-1);
}

void tblis_hip_debug_fill_block_scatter(int ndim, const len_type* len, const stride_type* stride,
                                        len_type BS, len_type off, len_type size,
                                        stride_type* scat, stride_type* bs, int pack_3d,
                                        int elem_size)
{
    guarded_void([&] {
        Side s;
        s.len.assign(len, len + ndim);
        s.stride.assign(stride, stride + ndim);
        s.pack_3d = pack_3d != 0;
        fill_scatter(s, BS, off, size, scat, elem_size);
        if (bs) fill_block_stride(BS, size, scat, bs);
    });
}

int64_t tblis_hip_debug_panel_side(int ndim, const len_type* len, const stride_type* stride, len_type G,
                                   len_type r0, stride_type* offs, stride_type* tab, int64_t max_out)
{
    return guarded(
        [&]() -> int64_t {
            if (ndim < 0 || (ndim > 0 && (!len || !stride)) || G < 1 || r0 < 0) fail("invalid side");
            Side s;
            s.len.assign(len, len + ndim);
            s.stride.assign(stride, stride + ndim);
            std::vector<int64_t> o, t;
            debug_side_tables(s, G, r0, o, t);
            if ((int64_t)o.size() > max_out || (int64_t)t.size() > max_out) return -2;
            std::copy(o.begin(), o.end(), offs);
            std::copy(t.begin(), t.end(), tab);
            return (int64_t)o.size();
        },
        (int64_t)-1);
}

void tblis_hip_debug_dpd_block(const tblis_dpd_tensor* t, const unsigned* irreps, len_type* len,
                               stride_type* offset, stride_type* stride)
{
    guarded_void([&] {
        if (!t) fail("null DPD descriptor");
        DpdLayout L;
        L.init(t->irrep, t->nirrep, t->ndim, t->len, t->layout);
        int64_t bl[DPD_MAXDIM], bs[DPD_MAXDIM], off;
        L.block(irreps, bl, &off, bs);
        for (unsigned i = 0; i < t->ndim; i++)
        {
            len[i] = bl[i];
            stride[i] = bs[i];
        }
        *offset = off;
    });
}

int tblis_hip_debug_dpd_patches(const tblis_dpd_tensor* A, const label_type* idx_A,
                                const tblis_dpd_tensor* B, const label_type* idx_B,
                                const tblis_dpd_tensor* C, const label_type* idx_C, int which,
                                unsigned irrep_ab, len_type BS_row, len_type BS_col, int* npatch,
                                len_type* rows, len_type* cols, stride_type* base,
                                stride_type* rscat, stride_type* cscat, int64_t max_patch,
                                int64_t max_scat)
{
    return guarded(
        [&] {
            return debug_dpd_patches(A, idx_A, B, idx_B, C, idx_C, which, irrep_ab, BS_row, BS_col,
                                     npatch, (int64_t*)rows, (int64_t*)cols, (int64_t*)base,
                                     (int64_t*)rscat, (int64_t*)cscat, max_patch, max_scat);
        },
        -1);
}

int tblis_hip_debug_mfma_probe(double* out_device)
{
    return guarded(
        [&] {
            hipError_t e = launch_mfma_probe(out_device, thread_state().stream);
            if (e == hipSuccess) e = hipStreamSynchronize(thread_state().stream);
            if (e != hipSuccess) fail(std::string("mfma probe: ") + hipGetErrorString(e));
            return 0;
        },
        -1);
}

int tblis_hip_debug_mfma_probe_bf16(const float* A, const float* B, float* D)
{
    return guarded(
        [&] {
            hipError_t e = launch_bf16_probe(A, B, D, thread_state().stream);
            if (e == hipSuccess) e = hipStreamSynchronize(thread_state().stream);
            if (e != hipSuccess) fail(std::string("bf16 mfma probe: ") + hipGetErrorString(e));
            return 0;
        },
        -1);
}

}  // extern "C"
