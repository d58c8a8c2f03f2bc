"""Python binding of libtblis_amd (ctypes over the C ABI in include/tblis.h).

Mirrors the reference's C++ front end ``tblis::mult<T>(alpha, A, idx_A, B,
idx_B, beta, C, idx_C)`` (src/iface/3t/mult.h:28-37) for torch tensors on a
HIP device: lengths and strides are taken from the tensors, labels are
strings, and the call goes through ``tblis_tensor_mult``.  ``dpd_mult`` and
``mult_batch`` wrap the DPD and batched extensions.  There is no CPU
fallback: a missing library is an error.
"""
import ctypes as C
import os

__all__ = ["lib", "mult", "mult_noalloc", "device_alloc_count", "matrix_mult", "dpd_mult", "mult_batch", "indexed_mult", "Indexed",
           "indexed_dpd_mult", "IndexedDpd",
           "prepare_indexed_mult", "prepare_mult", "prepare_dpd_mult",
           "prepare_batch", "Tensor", "DpdTensor", "get_config",
           "set_stream", "set_async", "last_error", "last_algorithm", "DPD_LAYOUTS", "TblisError"]

_HERE = os.path.dirname(os.path.abspath(__file__))
# TBLIS_AMD_LIB selects another build of the library (e.g. a tuning variant)
LIB_PATH = os.environ.get("TBLIS_AMD_LIB", os.path.join(_HERE, "libtblis_amd.so"))

TYPE_SINGLE, TYPE_DOUBLE, TYPE_SCOMPLEX, TYPE_DCOMPLEX = 0, 1, 2, 3

DPD_LAYOUTS = {"BALANCED_COLUMN_MAJOR": 0, "BALANCED_ROW_MAJOR": 1, "BLOCKED_COLUMN_MAJOR": 2,
               "BLOCKED_ROW_MAJOR": 3, "PREFIX_COLUMN_MAJOR": 4, "PREFIX_ROW_MAJOR": 5}


class TblisError(RuntimeError):
    pass


class Scalar(C.Structure):
    class _U(C.Union):
        _fields_ = [("s", C.c_float), ("d", C.c_double), ("c", C.c_float * 2), ("z", C.c_double * 2)]
    _fields_ = [("data", _U), ("type", C.c_int)]


class Tensor(C.Structure):
    _fields_ = [("type", C.c_int), ("conj", C.c_int), ("scalar", Scalar), ("data", C.c_void_p),
                ("ndim", C.c_uint), ("len", C.POINTER(C.c_ssize_t)),
                ("stride", C.POINTER(C.c_ssize_t))]


class Matrix(C.Structure):
    _fields_ = [("type", C.c_int), ("conj", C.c_int), ("scalar", Scalar), ("data", C.c_void_p),
                ("m", C.c_ssize_t), ("n", C.c_ssize_t), ("rs", C.c_ssize_t), ("cs", C.c_ssize_t)]


class DpdTensor(C.Structure):
    _fields_ = [("type", C.c_int), ("conj", C.c_int), ("scalar", Scalar), ("data", C.c_void_p),
                ("ndim", C.c_uint), ("irrep", C.c_uint), ("nirrep", C.c_uint), ("layout", C.c_int),
                ("len", C.POINTER(C.c_ssize_t))]


class IndexedTensor(C.Structure):
    _fields_ = [("type", C.c_int), ("conj", C.c_int), ("scalar", Scalar), ("ndim", C.c_uint),
                ("dense_ndim", C.c_uint), ("len", C.POINTER(C.c_ssize_t)),
                ("dense_stride", C.POINTER(C.c_ssize_t)), ("num_indices", C.c_ssize_t),
                ("data", C.POINTER(C.c_void_p)), ("indices", C.POINTER(C.c_ssize_t)),
                ("factor", C.c_void_p)]


class IndexedDpdTensor(C.Structure):
    _fields_ = [("type", C.c_int), ("conj", C.c_int), ("scalar", Scalar), ("ndim", C.c_uint),
                ("dense_ndim", C.c_uint), ("irrep", C.c_uint), ("nirrep", C.c_uint), ("layout", C.c_int),
                ("len", C.POINTER(C.c_ssize_t)), ("idx_irrep", C.POINTER(C.c_uint)),
                ("num_indices", C.c_ssize_t), ("data", C.POINTER(C.c_void_p)),
                ("indices", C.POINTER(C.c_ssize_t)), ("factor", C.c_void_p)]


assert C.sizeof(Scalar) == 24 and C.sizeof(Tensor) == 64 and C.sizeof(Matrix) == 72
assert C.sizeof(IndexedTensor) == 88 and C.sizeof(IndexedDpdTensor) == 104

_LIB = None

# every symbol include/tblis.h declares
EXPORTED = [
    "tblis_single", "tblis_init_scalar_s", "tblis_init_scalar_d", "tblis_init_scalar_c",
    "tblis_init_scalar_z", "tblis_init_tensor_scaled_s", "tblis_init_tensor_scaled_d",
    "tblis_init_tensor_scaled_c", "tblis_init_tensor_scaled_z", "tblis_init_tensor_s",
    "tblis_init_tensor_d", "tblis_init_tensor_c", "tblis_init_tensor_z", "tblis_tensor_mult",
    "tblis_init_matrix_scaled_s", "tblis_init_matrix_scaled_d", "tblis_init_matrix_scaled_c",
    "tblis_init_matrix_scaled_z", "tblis_init_matrix_s", "tblis_init_matrix_d", "tblis_init_matrix_c",
    "tblis_init_matrix_z", "tblis_matrix_mult", "tblis_tensor_add", "tblis_tensor_scale",
    "tblis_tensor_set",
    "tblis_get_num_threads", "tblis_set_num_threads", "tblis_get_config",
    "tblis_init_dpd_tensor_d", "tblis_init_dpd_tensor_s", "tblis_init_dpd_tensor_c",
    "tblis_init_dpd_tensor_z", "tblis_dpd_tensor_size",
    "tblis_dpd_tensor_mult", "tblis_init_indexed_tensor_d", "tblis_init_indexed_tensor_s",
    "tblis_init_indexed_tensor_c", "tblis_init_indexed_tensor_z",
    "tblis_indexed_tensor_mult", "tblis_indexed_dpd_tensor_mult", "tblis_tensor_mult_batch",
    "tblis_hip_set_stream",
    "tblis_hip_get_stream", "tblis_hip_set_async", "tblis_hip_set_error_mode",
    "tblis_hip_get_last_error", "tblis_hip_clear_plan_cache", "tblis_hip_last_algorithm",
    "tblis_hip_set_autotune", "tblis_hip_autotune_runs", "tblis_hip_last_tile", "tblis_hip_last_direct",
    "tblis_hip_tensor_mult_noalloc", "tblis_hip_device_alloc_count",
    "tblis_hip_debug_canonical", "tblis_hip_debug_fill_block_scatter", "tblis_hip_debug_panel_side",
    "tblis_hip_debug_dpd_block",
    "tblis_hip_debug_dpd_patches", "tblis_hip_debug_mfma_probe", "tblis_hip_debug_mfma_probe_bf16",
]


def lib():
    """Load libtblis_amd.so (built in-tree by __graft_entry__.build())."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise TblisError("libtblis_amd.so is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        T, D = C.POINTER(Tensor), C.POINTER(DpdTensor)
        P = C.POINTER
        L.tblis_tensor_mult.argtypes = [C.c_void_p, C.c_void_p, T, C.c_char_p, T, C.c_char_p, T,
                                        C.c_char_p]
        L.tblis_dpd_tensor_mult.argtypes = [C.c_void_p, C.c_void_p, D, C.c_char_p, D, C.c_char_p, D,
                                            C.c_char_p]
        L.tblis_hip_tensor_mult_noalloc.argtypes = L.tblis_tensor_mult.argtypes
        L.tblis_hip_device_alloc_count.restype = C.c_int64
        IX = C.POINTER(IndexedTensor)
        L.tblis_indexed_tensor_mult.argtypes = [C.c_void_p, C.c_void_p, IX, C.c_char_p, IX, C.c_char_p,
                                                IX, C.c_char_p]
        IXD = C.POINTER(IndexedDpdTensor)
        L.tblis_indexed_dpd_tensor_mult.argtypes = [C.c_void_p, C.c_void_p, IXD, C.c_char_p, IXD,
                                                    C.c_char_p, IXD, C.c_char_p]
        L.tblis_tensor_add.argtypes = [C.c_void_p, C.c_void_p, T, C.c_char_p, T, C.c_char_p]
        L.tblis_tensor_scale.argtypes = [C.c_void_p, C.c_void_p, T, C.c_char_p]
        L.tblis_tensor_set.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(Scalar), T, C.c_char_p]
        M = C.POINTER(Matrix)
        L.tblis_matrix_mult.argtypes = [C.c_void_p, C.c_void_p, M, M, M]
        L.tblis_tensor_mult_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_ssize_t, P(T),
                                              P(C.c_char_p), P(T), P(C.c_char_p), P(T),
                                              P(C.c_char_p)]
        L.tblis_hip_debug_panel_side.restype = C.c_int64
        L.tblis_hip_debug_panel_side.argtypes = [C.c_int, P(C.c_ssize_t), P(C.c_ssize_t), C.c_ssize_t,
                                                 C.c_ssize_t, P(C.c_ssize_t), P(C.c_ssize_t), C.c_int64]
        L.tblis_get_config.restype = C.c_void_p
        L.tblis_get_config.argtypes = [C.c_char_p]
        L.tblis_hip_get_last_error.restype = C.c_char_p
        L.tblis_hip_last_algorithm.restype = C.c_char_p
        L.tblis_hip_get_stream.restype = C.c_void_p
        L.tblis_hip_set_stream.argtypes = [C.c_void_p]
        L.tblis_dpd_tensor_size.restype = C.c_ssize_t
        L.tblis_dpd_tensor_size.argtypes = [D]
        L.tblis_hip_debug_canonical.restype = C.c_int64
        L.tblis_hip_debug_canonical.argtypes = [T, C.c_char_p, T, C.c_char_p, T, C.c_char_p,
                                                P(C.c_int64), C.c_int64]
        L.tblis_hip_debug_fill_block_scatter.argtypes = [C.c_int, P(C.c_ssize_t), P(C.c_ssize_t),
                                                         C.c_ssize_t, C.c_ssize_t, C.c_ssize_t,
                                                         P(C.c_ssize_t), P(C.c_ssize_t), C.c_int,
                                                         C.c_int]
        L.tblis_hip_debug_dpd_block.argtypes = [D, P(C.c_uint), P(C.c_ssize_t), P(C.c_ssize_t),
                                                P(C.c_ssize_t)]
        L.tblis_hip_debug_dpd_patches.restype = C.c_int
        L.tblis_hip_debug_dpd_patches.argtypes = [D, C.c_char_p, D, C.c_char_p, D, C.c_char_p,
                                                  C.c_int, C.c_uint, C.c_ssize_t, C.c_ssize_t,
                                                  P(C.c_int), P(C.c_ssize_t), P(C.c_ssize_t),
                                                  P(C.c_ssize_t), P(C.c_ssize_t), P(C.c_ssize_t),
                                                  C.c_int64, C.c_int64]
        L.tblis_hip_debug_mfma_probe.argtypes = [C.c_void_p]
        L.tblis_hip_debug_mfma_probe.restype = C.c_int
        L.tblis_hip_debug_mfma_probe_bf16.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.tblis_hip_debug_mfma_probe_bf16.restype = C.c_int
        _LIB = L
    return _LIB


def get_config(name):
    return lib().tblis_get_config(name.encode())


def set_stream(stream_ptr):
    lib().tblis_hip_set_stream(stream_ptr)


def set_async(flag):
    lib().tblis_hip_set_async(int(bool(flag)))


def set_error_mode(return_on_error):
    lib().tblis_hip_set_error_mode(int(bool(return_on_error)))


def last_error():
    return lib().tblis_hip_get_last_error().decode()


def last_algorithm():
    return lib().tblis_hip_last_algorithm().decode()


def set_autotune(flag):
    """Time the candidate panel tiles when a plan is built (see tblis.h)."""
    lib().tblis_hip_set_autotune(int(bool(flag)))


def autotune_runs():
    return lib().tblis_hip_autotune_runs()


def last_tile():
    return lib().tblis_hip_last_tile()


def last_direct():
    """Operands the last contraction's panel GEMM read straight from the
    tensors instead of packing them: bit 0 A, bit 1 B."""
    return lib().tblis_hip_last_direct()


def _type_code(dtype):
    import torch
    if dtype == torch.float32:
        return TYPE_SINGLE
    if dtype == torch.float64:
        return TYPE_DOUBLE
    if dtype == torch.complex64:
        return TYPE_SCOMPLEX
    if dtype == torch.complex128:
        return TYPE_DCOMPLEX
    raise TblisError("unsupported dtype %s" % dtype)


def _set_scalar(s, value, type_code):
    if type_code == TYPE_SINGLE:
        s.data.s = value
    elif type_code == TYPE_DOUBLE:
        s.data.d = value
    elif type_code == TYPE_SCOMPLEX:
        s.data.c[0], s.data.c[1] = complex(value).real, complex(value).imag
    else:
        s.data.z[0], s.data.z[1] = complex(value).real, complex(value).imag
    s.type = type_code


class _Desc:
    """tblis_tensor over a torch tensor (keeps the length/stride arrays alive)."""

    def __init__(self, t, scalar=1.0):
        self.t = t
        n = t.dim()
        self.len = (C.c_ssize_t * max(n, 1))(*t.shape)
        self.stride = (C.c_ssize_t * max(n, 1))(*t.stride())
        self.d = Tensor()
        self.d.type = _type_code(t.dtype)
        self.d.conj = 0
        _set_scalar(self.d.scalar, scalar, self.d.type)
        self.d.data = C.c_void_p(t.data_ptr())
        self.d.ndim = n
        self.d.len = C.cast(self.len, C.POINTER(C.c_ssize_t))
        self.d.stride = C.cast(self.stride, C.POINTER(C.c_ssize_t))


def _labels(s, ndim, name):
    if len(s) != ndim:
        raise TblisError("%s: %d labels for %d dimensions" % (name, len(s), ndim))
    return s.encode("latin1")


def _check(fn):
    err = last_error()
    if err:
        raise TblisError("%s: %s" % (fn, err))


class _Prepared:
    """A contraction whose descriptors are built once; calling it re-issues
    the C-ABI call (the tensors' storage must stay the same)."""

    def __init__(self, fn, args, keep):
        self._fn, self._args, self._keep = fn, args, keep

    def __call__(self):
        L = lib()
        L.tblis_hip_set_error_mode(1)
        self._fn(*self._args)
        err = last_error()
        if err:
            raise TblisError(err)


def prepare_mult(alpha, A, idx_A, B, idx_B, beta, C_, idx_C, config=None, conj_A=False, conj_B=False):
    """Build the descriptors of mult() once; returns a callable."""
    dA, dB, dC = _Desc(A, alpha), _Desc(B), _Desc(C_, beta)
    dA.d.conj, dB.d.conj = int(bool(conj_A)), int(bool(conj_B))
    cfg = get_config(config) if config else None
    args = (None, cfg, C.byref(dA.d), _labels(idx_A, A.dim(), "A"), C.byref(dB.d),
            _labels(idx_B, B.dim(), "B"), C.byref(dC.d), _labels(idx_C, C_.dim(), "C"))
    fn = lib().tblis_tensor_mult

    def run():
        # the library resets C's scalar to 1 after a call (as the reference
        # does): restore beta for the next call
        _set_scalar(dC.d.scalar, beta, dC.d.type)
        fn(*args)

    return _Prepared(run, (), (dA, dB, dC))


def mult(alpha, A, idx_A, B, idx_B, beta, C_, idx_C, config=None, conj_A=False, conj_B=False):
    """C_ <- alpha * A.B + beta * C_ (tblis::mult).  Tensors: torch float32,
    float64, complex64 or complex128 on the current HIP device, any strides;
    conj_A / conj_B conjugate a complex operand."""
    try:
        prepare_mult(alpha, A, idx_A, B, idx_B, beta, C_, idx_C, config, conj_A, conj_B)()
    except TblisError as e:
        raise TblisError("tblis_tensor_mult: %s" % e) from None
    return C_


def mult_noalloc(alpha, A, idx_A, B, idx_B, beta, C_, idx_C, config=None):
    """mult() through tblis_hip_tensor_mult_noalloc: no device allocation
    (plan tables and workspace in the library's 64 MiB device scratch)."""
    dA, dB, dC = _Desc(A, alpha), _Desc(B), _Desc(C_, beta)
    cfg = get_config(config) if config else None
    L = lib()
    L.tblis_hip_set_error_mode(1)
    L.tblis_hip_tensor_mult_noalloc(None, cfg, C.byref(dA.d), _labels(idx_A, A.dim(), "A"), C.byref(dB.d),
                                    _labels(idx_B, B.dim(), "B"), C.byref(dC.d), _labels(idx_C, C_.dim(), "C"))
    err = last_error()
    if err:
        raise TblisError("tblis_hip_tensor_mult_noalloc: %s" % err)
    return C_


def device_alloc_count():
    """Device allocations the library has made in this process."""
    return lib().tblis_hip_device_alloc_count()


def add(alpha, A, idx_A, beta, B, idx_B):
    """B <- alpha * A + beta * B over labels (tblis::add): labels only in A
    are summed, labels only in B replicated."""
    dA, dB = _Desc(A, alpha), _Desc(B, beta)
    L = lib()
    L.tblis_hip_set_error_mode(1)
    L.tblis_tensor_add(None, None, C.byref(dA.d), _labels(idx_A, A.dim(), "A"), C.byref(dB.d),
                       _labels(idx_B, B.dim(), "B"))
    _check("tblis_tensor_add")
    return B


def scale(alpha, A, idx_A):
    """A <- alpha * A (tblis::scale)."""
    dA = _Desc(A, alpha)
    L = lib()
    L.tblis_hip_set_error_mode(1)
    L.tblis_tensor_scale(None, None, C.byref(dA.d), _labels(idx_A, A.dim(), "A"))
    _check("tblis_tensor_scale")
    return A


def set(alpha, A, idx_A):
    """A <- alpha on every element its labels address (tblis::set)."""
    dA = _Desc(A)
    s = Scalar()
    _set_scalar(s, alpha, dA.d.type)
    L = lib()
    L.tblis_hip_set_error_mode(1)
    L.tblis_tensor_set(None, None, C.byref(s), C.byref(dA.d), _labels(idx_A, A.dim(), "A"))
    _check("tblis_tensor_set")
    return A


def _matrix(t, scalar):
    if t.dim() != 2:
        raise TblisError("matrix_mult: operands must be 2-D")
    d = Matrix()
    d.type = _type_code(t.dtype)
    d.conj = 0
    _set_scalar(d.scalar, scalar, d.type)
    d.data = C.c_void_p(t.data_ptr())
    d.m, d.n = t.shape
    d.rs, d.cs = t.stride()
    return d


def matrix_mult(alpha, A, B, beta, C_, config=None):
    """C_ <- alpha * A @ B + beta * C_ for 2-D torch tensors with any strides
    (tblis::mult over matrix views, src/iface/3m/mult.h:27-38)."""
    a, b, c = _matrix(A, alpha), _matrix(B, 1.0), _matrix(C_, beta)
    L = lib()
    L.tblis_hip_set_error_mode(1)
    L.tblis_matrix_mult(None, get_config(config) if config else None, C.byref(a), C.byref(b),
                        C.byref(c))
    _check("tblis_matrix_mult")
    return C_


class _DpdDesc:
    def __init__(self, t, irrep, nirrep, lens, layout, scalar=1.0):
        self.t = t
        ndim = len(lens)
        flat = [int(x) for row in lens for x in row]
        self.len = (C.c_ssize_t * max(len(flat), 1))(*flat)
        self.d = DpdTensor()
        self.d.type = _type_code(t.dtype)
        self.d.conj = 0
        _set_scalar(self.d.scalar, scalar, self.d.type)
        self.d.data = C.c_void_p(t.data_ptr())
        self.d.ndim = ndim
        self.d.irrep = irrep
        self.d.nirrep = nirrep
        self.d.layout = DPD_LAYOUTS[layout] if isinstance(layout, str) else int(layout)
        self.d.len = C.cast(self.len, C.POINTER(C.c_ssize_t))


def dpd_size(irrep, nirrep, lens, layout="PREFIX_ROW_MAJOR"):
    """Stored elements of a DPD tensor (lens: ndim x nirrep)."""
    import torch
    d = _DpdDesc(torch.empty(0, dtype=torch.float64), irrep, nirrep, lens, layout)
    return lib().tblis_dpd_tensor_size(C.byref(d.d))


def prepare_dpd_mult(alpha, A, irrep_A, idx_A, B, irrep_B, idx_B, beta, C_, irrep_C, idx_C, nirrep,
                     lens_A, lens_B, lens_C, layout="PREFIX_ROW_MAJOR", config=None, conj_A=False,
                     conj_B=False):
    """Build the descriptors of dpd_mult() once; returns a callable."""
    dA = _DpdDesc(A, irrep_A, nirrep, lens_A, layout, alpha)
    dB = _DpdDesc(B, irrep_B, nirrep, lens_B, layout)
    dC = _DpdDesc(C_, irrep_C, nirrep, lens_C, layout, beta)
    dA.d.conj, dB.d.conj = int(bool(conj_A)), int(bool(conj_B))
    cfg = get_config(config) if config else None
    args = (None, cfg, C.byref(dA.d), idx_A.encode("latin1"), C.byref(dB.d),
            idx_B.encode("latin1"), C.byref(dC.d), idx_C.encode("latin1"))
    fn = lib().tblis_dpd_tensor_mult

    def run():
        _set_scalar(dC.d.scalar, beta, dC.d.type)
        fn(*args)

    return _Prepared(run, (), (dA, dB, dC))


def dpd_mult(alpha, A, irrep_A, idx_A, B, irrep_B, idx_B, beta, C_, irrep_C, idx_C, nirrep,
             lens_A, lens_B, lens_C, layout="PREFIX_ROW_MAJOR", config=None, conj_A=False, conj_B=False):
    """DPD contraction over 1-D storage tensors A, B, C_ (tblis::mult over
    dpd_varray_view, src/iface/3t/mult.h:53-72); real or complex dtypes,
    conj_A / conj_B conjugate a complex operand."""
    try:
        prepare_dpd_mult(alpha, A, irrep_A, idx_A, B, irrep_B, idx_B, beta, C_, irrep_C, idx_C,
                         nirrep, lens_A, lens_B, lens_C, layout, config, conj_A, conj_B)()
    except TblisError as e:
        raise TblisError("tblis_dpd_tensor_mult: %s" % e) from None
    return C_


class Indexed:
    """An indexed tensor (MArray indexed_varray): entries are torch tensors of
    one dense shape and strides on the device; entry i sits at the index
    values indices[i] (one per indexed dimension, of lengths idx_len) and is
    multiplied by factors[i] (default 1)."""

    def __init__(self, entries, indices, idx_len, factors=None, dtype=None):
        self.entries = list(entries)
        self.indices = [tuple(int(v) for v in ix) for ix in indices]
        self.idx_len = [int(x) for x in idx_len]
        self.factors = None if factors is None else [complex(f) if isinstance(f, complex) else float(f)
                                                    for f in factors]
        if len(self.indices) != len(self.entries):
            raise TblisError("Indexed: one index tuple per entry")
        if self.entries:
            e0 = self.entries[0]
            for e in self.entries:
                if tuple(e.shape) != tuple(e0.shape) or tuple(e.stride()) != tuple(e0.stride()):
                    raise TblisError("Indexed: every entry must have the same shape and strides")
                if e.dtype != e0.dtype:
                    raise TblisError("Indexed: every entry must have the same dtype")
            self.dtype = e0.dtype
            self.dense_len = list(e0.shape)
            self.dense_stride = list(e0.stride())
        else:
            if dtype is None:
                raise TblisError("Indexed: dtype is needed when there are no entries")
            self.dtype = dtype
            self.dense_len, self.dense_stride = [], []

    def ndim(self):
        return len(self.dense_len) + len(self.idx_len)


class _IxDesc:
    """tblis_indexed_tensor over an Indexed (keeps the host arrays alive)."""

    def __init__(self, X, scalar=1.0):
        n, nd, nix = len(X.entries), X.ndim(), len(X.idx_len)
        self.len = (C.c_ssize_t * max(nd, 1))(*(X.dense_len + X.idx_len))
        self.stride = (C.c_ssize_t * max(len(X.dense_len), 1))(*X.dense_stride)
        self.data = (C.c_void_p * max(n, 1))(*[e.data_ptr() for e in X.entries])
        flat = [v for ix in X.indices for v in ix]
        if any(len(ix) != nix for ix in X.indices):
            raise TblisError("Indexed: each index tuple needs one value per indexed dimension")
        self.idx = (C.c_ssize_t * max(len(flat), 1))(*flat)
        self.d = IndexedTensor()
        self.d.type = _type_code(X.dtype)
        self.d.conj = 0
        _set_scalar(self.d.scalar, scalar, self.d.type)
        self.d.ndim = nd
        self.d.dense_ndim = len(X.dense_len)
        self.d.len = C.cast(self.len, C.POINTER(C.c_ssize_t))
        self.d.dense_stride = C.cast(self.stride, C.POINTER(C.c_ssize_t))
        self.d.num_indices = n
        self.d.data = C.cast(self.data, C.POINTER(C.c_void_p))
        self.d.indices = C.cast(self.idx, C.POINTER(C.c_ssize_t))
        self.fac = None
        if X.factors is not None:
            ft = C.c_double if self.d.type in (TYPE_DOUBLE, TYPE_DCOMPLEX) else C.c_float
            if self.d.type in (TYPE_SCOMPLEX, TYPE_DCOMPLEX):
                # interleaved (re, im) per entry
                vals = [v for f in X.factors for v in (complex(f).real, complex(f).imag)]
                self.fac = (ft * max(2 * n, 2))(*vals)
            else:
                self.fac = (ft * max(n, 1))(*[float(f) for f in X.factors])
            self.d.factor = C.cast(self.fac, C.c_void_p)


def prepare_indexed_mult(alpha, A, idx_A, B, idx_B, beta, C_, idx_C, config=None, conj_A=False, conj_B=False):
    """Build the descriptors of indexed_mult() once; returns a callable."""
    dA, dB, dC = _IxDesc(A, alpha), _IxDesc(B), _IxDesc(C_, beta)
    dA.d.conj, dB.d.conj = int(bool(conj_A)), int(bool(conj_B))
    cfg = get_config(config) if config else None
    args = (None, cfg, C.byref(dA.d), _labels(idx_A, A.ndim(), "A"), C.byref(dB.d),
            _labels(idx_B, B.ndim(), "B"), C.byref(dC.d), _labels(idx_C, C_.ndim(), "C"))
    fn = lib().tblis_indexed_tensor_mult

    def run():
        _set_scalar(dC.d.scalar, beta, dC.d.type)
        fn(*args)

    return _Prepared(run, (), (dA, dB, dC))


def indexed_mult(alpha, A, idx_A, B, idx_B, beta, C_, idx_C, config=None, conj_A=False, conj_B=False):
    """Indexed contraction (tblis::mult over indexed_varray_view,
    src/iface/3t/mult.h:74-86) over Indexed operands (real or complex;
    conj_A / conj_B conjugate a complex operand's entries); C_'s entries are
    updated in place."""
    try:
        prepare_indexed_mult(alpha, A, idx_A, B, idx_B, beta, C_, idx_C, config, conj_A, conj_B)()
    except TblisError as e:
        raise TblisError("tblis_indexed_tensor_mult: %s" % e) from None
    return C_


class IndexedDpd:
    """An indexed DPD tensor (MArray indexed_dpd_varray): entries are 1-D
    torch tensors holding the DPD storage of the dense part (the first
    dense_ndim dimensions, irrep irrep ^ idx_irrep[0] ^ ...); lens is
    ndim x nirrep; indexed dimension j has irrep idx_irrep[j] and entry i
    sits at the irrep-local values indices[i]."""

    def __init__(self, entries, indices, idx_irrep, irrep, nirrep, lens, dense_ndim, factors=None,
                 layout="PREFIX_ROW_MAJOR", dtype=None):
        self.entries = list(entries)
        self.indices = [tuple(int(v) for v in ix) for ix in indices]
        self.idx_irrep = [int(r) for r in idx_irrep]
        self.irrep, self.nirrep, self.dense_ndim = int(irrep), int(nirrep), int(dense_ndim)
        self.lens = [[int(x) for x in row] for row in lens]
        self.factors = None if factors is None else [complex(f) if isinstance(f, complex) else float(f)
                                                    for f in factors]
        self.layout = DPD_LAYOUTS[layout] if isinstance(layout, str) else int(layout)
        if len(self.indices) != len(self.entries):
            raise TblisError("IndexedDpd: one index tuple per entry")
        self.dtype = self.entries[0].dtype if self.entries else dtype
        if self.dtype is None:
            raise TblisError("IndexedDpd: dtype is needed when there are no entries")

    def ndim(self):
        return len(self.lens)


class _IxDpdDesc:
    def __init__(self, X, scalar=1.0):
        n, nd = len(X.entries), X.ndim()
        flat = [v for row in X.lens for v in row]
        self.len = (C.c_ssize_t * max(len(flat), 1))(*flat)
        self.irr = (C.c_uint * max(len(X.idx_irrep), 1))(*X.idx_irrep)
        self.data = (C.c_void_p * max(n, 1))(*[e.data_ptr() for e in X.entries])
        ix = [v for t in X.indices for v in t]
        self.idx = (C.c_ssize_t * max(len(ix), 1))(*ix)
        self.d = IndexedDpdTensor()
        self.d.type = _type_code(X.dtype)
        _set_scalar(self.d.scalar, scalar, self.d.type)
        self.d.ndim, self.d.dense_ndim = nd, X.dense_ndim
        self.d.irrep, self.d.nirrep, self.d.layout = X.irrep, X.nirrep, X.layout
        self.d.len = C.cast(self.len, C.POINTER(C.c_ssize_t))
        self.d.idx_irrep = C.cast(self.irr, C.POINTER(C.c_uint))
        self.d.num_indices = n
        self.d.data = C.cast(self.data, C.POINTER(C.c_void_p))
        self.d.indices = C.cast(self.idx, C.POINTER(C.c_ssize_t))
        self.fac = None
        if X.factors is not None:
            ft = C.c_double if self.d.type in (TYPE_DOUBLE, TYPE_DCOMPLEX) else C.c_float
            if self.d.type in (TYPE_SCOMPLEX, TYPE_DCOMPLEX):
                vals = [v for f in X.factors for v in (complex(f).real, complex(f).imag)]
                self.fac = (ft * max(2 * n, 2))(*vals)
            else:
                self.fac = (ft * max(n, 1))(*[float(f) for f in X.factors])
            self.d.factor = C.cast(self.fac, C.c_void_p)


def indexed_dpd_mult(alpha, A, idx_A, B, idx_B, beta, C_, idx_C, config=None, conj_A=False, conj_B=False):
    """Indexed DPD contraction (tblis::mult over indexed_dpd_varray_view,
    src/iface/3t/mult.h:95-114) over IndexedDpd operands, real or complex
    (conj_A / conj_B conjugate a complex operand)."""
    dA, dB, dC = _IxDpdDesc(A, alpha), _IxDpdDesc(B), _IxDpdDesc(C_, beta)
    dA.d.conj, dB.d.conj = int(bool(conj_A)), int(bool(conj_B))
    cfg = get_config(config) if config else None
    L = lib()
    L.tblis_hip_set_error_mode(1)
    L.tblis_indexed_dpd_tensor_mult(None, cfg, C.byref(dA.d), _labels(idx_A, A.ndim(), "A"), C.byref(dB.d),
                                    _labels(idx_B, B.ndim(), "B"), C.byref(dC.d), _labels(idx_C, C_.ndim(), "C"))
    _check("tblis_indexed_dpd_tensor_mult")
    return C_


def prepare_batch(alpha, As, idx_As, Bs, idx_Bs, beta, Cs, idx_Cs, config=None, conj_A=False, conj_B=False):
    """Build the descriptor arrays of mult_batch() once; returns a callable."""
    n = len(As)
    dA = [_Desc(a, alpha) for a in As]
    dB = [_Desc(b) for b in Bs]
    dC = [_Desc(c, beta) for c in Cs]
    for d in dA:
        d.d.conj = int(bool(conj_A))
    for d in dB:
        d.d.conj = int(bool(conj_B))
    T = C.POINTER(Tensor)
    pa = (T * n)(*[C.pointer(d.d) for d in dA])
    pb = (T * n)(*[C.pointer(d.d) for d in dB])
    pc = (T * n)(*[C.pointer(d.d) for d in dC])
    la = (C.c_char_p * n)(*[_labels(s, a.dim(), "A") for s, a in zip(idx_As, As)])
    lb = (C.c_char_p * n)(*[_labels(s, b.dim(), "B") for s, b in zip(idx_Bs, Bs)])
    lc = (C.c_char_p * n)(*[_labels(s, c.dim(), "C") for s, c in zip(idx_Cs, Cs)])
    cfg = get_config(config) if config else None
    args = (None, cfg, n, pa, la, pb, lb, pc, lc)

    def run():
        # the library resets each C scalar to 1 after a call: restore beta
        for d in dC:
            _set_scalar(d.d.scalar, beta, d.d.type)
        lib().tblis_tensor_mult_batch(*args)

    return _Prepared(run, (), (dA, dB, dC, pa, pb, pc, la, lb, lc))


def mult_batch(alpha, As, idx_As, Bs, idx_Bs, beta, Cs, idx_Cs, config=None, conj_A=False, conj_B=False):
    """Independent contractions C_i <- alpha*A_i.B_i + beta*C_i, one GEMM
    launch per kernel variant; real or complex dtypes (conj_A / conj_B
    conjugate a complex operand of every instance)."""
    try:
        prepare_batch(alpha, As, idx_As, Bs, idx_Bs, beta, Cs, idx_Cs, config, conj_A, conj_B)()
    except TblisError as e:
        raise TblisError("tblis_tensor_mult_batch: %s" % e) from None
    return Cs
