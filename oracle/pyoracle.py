"""ctypes bindings for the oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: the checker the HIP library is compared against.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

MAXDIM = 16
I64 = C.c_int64
P64 = C.POINTER(C.c_int64)
PD = C.POINTER(C.c_double)
PF = C.POINTER(C.c_float)


class Groups(C.Structure):
    _fields_ = [("valid", C.c_int), ("n_ab", C.c_int), ("n_ac", C.c_int), ("n_bc", C.c_int),
                ("n_abc", C.c_int)] + \
        [(n, C.c_byte * MAXDIM) for n in ("idx_ab", "idx_ac", "idx_bc", "idx_abc")] + \
        [(n, I64 * MAXDIM) for n in ("len_ab", "len_ac", "len_bc", "len_abc", "sA_ab", "sB_ab",
                                     "sA_ac", "sC_ac", "sB_bc", "sC_bc", "sA_abc", "sB_abc",
                                     "sC_abc")]

    def as_dict(self):
        d = {"valid": self.valid}
        for g in ("ab", "ac", "bc", "abc"):
            n = getattr(self, "n_" + g)
            d["idx_" + g] = bytes(x & 0xFF for x in getattr(self, "idx_" + g)[:n]).decode("latin1")
            d["len_" + g] = list(getattr(self, "len_" + g))[:n]
        for f in ("sA_ab", "sB_ab", "sA_ac", "sC_ac", "sB_bc", "sC_bc", "sA_abc", "sB_abc",
                  "sC_abc"):
            g = f.split("_")[1]
            d[f] = list(getattr(self, f))[:getattr(self, "n_" + g)]
        return d


class TMat(C.Structure):
    _fields_ = [("nm", C.c_int), ("nn", C.c_int), ("nk", C.c_int), ("nl", C.c_int)] + \
        [(n, I64 * MAXDIM) for n in ("len_m", "len_n", "len_k", "len_l", "sA_m", "sA_k", "sB_k",
                                     "sB_n", "sC_m", "sC_n", "sA_l", "sB_l", "sC_l")] + \
        [("pack_m_3d", C.c_int), ("pack_n_3d", C.c_int), ("pack_k_3d", C.c_int)]

    def as_dict(self):
        d = {}
        for k, n in (("m", self.nm), ("n", self.nn), ("k", self.nk), ("l", self.nl)):
            d["len_" + k] = list(getattr(self, "len_" + k))[:n]
        for f, n in (("sA_m", self.nm), ("sA_k", self.nk), ("sB_k", self.nk), ("sB_n", self.nn),
                     ("sC_m", self.nm), ("sC_n", self.nn), ("sA_l", self.nl), ("sB_l", self.nl),
                     ("sC_l", self.nl)):
            d[f] = list(getattr(self, f))[:n]
        d["pack_3d"] = (self.pack_m_3d, self.pack_n_3d, self.pack_k_3d)
        return d


class Dpd(C.Structure):
    _fields_ = [("ndim", C.c_int), ("nirrep", C.c_int), ("irrep", C.c_int), ("row_major", C.c_int),
                ("len", (I64 * 8) * MAXDIM), ("ext", (I64 * 8) * MAXDIM),
                ("leaf", C.c_int * MAXDIM), ("parent", C.c_int * (2 * MAXDIM)),
                ("perm", C.c_int * MAXDIM), ("depth", C.c_int * MAXDIM),
                ("size", (I64 * 8) * (2 * MAXDIM))]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        L.orc_canonicalize.restype = C.c_int
        L.orc_diagonal.restype = C.c_int
        L.orc_fold.restype = C.c_int
        L.orc_unit_dim.restype = C.c_int
        L.orc_free_idx.restype = C.c_char
        L.orc_gcd.restype = I64
        L.orc_gcd.argtypes = [I64, I64]
        L.orc_lcm.restype = I64
        L.orc_lcm.argtypes = [I64, I64]
        L.orc_dpd_size.restype = I64
        L.orc_dpd_patch_scatter.restype = C.c_int
        L.orc_blis_contract_d.restype = C.c_double
        _LIB = L
    return _LIB


def _i64(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.int64))
    return a, a.ctypes.data_as(P64)


def _lab(s):
    return C.c_char_p(s.encode("latin1") if isinstance(s, str) else bytes(s))


def canonicalize(lenA, strA, idxA, lenB, strB, idxB, lenC, strC, idxC):
    g = Groups()
    args = []
    keep = []
    for l, s, i in ((lenA, strA, idxA), (lenB, strB, idxB), (lenC, strC, idxC)):
        la, lp = _i64(l)
        sa, sp = _i64(s)
        keep += [la, sa]
        args += [len(i), lp, sp, _lab(i)]
    ok = lib().orc_canonicalize(*args, C.byref(g))
    return g if ok else None


def blis_reorder(g):
    t = TMat()
    lib().orc_blis_reorder(C.byref(g), C.byref(t))
    return t


def diagonal(lens, strides, idx):
    n = len(idx)
    lo = np.zeros(n, np.int64)
    so = np.zeros(n, np.int64)
    io = C.create_string_buffer(n + 1)
    la, lp = _i64(lens)
    sa, sp = _i64(strides)
    r = lib().orc_diagonal(n, lp, sp, _lab(idx), lo.ctypes.data_as(P64), so.ctypes.data_as(P64), io)
    if r < 0:
        return None
    return list(lo[:r]), list(so[:r]), io.raw[:r].decode("latin1")


def fold(lens, idx, strides_list):
    n = len(idx)
    la = np.ascontiguousarray(np.asarray(lens, np.int64))
    ib = C.create_string_buffer(idx.encode("latin1"), n + 1)
    sarrs = [np.ascontiguousarray(np.asarray(s, np.int64)) for s in strides_list]
    ptrs = (P64 * len(sarrs))(*[s.ctypes.data_as(P64) for s in sarrs])
    r = lib().orc_fold(n, la.ctypes.data_as(P64), ib, len(sarrs), ptrs)
    return list(la[:r]), ib.raw[:r].decode("latin1"), [list(s[:r]) for s in sarrs]


def sort_by_stride(strides_list):
    n = len(strides_list[0])
    sarrs = [np.ascontiguousarray(np.asarray(s, np.int64)) for s in strides_list]
    ptrs = (P64 * len(sarrs))(*[s.ctypes.data_as(P64) for s in sarrs])
    perm = (C.c_int * max(n, 1))()
    lib().orc_sort_by_stride(n, len(sarrs), ptrs, perm)
    return list(perm)[:n]


def free_idx(labels):
    return lib().orc_free_idx(len(labels), _lab(labels))


def fill_scatter(lens, strides, BS, off, size, pack_3d=False, elem_size=8):
    out = np.zeros(max(size, 1), np.int64)
    la, lp = _i64(lens)
    sa, sp = _i64(strides)
    lib().orc_fill_scatter(len(lens), lp, sp, I64(BS), I64(off), I64(size),
                           out.ctypes.data_as(P64), int(bool(pack_3d)), elem_size)
    return out[:size]


def fill_block_stride(BS, scat):
    scat = np.ascontiguousarray(np.asarray(scat, np.int64))
    bs = np.zeros(max(len(scat), 1), np.int64)
    lib().orc_fill_block_stride(I64(BS), I64(len(scat)), scat.ctypes.data_as(P64),
                                bs.ctypes.data_as(P64))
    return bs[:len(scat)]


def gcd(a, b):
    return lib().orc_gcd(a, b)


def lcm(a, b):
    return lib().orc_lcm(a, b)


def _tensor_args(lens, strides, idx):
    la, lp = _i64(lens)
    sa, sp = _i64(strides)
    return (la, sa), [len(idx), lp, sp, _lab(idx)]


def tensor_mult(alpha, A, lenA, strA, idxA, B, lenB, strB, idxB, beta, C_, lenC, strC, idxC,
                acc=1):
    """C_ (numpy, modified in place for float64) <- alpha*A.B + beta*C_.

    For float32 inputs returns a float64 array (the exact-ish result)."""
    kA, aA = _tensor_args(lenA, strA, idxA)
    kB, aB = _tensor_args(lenB, strB, idxB)
    kC, aC = _tensor_args(lenC, strC, idxC)
    if A.dtype == np.float64:
        lib().orc_tensor_mult_d(acc, C.c_double(alpha), *aA, A.ctypes.data_as(PD), *aB,
                                B.ctypes.data_as(PD), C.c_double(beta), *aC,
                                C_.ctypes.data_as(PD))
        return C_
    out = C_.astype(np.float64)
    lib().orc_tensor_mult_s(C.c_double(alpha), *aA, A.ctypes.data_as(PF), *aB,
                            B.ctypes.data_as(PF), C.c_double(beta), *aC,
                            out.ctypes.data_as(PD), C_.ctypes.data_as(PF))
    return out


def mult_entries(g, alpha, A, B, beta, C0, ent_l, ent_m, ent_n):
    """Extended-precision values of selected entries (l, m, n flat logical
    indices of the canonical form g)."""
    el, pl = _i64(ent_l)
    em, pm = _i64(ent_m)
    en, pn = _i64(ent_n)
    out = np.zeros(len(el), np.float64)
    is_float = int(A.dtype == np.float32)
    lib().orc_mult_entries_d(C.byref(g), C.c_double(alpha), C.c_void_p(A.ctypes.data),
                             C.c_void_p(B.ctypes.data), C.c_double(beta),
                             C.c_void_p(C0.ctypes.data), is_float, I64(len(el)), pl, pm, pn,
                             out.ctypes.data_as(PD))
    return out


def blis_contract(nthreads, alpha, A, lenA, strA, idxA, B, lenB, strB, idxB, beta, C_, lenC, strC,
                  idxC, n_limit=0):
    kA, aA = _tensor_args(lenA, strA, idxA)
    kB, aB = _tensor_args(lenB, strB, idxB)
    kC, aC = _tensor_args(lenC, strC, idxC)
    return lib().orc_blis_contract_d(nthreads, C.c_double(alpha), *aA, A.ctypes.data_as(PD), *aB,
                                     B.ctypes.data_as(PD), C.c_double(beta), *aC,
                                     C_.ctypes.data_as(PD), I64(n_limit))


# ---------------------------------------------------------------- DPD
def dpd_init(irrep, nirrep, lens, layout):
    """lens: ndim x nirrep nested list.  layout codes: 0 BALANCED_COL,
    1 BALANCED_ROW, 2 BLOCKED_COL, 3 BLOCKED_ROW, 4 PREFIX_COL, 5 PREFIX_ROW."""
    d = Dpd()
    la = np.ascontiguousarray(np.asarray(lens, np.int64).reshape(-1))
    lib().orc_dpd_init(C.byref(d), irrep, nirrep, len(lens), la.ctypes.data_as(P64), layout)
    return d


def dpd_size(d):
    return lib().orc_dpd_size(C.byref(d))


def dpd_get_block(d, irreps):
    n = d.ndim
    ir = (C.c_int * n)(*irreps)
    ln = np.zeros(n, np.int64)
    st = np.zeros(n, np.int64)
    off = I64()
    lib().orc_dpd_get_block(C.byref(d), ir, ln.ctypes.data_as(P64), C.byref(off),
                            st.ctypes.data_as(P64))
    return list(ln), off.value, list(st)


def dpd_mult(alpha, dA, idxA, A, dB, idxB, B, beta, dC, idxC, Cd, mode=1):
    """mode 0: double block-wise, 1: extended block-wise, 2: full dense
    expansion.  float64 arrays: Cd updated in place.  float32: returns a
    float64 array."""
    if A.dtype == np.float64:
        lib().orc_dpd_mult_d(mode, C.c_double(alpha), C.byref(dA), _lab(idxA), A.ctypes.data_as(PD),
                             C.byref(dB), _lab(idxB), B.ctypes.data_as(PD), C.c_double(beta),
                             C.byref(dC), _lab(idxC), Cd.ctypes.data_as(PD))
        return Cd
    out = Cd.astype(np.float64)
    lib().orc_dpd_mult_s(C.c_double(alpha), C.byref(dA), _lab(idxA), A.ctypes.data_as(PF),
                         C.byref(dB), _lab(idxB), B.ctypes.data_as(PF), C.c_double(beta),
                         C.byref(dC), _lab(idxC), out.ctypes.data_as(PD), Cd.ctypes.data_as(PF))
    return out


def complex_dpd_mult(alpha, dA, idxA, A, dB, idxB, B, beta, dC, idxC, C_, conj_A=False, conj_B=False):
    """Complex DPD contraction (the reference instantiates DPD mult for its
    complex types, src/iface/3t/mult.cxx:292-297) on complex128 storage
    arrays: four real block-wise contractions (dpd_mult mode 1) of the
    real and imaginary parts, combined by complex arithmetic,
    P = op(A) op(B), returns alpha P + beta C (op conjugates when conj_X)."""
    sa = -1.0 if conj_A else 1.0
    sb = -1.0 if conj_B else 1.0
    A = np.asarray(A, np.complex128)
    B = np.asarray(B, np.complex128)
    nc = int(np.asarray(C_).size)
    parts = lambda x: (np.ascontiguousarray(x.real), np.ascontiguousarray(x.imag))
    (ar, ai), (br, bi) = parts(A), parts(B)

    def prod(x, y):
        out = np.zeros(nc)
        dpd_mult(1.0, dA, idxA, x, dB, idxB, y, 0.0, dC, idxC, out, mode=1)
        return out

    P = (prod(ar, br) - sa * sb * prod(ai, bi)) + 1j * (sb * prod(ar, bi) + sa * prod(ai, br))
    # C's storage is its dpd_size elements; a longer array keeps the rest
    out = np.array(C_, np.complex128)
    n = dpd_size(dC)
    out[:n] = alpha * P[:n] + beta * out[:n]
    return out


def dpd_patch_scatter(dA, idxA, dB, idxB, dC, idxC, which, irrep_ab, BS_row, BS_col,
                      max_scat=1 << 22):
    npatch = (C.c_int * 2)()
    maxp = 1 << 16
    pr = np.zeros(maxp, np.int64)
    pc = np.zeros(maxp, np.int64)
    pb = np.zeros(maxp, np.int64)
    rs = np.zeros(max_scat, np.int64)
    cs = np.zeros(max_scat, np.int64)
    n = lib().orc_dpd_patch_scatter(C.byref(dA), _lab(idxA), C.byref(dB), _lab(idxB), C.byref(dC),
                                    _lab(idxC), which, irrep_ab, I64(BS_row), I64(BS_col), npatch,
                                    pr.ctypes.data_as(P64), pc.ctypes.data_as(P64),
                                    pb.ctypes.data_as(P64), rs.ctypes.data_as(P64),
                                    cs.ctypes.data_as(P64), I64(max_scat))
    if n < 0:
        raise ValueError("dpd_patch_scatter failed: %d" % n)
    nr = int(pr[:n].sum())
    nc = int(pc[:n].sum())
    return {"npatch": (npatch[0], npatch[1]), "rows": pr[:n].copy(), "cols": pc[:n].copy(),
            "base": pb[:n].copy(), "rscat": rs[:nr].copy(), "cscat": cs[:nc].copy()}


def level1(mode, alpha, A, lenA, strA, idxA, beta, B, lenB, strB, idxB):
    """Level-1 add (mode 0), scale (1), set (2) on host arrays; returns the
    new B storage (float64)."""
    L = lib()
    la, pla = _i64(list(lenA) or [0])
    sa, psa = _i64(list(strA) or [0])
    lb, plb = _i64(lenB)
    sb, psb = _i64(strB)
    if B.dtype == np.float64:
        out = B.copy()
        Ap = A.ctypes.data_as(PD) if A is not None else None
        L.orc_level1_d(mode, C.c_double(alpha), len(lenA), pla, psa, _lab(idxA), Ap,
                       C.c_double(beta), len(lenB), plb, psb, _lab(idxB), out.ctypes.data_as(PD))
        return out
    out = B.astype(np.float64)
    Ap = A.ctypes.data_as(PF) if A is not None else None
    L.orc_level1_s(mode, C.c_double(alpha), len(lenA), pla, psa, _lab(idxA), Ap, C.c_double(beta),
                   len(lenB), plb, psb, _lab(idxB), B.ctypes.data_as(PF), out.ctypes.data_as(PD))
    return out


# ---------------------------------------------------------------- complex

def complex_tensor_mult(alpha, A, lenA, strA, idxA, B, lenB, strB, idxB, beta, C_, lenC, strC, idxC,
                        conj_A=False, conj_B=False):
    """Complex contraction on interleaved host storage (complex128 arrays,
    strides in complex elements) from four real contractions of the oracle
    (oracle.c) over the real and imaginary views (strides doubled, offset 0
    or 1), combined by complex arithmetic: P = op(A) op(B),
    C <- alpha P + beta C (the reference's dense call for its complex types,
    src/iface/3t/mult.cxx:24-165; op conjugates when conj_X is set)."""
    ra = np.ascontiguousarray(A).view(np.float64)
    rb = np.ascontiguousarray(B).view(np.float64)
    s2 = lambda st: [2 * x for x in st]
    sa = -1.0 if conj_A else 1.0
    sb = -1.0 if conj_B else 1.0
    nc = int(np.asarray(C_).size)

    def real_prod(pa, pb):
        out = np.zeros(2 * nc)
        return np.asarray(tensor_mult(1.0, ra[pa:], lenA, s2(strA), idxA, rb[pb:], lenB, s2(strB), idxB,
                                      0.0, out, lenC, s2(strC), idxC))

    rr = real_prod(0, 0)
    ii = real_prod(1, 1)
    ri = real_prod(0, 1)
    ir = real_prod(1, 0)
    # the real contractions write C's layout at even offsets (stride 2)
    P = (rr - sa * sb * ii)[0::2] + 1j * (sb * ri + sa * ir)[0::2]
    c = np.array(C_, np.complex128)
    # elements C's labels address: a repeated label is one index over the
    # sum of its strides (the diagonal, src/iface/3t/mult.cxx:44-48); the
    # rest of the storage is untouched
    if any(n == 0 for n in lenC):
        return c
    diag = {}
    for n, st, lab in zip(lenC, strC, idxC):
        diag[lab] = (n, diag.get(lab, (n, 0))[1] + st)
    off = np.zeros(1, np.int64)
    for n, st in diag.values():
        off = (off[:, None] + np.arange(n, dtype=np.int64)[None, :] * st).ravel()
    hit = np.unique(off)
    c[hit] = complex(alpha) * P[hit] + (complex(beta) * c[hit] if beta != 0 else 0)
    return c


# ---------------------------------------------------------------- indexed

def _dt(*xs):
    """complex128 when any of xs (arrays, lists of arrays, scalars) is
    complex, else float64."""
    def cx(x):
        if isinstance(x, (list, tuple)):
            return any(cx(y) for y in x)
        return np.iscomplexobj(x)
    return np.complex128 if any(cx(x) for x in xs) else np.float64


def indexed_full(entries, indices, factors, dense_len, idx_len, conj=False):
    """The full tensor an indexed tensor describes (dense dimensions first),
    as block_to_full builds it (src/internal/1t/indexed/util.hpp:14-43):
    each entry, times its factor, written at its index values (conj: the
    complex conjugate of each entry, the factor as given)."""
    dt = _dt(entries, factors)
    full = np.zeros(list(dense_len) + list(idx_len), dt)
    nd = len(dense_len)
    for e, ix, f in zip(entries, indices, factors):
        e = np.asarray(e, dt)
        full[(slice(None),) * nd + tuple(ix)] = f * (np.conj(e) if conj else e)
    return full


def indexed_mult(alpha, A, idx_A, B, idx_B, beta, C_, idx_C, conj_A=False, conj_B=False):
    """Indexed contraction on host arrays, the reference's FULL algorithm
    (src/internal/3t/indexed/mult.cxx:15-59) behind its front end
    (src/iface/3t/mult.cxx:299-402) and scaling rule
    (src/internal/3t/indexed/mult.cxx:288-296, src/internal/1t/indexed/
    scale.cxx:12-30, set.cxx:12-20).  Operands are dicts with "entries"
    (dense numpy arrays, logical shape), "indices" (tuples), "factors",
    "dense_len", "idx_len"; returns C's new entries (float64, or complex128
    when any input is complex; conj_X conjugates X's entries)."""
    dt = _dt(alpha, beta, A["entries"], A["factors"], B["entries"], B["factors"], C_["entries"],
             C_["factors"])
    out = []
    for e, f in zip(C_["entries"], C_["factors"]):
        e = np.asarray(e, dt)
        if beta == 0.0 or (beta != 1.0 and f == 0.0):
            out.append(np.zeros_like(e))
        elif beta != 1.0:
            out.append(e * (f * beta))
        else:
            out.append(e.copy())
    if alpha == 0.0:
        return out
    full = {}
    for name, X, cj in (("A", A, conj_A), ("B", B, conj_B)):
        full[name] = indexed_full(X["entries"], X["indices"], X["factors"], X["dense_len"],
                                  X["idx_len"], cj)
    prod = alpha * np.einsum("%s,%s->%s" % (idx_A, idx_B, idx_C), full["A"], full["B"])
    nd = len(C_["dense_len"])
    for i, (ix, f) in enumerate(zip(C_["indices"], C_["factors"])):
        out[i] = out[i] + f * prod[(slice(None),) * nd + tuple(ix)]
    return out


# ---------------------------------------------------------------- indexed DPD

def _irrep_blocks(nirrep, ndim, irrep):
    """Every irrep combination of ndim dimensions whose XOR is irrep."""
    import itertools
    return [c for c in itertools.product(range(nirrep), repeat=ndim)
            if (np.bitwise_xor.reduce(c) if c else 0) == irrep]


def indexed_dpd_expand(X):
    """An indexed DPD operand as a plain indexed one over full extents
    (block_to_full, src/internal/1t/indexed_dpd/util.hpp): each entry's DPD
    storage spread into a dense array of the summed per-irrep lengths,
    index values shifted by the offset of their irrep."""
    lens, nir, nd = X["lens"], X["nirrep"], X["dense_ndim"]
    full = [int(sum(l)) for l in lens]
    ioff = [np.concatenate([[0], np.cumsum(l)[:-1]]).astype(int) for l in lens]
    dense_irrep = X["irrep"]
    for r in X["idx_irrep"]:
        dense_irrep ^= r
    d = dpd_init(dense_irrep, nir, [lens[i] for i in range(nd)], X["layout"]) if nd else None
    entries = []
    dt = _dt(X["entries"])
    for e in X["entries"]:
        out = np.zeros(full[:nd], dt)
        ea = np.ascontiguousarray(np.asarray(e, dt))
        if nd:
            for irr in _irrep_blocks(nir, nd, dense_irrep):
                bl, off, st = dpd_get_block(d, list(irr))
                if 0 in bl:
                    continue
                src = np.lib.stride_tricks.as_strided(ea[off:], bl, [ea.itemsize * x for x in st])
                out[tuple(slice(ioff[i][irr[i]], ioff[i][irr[i]] + bl[i]) for i in range(nd))] = src
        else:
            out = ea.reshape(())
        entries.append(out)
    idx = [tuple(ioff[nd + j][X["idx_irrep"][j]] + v for j, v in enumerate(ix)) for ix in X["indices"]]
    return dict(entries=entries, indices=idx, factors=list(X["factors"]), dense_len=full[:nd],
                idx_len=full[nd:]), d, dense_irrep, ioff


def indexed_dpd_mult(alpha, A, idx_A, B, idx_B, beta, C_, idx_C, conj_A=False, conj_B=False):
    """Indexed DPD contraction on host storage by the reference's FULL
    algorithm (src/internal/3t/indexed_dpd/mult.cxx:68-118): A, B and C
    expanded to full indexed tensors, the indexed oracle above, and C's
    symmetry-allowed blocks read back.  Operands are dicts with "entries"
    (1-D DPD storage), "indices" (irrep-local), "factors", "irrep",
    "nirrep", "lens" (ndim x nirrep), "dense_ndim", "idx_irrep", "layout";
    returns C's new entries (float64 storage)."""
    ea, _, _, _ = indexed_dpd_expand(A)
    eb, _, _, _ = indexed_dpd_expand(B)
    ec, dc, cirr, ioff = indexed_dpd_expand(C_)
    res = indexed_mult(alpha, ea, idx_A, eb, idx_B, beta, ec, idx_C, conj_A, conj_B)
    nd, nir = C_["dense_ndim"], C_["nirrep"]
    out = []
    dt = _dt(alpha, beta, A["entries"], B["entries"], C_["entries"], A["factors"], B["factors"], C_["factors"])
    for e0, full in zip(C_["entries"], res):
        e = np.array(e0, dt)
        if nd:
            for irr in _irrep_blocks(nir, nd, cirr):
                bl, off, st = dpd_get_block(dc, list(irr))
                if 0 in bl:
                    continue
                dst = np.lib.stride_tricks.as_strided(e[off:], bl, [e.itemsize * x for x in st])
                dst[...] = full[tuple(slice(ioff[i][irr[i]], ioff[i][irr[i]] + bl[i]) for i in range(nd))]
        else:
            e[...] = full
        out.append(e)
    return out


def fill_uniform(n, seed):
    x = np.empty(n, np.float64)
    lib().orc_fill_uniform_d(x.ctypes.data_as(PD), I64(n), C.c_uint64(seed))
    return x


# ------------------------------------------------------------ FP64-accurate mode
# No counterpart in the reference.  Restates the Ozaki scheme as the library
# runs it (Ozaki, Ogita, Oishi and Rump, Numer. Algorithms 59 (2012); integer
# form after Ootomo, Ozaki and Yokota, IJHPCA 2024), with the same operation
# order, so that the library's result can be compared bit for bit.

def ozaki_exponents(X):
    """Per row of X: the smallest e with max|x| < 2^e, from the bit pattern of
    the row maximum (0 for an all-zero row, -1022 for a subnormal maximum)."""
    m = np.max(np.abs(X), axis=1) if X.shape[1] else np.zeros(X.shape[0])
    biased = (m.view(np.uint64) >> np.uint64(52)).astype(np.int64)
    e = np.where(biased == 0, -1022, biased - 1022)
    return np.where(m == 0, 0, e).astype(np.int64)


def ozaki_digits(X, S):
    """S signed 7-bit digit planes of the row-scaled X: X[i] 2^-e_i =
    sum_t D[t][i] 2^(-7 (t + 1)) + r, each step exact in float64."""
    e = ozaki_exponents(X)
    v = np.ldexp(X.astype(np.float64), -e[:, None])
    D = np.zeros((S,) + X.shape, dtype=np.int64)
    for t in range(S):
        v = v * 128.0
        d = np.trunc(v)
        v = v - d
        D[t] = d.astype(np.int64)
    return D, e


def ozaki_gemm(A, B, S=8, flush_k=8160):
    """C = A B (m x k times k x n) as the FP64-accurate mode forms it: digit
    pairs t + u <= S + 1 (1-based), exact integer sums per diagonal over
    blocks of flush_k, added to an FP64 sum in diagonal order per block,
    then scaled by 2^(e_i + f_j)."""
    A = np.asarray(A, np.float64)
    B = np.asarray(B, np.float64)
    DA, ea = ozaki_digits(A, S)
    DB, eb = ozaki_digits(B.T, S)
    m, k = A.shape
    n = B.shape[1]
    tot = np.zeros((m, n))
    for k0 in range(0, max(k, 1), flush_k):
        ks = slice(k0, min(k0 + flush_k, k))
        acc = [np.zeros((m, n), dtype=np.int64) for _ in range(S)]
        for t in range(S):
            for u in range(S - t):
                # integer digits, |sum| <= 127^2 * flush_k < 2^53: a float64
                # product is exact
                acc[t + u] += (DA[t][:, ks].astype(np.float64) @ DB[u][:, ks].T.astype(np.float64)).astype(np.int64)
        for d in range(S):
            tot = tot + np.ldexp(1.0, -7 * (d + 2)) * acc[d].astype(np.float64)
    return np.ldexp(tot, ea[:, None] + eb[None, :])
