"""The C-ABI boundary on the CPU: the library loads, exports every declared
symbol, and its host-side integer work (canonical forms, block-scatter
vectors, DPD layout and patch lists) is bit-exact with the oracle.  No
device work is started here."""
import ctypes as C
import itertools
import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import tblis_amd as T
from oracle import pyoracle as po
from tests.helpers import layout_for, random_contraction, storage_size

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "tblis.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(tblis_[a-z0-9_]+)\s*\(", text))
    names |= set(re.findall(r"extern\s+const\s+tblis_comm\s*\*\s*const\s+(tblis_\w+)", text))
    return names


def test_library_exports_every_declared_symbol():
    L = T.lib()
    declared = _declared_symbols()
    assert declared == set(T.EXPORTED)
    for name in declared:
        assert hasattr(L, name), name


def test_abi_layout():
    assert C.sizeof(T.Scalar) == 24 and C.sizeof(T.Tensor) == 64
    assert T.Tensor.scalar.offset == 8 and T.Tensor.data.offset == 32
    assert T.Tensor.ndim.offset == 40 and T.Tensor.len.offset == 48 and T.Tensor.stride.offset == 56
    assert T.Scalar.type.offset == 16


def test_matrix_abi_layout_and_init():
    """tblis_matrix: 72 bytes, type@0 conj@4 scalar@8 data@32 m@40 n@48 rs@56
    cs@64 (src/util/basic_types.h:352-428) -- checked in ctypes and by
    compiling the header."""
    assert C.sizeof(T.Matrix) == 72
    assert (T.Matrix.scalar.offset, T.Matrix.data.offset, T.Matrix.m.offset, T.Matrix.n.offset,
            T.Matrix.rs.offset, T.Matrix.cs.offset) == (8, 32, 40, 48, 56, 64)
    src = os.path.join(ROOT, "_build_abi_probe.c")
    exe = os.path.join(ROOT, "_build_abi_probe")
    try:
        with open(src, "w") as f:
            f.write('#include <stddef.h>\n#include <stdio.h>\n#include "tblis.h"\n'
                    'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(tblis_matrix),'
                    'offsetof(tblis_matrix, scalar), offsetof(tblis_matrix, data), offsetof(tblis_matrix, m),'
                    'offsetof(tblis_matrix, cs), sizeof(tblis_tensor), offsetof(tblis_tensor, stride));'
                    'return 0;}\n')
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
        assert [int(x) for x in out] == [72, 8, 32, 40, 64, 64, 56]
    finally:
        for f in (src, exe):
            if os.path.exists(f):
                os.remove(f)
    L = T.lib()
    m = T.Matrix()
    L.tblis_init_matrix_scaled_d(C.byref(m), C.c_double(-1.5), 3, 4, None, 1, 3)
    assert (m.type, m.conj, m.scalar.data.d, m.m, m.n, m.rs, m.cs) == (T.TYPE_DOUBLE, 0, -1.5, 3, 4, 1, 3)
    L.tblis_init_matrix_s(C.byref(m), 5, 2, None, 2, 1)
    assert (m.type, m.scalar.data.s, m.m, m.n, m.rs, m.cs) == (T.TYPE_SINGLE, 1.0, 5, 2, 2, 1)


def test_matrix_mult_argument_errors():
    """Shape and conjugation errors are reported before any device access."""
    L = T.lib()
    L.tblis_hip_set_error_mode(1)
    try:
        a, b, c = T.Matrix(), T.Matrix(), T.Matrix()
        L.tblis_init_matrix_d(C.byref(a), 3, 4, None, 1, 3)
        L.tblis_init_matrix_d(C.byref(b), 5, 2, None, 1, 5)
        L.tblis_init_matrix_d(C.byref(c), 3, 2, None, 1, 3)
        L.tblis_matrix_mult(None, None, C.byref(a), C.byref(b), C.byref(c))
        assert "shapes do not agree" in T.last_error()
        L.tblis_init_matrix_d(C.byref(b), 4, 2, None, 1, 4)
        a.conj = 1
        L.tblis_matrix_mult(None, None, C.byref(a), C.byref(b), C.byref(c))
        assert "conjugation" in T.last_error()
    finally:
        L.tblis_hip_set_error_mode(0)


def test_init_tensor_and_threads():
    L = T.lib()
    lens = (C.c_ssize_t * 2)(3, 4)
    strides = (C.c_ssize_t * 2)(1, 3)
    t = T.Tensor()
    L.tblis_init_tensor_scaled_d(C.byref(t), C.c_double(2.5), 2, lens, None, strides)
    assert t.type == T.TYPE_DOUBLE and t.scalar.data.d == 2.5 and t.ndim == 2 and t.conj == 0
    L.tblis_init_tensor_s(C.byref(t), 1, lens, None, strides)
    assert t.type == T.TYPE_SINGLE and t.scalar.data.s == 1.0
    L.tblis_set_num_threads(7)
    assert L.tblis_get_num_threads() == 7
    L.tblis_set_num_threads(0)


class _HostDesc:
    """A tblis_tensor with host arrays only (data never touched)."""

    def __init__(self, lens, strides, dtype=T.TYPE_DOUBLE):
        n = len(lens)
        self.l = (C.c_ssize_t * max(n, 1))(*lens)
        self.s = (C.c_ssize_t * max(n, 1))(*strides)
        self.d = T.Tensor()
        self.d.type = dtype
        self.d.scalar.data.d = 1.0
        self.d.ndim = n
        self.d.len = C.cast(self.l, C.POINTER(C.c_ssize_t))
        self.d.stride = C.cast(self.s, C.POINTER(C.c_ssize_t))


def _product_canonical(lA, sA, iA, lB, sB, iB, lC, sC, iC):
    a, b, c = _HostDesc(lA, sA), _HostDesc(lB, sB), _HostDesc(lC, sC)
    out = (C.c_int64 * 4096)()
    T.lib().tblis_hip_set_error_mode(1)
    n = T.lib().tblis_hip_debug_canonical(C.byref(a.d), iA.encode("latin1"), C.byref(b.d),
                                          iB.encode("latin1"), C.byref(c.d), iC.encode("latin1"),
                                          out, 4096)
    if n < 0:
        return None
    v = list(out)[:n]
    res, pos = [], 0

    def take(nslot):
        nonlocal pos
        k = v[pos]
        pos += 1
        lens = v[pos:pos + k]
        pos += k
        st = []
        for _ in range(nslot):
            st.append(v[pos:pos + k])
            pos += k
        return lens, st

    for nslot in (2, 2, 2, 3, 2, 2, 2, 3):
        res.append(take(nslot))
    res.append(tuple(v[pos:pos + 3]))
    return res


def _oracle_canonical(lA, sA, iA, lB, sB, iB, lC, sC, iC):
    g = po.canonicalize(lA, sA, iA, lB, sB, iB, lC, sC, iC)
    if g is None:
        return None
    d = g.as_dict()
    t = po.blis_reorder(g).as_dict()
    return [(d["len_ab"], [d["sA_ab"], d["sB_ab"]]), (d["len_ac"], [d["sA_ac"], d["sC_ac"]]),
            (d["len_bc"], [d["sB_bc"], d["sC_bc"]]),
            (d["len_abc"], [d["sA_abc"], d["sB_abc"], d["sC_abc"]]),
            (t["len_m"], [t["sA_m"], t["sC_m"]]), (t["len_n"], [t["sB_n"], t["sC_n"]]),
            (t["len_k"], [t["sA_k"], t["sB_k"]]), (t["len_l"], [t["sA_l"], t["sB_l"], t["sC_l"]]),
            tuple(t["pack_3d"])]


@pytest.mark.parametrize("seed", range(150))
def test_canonical_form_bit_exact(seed):
    rng = np.random.default_rng(1000 + seed)
    iA, iB, iC, ext = random_contraction(rng, max_dims=3, max_len=5, allow_diag=seed % 4 == 0)
    kinds = ["col", "row", "random"]
    (lA, sA), (lB, sB), (lC, sC) = [layout_for(i, ext, rng, kinds[(seed + j) % 3])
                                    for j, i in enumerate((iA, iB, iC))]
    assert _product_canonical(lA, sA, iA, lB, sB, iB, lC, sC, iC) == \
        _oracle_canonical(lA, sA, iA, lB, sB, iB, lC, sC, iC)


def test_canonical_workloads_bit_exact():
    cases = [("abef", "efij", "abij", dict(a=24, b=24, e=24, f=24, i=24, j=24)),
             ("aebf", "fjei", "aibj", dict(a=64, b=64, e=64, f=64, i=32, j=32)),
             ("aebf", "ecfd", "abcd", dict(a=96, b=96, c=96, d=96, e=96, f=96))]
    from tests.helpers import col_major
    for iA, iB, iC, e in cases:
        args = []
        for idx in (iA, iB, iC):
            lens = [e[c] for c in idx]
            args += [lens, col_major(lens), idx]
        assert _product_canonical(*args) == _oracle_canonical(*args)


def test_invalid_contraction_reports_error():
    a, b, c = _HostDesc([2, 3], [1, 2]), _HostDesc([3], [1]), _HostDesc([3], [1])
    L = T.lib()
    L.tblis_hip_set_error_mode(1)
    L.tblis_tensor_mult(None, None, C.byref(a.d), b"ab", C.byref(b.d), b"c", C.byref(c.d), b"c")
    assert "only one tensor" in T.last_error()
    # unequal lengths for a shared label
    a, b, c = _HostDesc([2, 3], [1, 2]), _HostDesc([4], [1]), _HostDesc([2], [1])
    L.tblis_tensor_mult(None, None, C.byref(a.d), b"ab", C.byref(b.d), b"b", C.byref(c.d), b"a")
    assert "unequal" in T.last_error()
    # mixed types
    a2 = _HostDesc([2], [1], T.TYPE_SINGLE)
    L.tblis_tensor_mult(None, None, C.byref(a2.d), b"a", C.byref(b.d), b"b", C.byref(c.d), b"ab")
    assert "same type" in T.last_error()


def test_abort_mode_and_unknown_config():
    code = ("import ctypes as C, tblis_amd as T; L=T.lib(); L.tblis_get_config(b'nope')")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True)
    assert r.returncode != 0 and "No configuration named nope" in r.stderr
    code = ("import ctypes as C, tblis_amd as T\n"
            "from tests.test_boundary import _HostDesc\n"
            "a=_HostDesc([2,3],[1,2]); b=_HostDesc([3],[1]); c=_HostDesc([3],[1])\n"
            "T.lib().tblis_tensor_mult(None,None,C.byref(a.d),b'ab',C.byref(b.d),b'c',"
            "C.byref(c.d),b'c')\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True)
    assert r.returncode != 0 and "tblis: invalid contraction" in r.stderr
    for name in ("mi355x", "fp64_mfma", "bf16x3", "bf16x3_fast", "reference", "ozaki"):
        assert T.get_config(name)


def _product_scatter(lens, strides, BS, off, size, pack_3d, es):
    n = len(lens)
    out = (C.c_ssize_t * max(size, 1))()
    bs = (C.c_ssize_t * max(size, 1))()
    T.lib().tblis_hip_set_error_mode(1)
    T.lib().tblis_hip_debug_fill_block_scatter(n, (C.c_ssize_t * max(n, 1))(*lens),
                                               (C.c_ssize_t * max(n, 1))(*strides), BS, off, size,
                                               out, bs, int(pack_3d), es)
    assert T.last_error() == ""
    return np.array(out[:size]), np.array(bs[:size])


def _panel_side(lens, strides, G, r0):
    n = len(lens)
    cap = int(np.prod(lens)) + 2 * (int(np.prod(lens)) // G + 3)
    offs = (C.c_ssize_t * cap)()
    tab = (C.c_ssize_t * cap)()
    got = T.lib().tblis_hip_debug_panel_side(n, (C.c_ssize_t * max(n, 1))(*lens),
                                             (C.c_ssize_t * max(n, 1))(*strides), G, r0, offs, tab, cap)
    assert got == int(np.prod(lens)), got
    ng = -(-(r0 + got) // G) - r0 // G
    return np.array(offs[:got]), np.array(tab[:2 * ng]).reshape(ng, 2)


@pytest.mark.parametrize("seed", range(120))
def test_panel_side_tables_bit_exact(seed):
    """The tables the default (panel) path builds, against the oracle: the
    offset vector equals the reference's plain scatter vector
    (block_scatter_matrix.hpp:68-113), and every fragment group's stride is
    the reference's block stride at that block size (:232-247, 0 where the
    group is not affine or has stride 0, 1 for one element) with the group's
    base at its first element; with the side placed at panel row r0 the
    groups are cut at multiples of G from the panel's row 0."""
    rng = np.random.default_rng(5000 + seed)
    nd = int(rng.integers(1, 5))
    lens = [int(rng.integers(1, 12)) for _ in range(nd)]
    strides = [int(rng.choice([0, 1, -3, 7, 64, 4096, int(rng.integers(-50, 5000))])) for _ in range(nd)]
    n = int(np.prod(lens))
    G = int(rng.choice([8, 16, 32]))
    r0 = 0 if seed % 2 == 0 else int(rng.integers(0, 3 * G))
    offs, tab = _panel_side(lens, strides, G, r0)
    assert (offs == po.fill_scatter(lens, strides, G, 0, n, False)).all()
    SC = -(1 << 63)
    for gi in range(len(tab)):
        lo = max((r0 // G + gi) * G - r0, 0)
        hi = min((r0 // G + gi) * G - r0 + G, n)
        blk = offs[lo:hi]
        if len(blk) == 1:
            want = 1
        else:
            d = np.diff(blk)
            want = int(d[0]) if (d == d[0]).all() else SC
        assert tab[gi, 1] == want, (gi, blk, tab[gi])
        if want != SC:
            # base + (element index from the group's first panel row) * stride
            first = (r0 // G + gi) * G
            assert tab[gi, 0] + (lo + r0 - first) * want == blk[0]
        if r0 == 0:
            ref_bs = po.fill_block_stride(G, offs)[lo]
            assert ref_bs == (0 if want in (SC, 0) else want)


@pytest.mark.parametrize("seed", range(200))
def test_block_scatter_bit_exact(seed):
    rng = np.random.default_rng(2000 + seed)
    pack = seed % 2 == 1
    nd = int(rng.integers(2 if pack else 0, 6))
    lens = [int(rng.integers(1, 14)) for _ in range(nd)]
    strides = [int(rng.integers(-50, 5000)) for _ in range(nd)]
    tot = int(np.prod(lens)) if nd else 1
    off = int(rng.integers(0, tot)) if seed % 3 else 0
    size = int(rng.integers(1, tot - off + 1)) if seed % 5 else tot - off
    if nd == 0:
        off, size = 0, 1
    BS = int(rng.choice([1, 2, 3, 4, 6, 8, 16, 32, 64, 128]))
    es = int(rng.choice([4, 8]))
    got, gbs = _product_scatter(lens, strides, BS, off, size, pack, es)
    want = po.fill_scatter(lens, strides, BS, off, size, pack, es)
    assert (got == want).all()
    wbs = po.fill_block_stride(BS, want)
    assert (gbs[::BS] == wbs[::BS]).all()


def test_block_scatter_rejects_reference_crash_case():
    T.lib().tblis_hip_set_error_mode(1)
    out = (C.c_ssize_t * 30)()
    T.lib().tblis_hip_debug_fill_block_scatter(2, (C.c_ssize_t * 2)(10, 3), (C.c_ssize_t * 2)(1, 10),
                                               5, 0, 30, out, None, 1, 8)
    assert "gcd" in T.last_error()


# ---------------------------------------------------------------- DPD
def _dpd_desc(irrep, nirrep, lens, layout):
    import torch
    return T._DpdDesc(torch.empty(0, dtype=torch.float64), irrep, nirrep, lens, layout)


def test_dpd_layout_known_answers_product():
    kat = json.load(open(os.path.join(GOLD, "dpd_varray_kat.json")))
    for j, code in enumerate(kat["layout_codes"]):
        d = _dpd_desc(kat["irrep"], kat["nirrep"], kat["len"], code)
        assert T.lib().tblis_dpd_tensor_size(C.byref(d.d)) == kat["total_size"]
        for i, irr in enumerate(kat["irreps"]):
            ln = (C.c_ssize_t * 4)()
            st = (C.c_ssize_t * 4)()
            off = C.c_ssize_t()
            T.lib().tblis_hip_debug_dpd_block(C.byref(d.d), (C.c_uint * 4)(*irr), ln, C.byref(off), st)
            assert list(ln) == kat["lengths"][i]
            assert list(st) == kat["strides"][j][i]
            assert off.value == kat["offsets"][j][i]


@pytest.mark.parametrize("seed", range(16))
def test_dpd_layout_matches_oracle(seed):
    rng = np.random.default_rng(3000 + seed)
    nirrep = [1, 2, 4, 8][seed % 4]
    ndim = int(rng.integers(1, 6))
    lens = [[int(rng.integers(0, 5)) for _ in range(nirrep)] for _ in range(ndim)]
    irrep = int(rng.integers(0, nirrep))
    for layout in range(6):
        d = _dpd_desc(irrep, nirrep, lens, layout)
        od = po.dpd_init(irrep, nirrep, lens, layout)
        assert T.lib().tblis_dpd_tensor_size(C.byref(d.d)) == po.dpd_size(od)
        for combo in itertools.islice(itertools.product(range(nirrep), repeat=ndim), 64):
            if np.bitwise_xor.reduce(combo) != irrep:
                continue
            ln = (C.c_ssize_t * ndim)()
            st = (C.c_ssize_t * ndim)()
            off = C.c_ssize_t()
            T.lib().tblis_hip_debug_dpd_block(C.byref(d.d), (C.c_uint * ndim)(*combo), ln,
                                              C.byref(off), st)
            assert (list(ln), off.value, list(st)) == po.dpd_get_block(od, list(combo))


def _product_patches(dA, iA, dB, iB, dC, iC, which, r, BSr, BSc):
    npatch = (C.c_int * 2)()
    mp, ms = 1 << 14, 1 << 20
    rows = (C.c_ssize_t * mp)()
    cols = (C.c_ssize_t * mp)()
    base = (C.c_ssize_t * mp)()
    rs = (C.c_ssize_t * ms)()
    cs = (C.c_ssize_t * ms)()
    T.lib().tblis_hip_set_error_mode(1)
    n = T.lib().tblis_hip_debug_dpd_patches(C.byref(dA.d), iA.encode(), C.byref(dB.d), iB.encode(),
                                            C.byref(dC.d), iC.encode(), which, r, BSr, BSc, npatch,
                                            rows, cols, base, rs, cs, mp, ms)
    assert n >= 0, T.last_error()
    nr, nc = sum(rows[:n]), sum(cols[:n])
    return {"npatch": (npatch[0], npatch[1]), "rows": np.array(rows[:n]),
            "cols": np.array(cols[:n]), "base": np.array(base[:n]), "rscat": np.array(rs[:nr]),
            "cscat": np.array(cs[:nc])}


@pytest.mark.parametrize("seed", range(12))
def test_dpd_patch_scatter_bit_exact(seed):
    rng = np.random.default_rng(4000 + seed)
    nirrep = [2, 4, 8][seed % 3]
    layout = seed % 6
    cases = [("abef", "efij", "abij"), ("aebf", "fjei", "aibj"), ("abe", "eij", "abij")]
    iA, iB, iC = cases[seed % 3]
    ext = {c: [int(rng.integers(0, 4)) for _ in range(nirrep)] for c in sorted(set(iA + iB + iC))}
    irA, irB = int(rng.integers(0, nirrep)), int(rng.integers(0, nirrep))
    irC = irA ^ irB
    lens = lambda idx: [ext[c] for c in idx]
    oA, oB, oC = (po.dpd_init(irA, nirrep, lens(iA), layout), po.dpd_init(irB, nirrep, lens(iB), layout),
                  po.dpd_init(irC, nirrep, lens(iC), layout))
    pA, pB, pC = (_dpd_desc(irA, nirrep, lens(iA), layout), _dpd_desc(irB, nirrep, lens(iB), layout),
                  _dpd_desc(irC, nirrep, lens(iC), layout))
    for which in range(3):
        for r in range(nirrep):
            BSr, BSc = [(4, 2), (8, 4), (2, 8)][which]
            want = po.dpd_patch_scatter(oA, iA, oB, iB, oC, iC, which, r, BSr, BSc)
            got = _product_patches(pA, iA, pB, iB, pC, iC, which, r, BSr, BSc)
            assert got["npatch"] == want["npatch"]
            for k in ("rows", "cols", "base", "rscat", "cscat"):
                assert (got[k] == want[k]).all(), (which, r, k)


def test_level1_argument_errors():
    """Level-1 label errors are reported before any device access."""
    L = T.lib()
    L.tblis_hip_set_error_mode(1)
    try:
        a, b = _HostDesc([3, 4], [1, 3]), _HostDesc([3, 5], [1, 3])
        L.tblis_tensor_add(None, None, C.byref(a.d), b"ab", C.byref(b.d), b"ac")
        assert "only in A and labels present only in B" in T.last_error()
        b = _HostDesc([3, 5], [1, 3])
        L.tblis_tensor_add(None, None, C.byref(a.d), b"ab", C.byref(b.d), b"ab")
        assert "different lengths" in T.last_error()
        a = _HostDesc([3, 4], [1, 3])
        L.tblis_tensor_scale(None, None, C.byref(a.d), b"aa")
        assert "unequal lengths" in T.last_error()
        s = T.Scalar()
        s.type = T.TYPE_SINGLE
        s.data.s = 1.0
        L.tblis_tensor_set(None, None, C.byref(s), C.byref(a.d), b"ab")
        assert "type differs" in T.last_error()
        a.d.type = T.TYPE_DCOMPLEX
        L.tblis_tensor_scale(None, None, C.byref(a.d), b"ab")
        assert "complex" in T.last_error().lower()
    finally:
        L.tblis_hip_set_error_mode(0)


def test_indexed_abi_layout_and_init():
    """tblis_indexed_tensor: 88 bytes; the ctypes mirror and the compiled
    header agree on every offset, and the init calls fill it."""
    import tempfile
    fields = ["type", "conj", "scalar", "ndim", "dense_ndim", "len", "dense_stride", "num_indices",
              "data", "indices", "factor"]
    assert C.sizeof(T.IndexedTensor) == 88
    want = [C.sizeof(T.IndexedTensor)] + [getattr(T.IndexedTensor, f).offset for f in fields]
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "probe.c"), os.path.join(d, "probe")
        with open(src, "w") as f:
            f.write('#include <stddef.h>\n#include <stdio.h>\n#include "tblis.h"\nint main(void){'
                    'printf("%zu", sizeof(tblis_indexed_tensor));' +
                    "".join('printf(" %%zu", offsetof(tblis_indexed_tensor, %s));' % x for x in fields) +
                    'printf("\\n"); return 0;}\n')
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    assert [int(x) for x in out] == want
    L = T.lib()
    lens = (C.c_ssize_t * 3)(4, 5, 3)
    strides = (C.c_ssize_t * 2)(1, 4)
    idx = (C.c_ssize_t * 2)(0, 2)
    data = (C.c_void_p * 2)(0x1000, 0x2000)
    fac = (C.c_double * 2)(1.0, 0.5)
    t = T.IndexedTensor()
    L.tblis_init_indexed_tensor_d(C.byref(t), C.c_double(2.0), 3, 2, lens, strides, 2, data, idx, fac)
    assert (t.type, t.scalar.data.d, t.ndim, t.dense_ndim, t.num_indices) == (T.TYPE_DOUBLE, 2.0, 3, 2, 2)
    assert t.len[2] == 3 and t.dense_stride[1] == 4 and t.indices[1] == 2 and t.data[1] == 0x2000


def _ix(lens, dense_ndim, strides, n, indices=None, dtype=T.TYPE_DOUBLE):
    d = T.IndexedTensor()
    keep = [(C.c_ssize_t * max(len(lens), 1))(*lens), (C.c_ssize_t * max(len(strides), 1))(*strides),
            (C.c_void_p * max(n, 1))(*([0x10000 + 0x1000 * i for i in range(n)])),
            (C.c_ssize_t * max(len(indices or []), 1))(*(indices or []))]
    d.type = dtype
    d.scalar.type = dtype
    d.scalar.data.d = 1.0
    d.ndim, d.dense_ndim, d.num_indices = len(lens), dense_ndim, n
    d.len = C.cast(keep[0], C.POINTER(C.c_ssize_t))
    d.dense_stride = C.cast(keep[1], C.POINTER(C.c_ssize_t))
    d.data = C.cast(keep[2], C.POINTER(C.c_void_p))
    d.indices = C.cast(keep[3], C.POINTER(C.c_ssize_t))
    return d, keep


def test_indexed_argument_errors():
    """Label and index errors are reported before any device access."""
    L = T.lib()
    L.tblis_hip_set_error_mode(1)
    try:
        b, kb = _ix([4, 5], 2, [1, 4], 1)
        c, kc = _ix([3, 5], 2, [1, 3], 1)
        a, ka = _ix([3, 4], 1, [1], 1, [7])
        L.tblis_indexed_tensor_mult(None, None, C.byref(a), b"ae", C.byref(b), b"ej", C.byref(c), b"aj")
        assert "index value out of range" in T.last_error()
        a, ka = _ix([3, 4], 1, [1], 1, [2])
        L.tblis_indexed_tensor_mult(None, None, C.byref(a), b"aa", C.byref(b), b"ej", C.byref(c), b"aj")
        assert "repeated label" in T.last_error()
        L.tblis_indexed_tensor_mult(None, None, C.byref(a), b"ax", C.byref(b), b"ej", C.byref(c), b"aj")
        assert "only one tensor" in T.last_error()
        b2, kb2 = _ix([5, 5], 2, [1, 5], 1)
        L.tblis_indexed_tensor_mult(None, None, C.byref(a), b"ae", C.byref(b2), b"ej", C.byref(c), b"aj")
        assert "unequal lengths" in T.last_error()
        a3, ka3 = _ix([3, 4], 1, [1], 2, [0, 1], dtype=T.TYPE_SINGLE)
        L.tblis_indexed_tensor_mult(None, None, C.byref(a3), b"ae", C.byref(b), b"ej", C.byref(c), b"aj")
        assert "same type" in T.last_error()
    finally:
        L.tblis_hip_set_error_mode(0)


def test_cpp_header_compiles_links_and_reaches_library():
    """include/tblis.hpp: every tblis::mult overload instantiates for float
    and double, links against the library, and an invalid call made through
    it comes back with the library's message (no device access)."""
    import tempfile
    prog = r'''
#include <cstdio>
#include "tblis.hpp"
template <typename T> void all()
{
    tblis::varray_view<T> t{nullptr, {2, 3}, {1, 2}};
    tblis::matrix_view<T> m{nullptr, 2, 2, 1, 2};
    tblis::dpd_varray_view<T> d{nullptr, 0, 1, {{2}, {2}}};
    tblis::indexed_varray_view<T> x{{nullptr}, {2}, {1}, {2}, {0}, {}};
    tblis::indexed_dpd_varray_view<T> y{{nullptr}, 0, 1, {{2}, {2}}, 1, {0}, {0}, {}};
    if (false)
    {
        tblis::mult<T>(1, y, "az", y, "bz", 0, y, "ab");
        tblis::mult<T>(1, t, "ab", t, "bc", 0, t, "ac");
        tblis::mult<T>(1, m, m, 0, m);
        tblis::mult<T>(1, d, "ab", d, "bc", 0, d, "ac");
    }
    tblis::mult<T>(1, x, "az", x, "bz", 0, x, "ab");
    std::printf("%s\n", tblis_hip_get_last_error());
}
int main() { tblis_hip_set_error_mode(1); all<float>(); all<double>(); return 0; }
'''
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "probe.cpp"), os.path.join(d, "probe")
        with open(src, "w") as f:
            f.write(prog)
        libdir = os.path.join(ROOT, "tblis_amd")
        subprocess.run(["g++", "-std=c++14", "-I", os.path.join(ROOT, "include"), src, "-o", exe,
                        "-L", libdir, "-ltblis_amd", "-Wl,-rpath," + libdir], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.splitlines()
    assert out == ["A: null entry pointer"] * 2, out


def test_indexed_dpd_abi_layout():
    """tblis_indexed_dpd_tensor: 104 bytes; ctypes mirror and compiled header agree."""
    import tempfile
    fields = ["type", "conj", "scalar", "ndim", "dense_ndim", "irrep", "nirrep", "layout", "len",
              "idx_irrep", "num_indices", "data", "indices", "factor"]
    assert C.sizeof(T.IndexedDpdTensor) == 104
    want = [C.sizeof(T.IndexedDpdTensor)] + [getattr(T.IndexedDpdTensor, f).offset for f in fields]
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "probe.c"), os.path.join(d, "probe")
        with open(src, "w") as f:
            f.write('#include <stddef.h>\n#include <stdio.h>\n#include "tblis.h"\nint main(void){'
                    'printf("%zu", sizeof(tblis_indexed_dpd_tensor));' +
                    "".join('printf(" %%zu", offsetof(tblis_indexed_dpd_tensor, %s));' % x for x in fields) +
                    'printf("\\n"); return 0;}\n')
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    assert [int(x) for x in out] == want
