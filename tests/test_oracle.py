"""Pins the oracle (oracle/oracle.c) against the reference's own known-answer
tests, closed forms, brute force and invariants.  CPU only."""
import itertools
import json
import math
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import col_major, layout_for, random_contraction, row_major, storage_size

GOLD = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------- DPD layout
def test_dpd_layout_known_answers():
    """src/external/marray/test/dpd_varray.cxx:7-55: 6 layouts x 8 blocks."""
    kat = json.load(open(os.path.join(GOLD, "dpd_varray_kat.json")))
    for j, code in enumerate(kat["layout_codes"]):
        d = po.dpd_init(kat["irrep"], kat["nirrep"], kat["len"], code)
        assert po.dpd_size(d) == kat["total_size"]
        assert list(d.perm)[:4] == kat["perms"][j]
        for i, irr in enumerate(kat["irreps"]):
            ln, off, st = po.dpd_get_block(d, irr)
            assert ln == kat["lengths"][i], (kat["layouts"][j], irr)
            assert st == kat["strides"][j][i], (kat["layouts"][j], irr)
            assert off == kat["offsets"][j][i], (kat["layouts"][j], irr)


def _closed_form_size(irrep, lens, nirrep):
    # dpd_marray_base.hpp:460-496
    size = [lens[0][r] for r in range(nirrep)]
    for row in lens[1:]:
        size = [sum(size[r1 ^ r2] * row[r2] for r2 in range(nirrep)) for r1 in range(nirrep)]
    return size[irrep]


@pytest.mark.parametrize("seed", range(12))
def test_dpd_layout_tiles_storage(seed):
    """Every allowed block is a dense box; the blocks tile [0, size) exactly."""
    rng = np.random.default_rng(seed)
    nirrep = [1, 2, 4, 8][seed % 4]
    ndim = int(rng.integers(1, 5))
    lens = [[int(rng.integers(0, 4)) for _ in range(nirrep)] for _ in range(ndim)]
    irrep = int(rng.integers(0, nirrep))
    for layout in range(6):
        d = po.dpd_init(irrep, nirrep, lens, layout)
        total = po.dpd_size(d)
        assert total == _closed_form_size(irrep, lens, nirrep)
        seen = np.zeros(total, np.int64)
        for combo in itertools.product(range(nirrep), repeat=ndim):
            if np.bitwise_xor.reduce(combo) != irrep:
                continue
            ln, off, st = po.dpd_get_block(d, list(combo))
            assert ln == [lens[i][combo[i]] for i in range(ndim)]
            if 0 in ln:
                continue
            for idx in itertools.product(*[range(l) for l in ln]):
                seen[off + sum(i * s for i, s in zip(idx, st))] += 1
        assert (seen == 1).all()


# ---------------------------------------------------------------- scatter
def _plain_offsets(lens, strides, off, size):
    out = []
    for i in range(off, off + size):
        f, o = i, 0
        for l, s in zip(lens, strides):
            o += (f % l) * s
            f //= l
        out.append(o)
    return np.array(out, np.int64)


@pytest.mark.parametrize("seed", range(40))
def test_fill_scatter_plain_closed_form(seed):
    rng = np.random.default_rng(seed)
    nd = int(rng.integers(0, 5))
    lens = [int(rng.integers(1, 6)) for _ in range(nd)]
    strides = [int(rng.integers(-20, 40)) for _ in range(nd)]
    tot = int(np.prod(lens)) if nd else 1
    off = int(rng.integers(0, tot))
    size = int(rng.integers(1, tot - off + 1))
    if nd == 0:
        off, size = 0, 1
    got = po.fill_scatter(lens, strides, int(rng.integers(1, 9)), off, size, False)
    assert (got == _plain_offsets(lens, strides, off, size)).all()


def test_fill_scatter_pack3d_hand_example():
    """m0=10, m1=3 column-major, BS=2, CL=8 doubles: BS0=min(lcm(2,8),10)=8,
    BS1=min(8,max(2,2))=2; full 8x2 and 8x1 tiles first, then the two
    leftover columns of dim 0 row by row."""
    got = po.fill_scatter([10, 3], [1, 10], 2, 0, 30, True, 8)
    want = list(range(0, 8)) + list(range(10, 18)) + list(range(20, 28)) + [8, 9, 18, 19, 28, 29]
    assert got.tolist() == want


@pytest.mark.parametrize("seed", range(60))
def test_fill_scatter_pack3d_is_permutation_of_range(seed):
    """The pack_3d order covers exactly the multi-indices of [off, off+size)."""
    rng = np.random.default_rng(100 + seed)
    nd = int(rng.integers(2, 5))
    lens = [int(rng.integers(1, 13)) for _ in range(nd)]
    strides = col_major(lens)  # unique offsets
    tot = int(np.prod(lens))
    off = int(rng.integers(0, tot)) if seed % 2 else 0
    size = int(rng.integers(1, tot - off + 1)) if seed % 3 else tot - off
    BS = int(rng.choice([1, 2, 3, 4, 6, 8, 16]))
    es = int(rng.choice([4, 8, 16]))
    got = po.fill_scatter(lens, strides, BS, off, size, True, es)
    want = _plain_offsets(lens, strides, off, size)
    assert sorted(got.tolist()) == sorted(want.tolist())


@pytest.mark.parametrize("seed", range(30))
def test_fill_block_stride_brute_force(seed):
    rng = np.random.default_rng(200 + seed)
    n = int(rng.integers(1, 40))
    scat = np.cumsum(rng.choice([1, 1, 1, 2, 5], size=n)) - 1
    BS = int(rng.integers(1, 9))
    bs = po.fill_block_stride(BS, scat)
    for i in range(0, n, BS):
        blk = scat[i:i + BS]
        if len(blk) == 1:
            assert bs[i] == 1
        else:
            d = np.diff(blk)
            assert bs[i] == (d[0] if (d == d[0]).all() else 0)


def test_gcd_lcm():
    """The reference's binary gcd (block_scatter_matrix.hpp:18-42) shifts once
    per subtraction.  Worked by hand: gcd(7,5): 7,5 -> 1,5 -> 1,2 -> 1,0 ->
    0,0 -> 0.  With a power-of-two second argument (the cache-line length it
    is used with) it is either exact or 0."""
    assert po.gcd(12, 18) == 6 and po.gcd(64, 8) == 8 and po.gcd(0, 9) == 9
    assert po.gcd(6, 8) == 2 and po.gcd(3, 8) == 1 and po.gcd(24, 8) == 8
    assert po.gcd(7, 5) == 0 and po.gcd(5, 8) == 0
    for CL in (4, 8, 16):
        for BS in range(1, 129):
            assert po.gcd(BS, CL) in (0, math.gcd(BS, CL))
            if BS & (BS - 1) == 0:
                assert po.lcm(BS, CL) == max(BS, CL)
    assert po.lcm(5, 8) == 0


# ---------------------------------------------------------------- canonical form
def _ext(**kw):
    return kw


def _T(idx, ext, layout=col_major):
    lens = [ext[c] for c in idx]
    return lens, layout(lens), idx


def test_canonical_worked_forms():
    """Hand-derived canonical forms of the dense workloads."""
    e = _ext(a=24, b=24, e=24, f=24, i=24, j=24)
    t = po.blis_reorder(po.canonicalize(*_T("abef", e), *_T("efij", e), *_T("abij", e))).as_dict()
    assert (t["len_m"], t["len_n"], t["len_k"]) == ([576], [576], [576])
    assert (t["sA_m"], t["sC_m"], t["sB_n"], t["sC_n"], t["sA_k"], t["sB_k"]) == \
        ([1], [1], [576], [576], [576], [1])
    e = _ext(a=64, b=64, e=64, f=64, i=32, j=32)
    t = po.blis_reorder(po.canonicalize(*_T("aebf", e), *_T("fjei", e), *_T("aibj", e))).as_dict()
    assert t["len_m"] == [64, 64] and t["sA_m"] == [1, 4096] and t["sC_m"] == [1, 2048]
    assert t["len_n"] == [32, 32] and t["sB_n"] == [131072, 64] and t["sC_n"] == [64, 131072]
    assert t["len_k"] == [64, 64] and t["sA_k"] == [262144, 64] and t["sB_k"] == [1, 2048]
    assert t["pack_3d"] == (0, 0, 0)
    e = _ext(a=96, b=96, c=96, d=96, e=96, f=96)
    t = po.blis_reorder(po.canonicalize(*_T("aebf", e), *_T("ecfd", e), *_T("abcd", e))).as_dict()
    assert t["sA_m"] == [1, 9216] and t["sC_m"] == [1, 96]
    assert t["sB_n"] == [96, 884736] and t["sC_n"] == [9216, 884736]
    assert t["sA_k"] == [96, 884736] and t["sB_k"] == [1, 9216]


def test_canonical_dummy_dims_and_invalid():
    # C holds only batch labels: dummy labels \0 (A, C) and \1 (B, C) added
    g = po.canonicalize([3], [1], "a", [3], [1], "a", [3], [1], "a").as_dict()
    assert g["idx_abc"] == "a" and g["idx_ac"] == "\x00" and g["idx_bc"] == "\x01"
    assert g["len_ac"] == [1] and g["len_bc"] == [1]
    # A-only label
    assert po.canonicalize([2, 3], [1, 2], "ab", [3], [1], "c", [3], [1], "c") is None
    # free_idx
    assert po.free_idx("abc") == b"\x00"
    assert po.free_idx("\x00\x01\x03") == b"\x02"


def test_diagonal_and_fold():
    ln, st, ix = po.diagonal([3, 1, 3, 4], [1, 3, 3, 9], "aeab")
    assert (ln, st, ix) == ([3, 4], [4, 9], "ab")
    ln, ix, ss = po.fold([4, 5, 6], "abc", [[1, 4, 20], [1, 4, 99]])
    assert (ln, ix, ss) == ([20, 6], "ac", [[1, 20], [1, 99]])
    assert po.sort_by_stride([[5, 1, 5], [2, 9, 3]]) == [1, 0, 2]


def _element_triples(lensA, sA, iA, lensB, sB, iB, lensC, sC, iC):
    labels = sorted(set(iA) | set(iB) | set(iC))
    ext = {}
    for lens, idx in ((lensA, iA), (lensB, iB), (lensC, iC)):
        for l, c in zip(lens, idx):
            ext[c] = l
    out = []
    for vals in itertools.product(*[range(ext[c]) for c in labels]):
        v = dict(zip(labels, vals))
        out.append((sum(v[c] * s for c, s in zip(iA, sA)), sum(v[c] * s for c, s in zip(iB, sB)),
                    sum(v[c] * s for c, s in zip(iC, sC))))
    return sorted(out)


def _canonical_triples(g):
    out = []
    def offs(lens, strides):
        res = []
        for vals in itertools.product(*[range(l) for l in lens]):
            res.append(sum(v * s for v, s in zip(vals, strides)))
        return res
    d = g.as_dict()
    for ab in zip(offs(d["len_ab"], d["sA_ab"]), offs(d["len_ab"], d["sB_ab"])):
        for ac in zip(offs(d["len_ac"], d["sA_ac"]), offs(d["len_ac"], d["sC_ac"])):
            for bc in zip(offs(d["len_bc"], d["sB_bc"]), offs(d["len_bc"], d["sC_bc"])):
                for abc in zip(offs(d["len_abc"], d["sA_abc"]), offs(d["len_abc"], d["sB_abc"]),
                               offs(d["len_abc"], d["sC_abc"])):
                    out.append((ab[0] + ac[0] + abc[0], ab[1] + bc[0] + abc[1],
                                ac[1] + bc[1] + abc[2]))
    return sorted(out)


@pytest.mark.parametrize("seed", range(40))
def test_canonical_preserves_every_element_address(seed):
    """diagonal + group split + fold keep the (A, B, C) address triple of every
    term of the contraction."""
    rng = np.random.default_rng(300 + seed)
    iA, iB, iC, ext = random_contraction(rng, max_dims=2, max_len=3)
    (lA, sA), (lB, sB), (lC, sC) = [layout_for(i, ext, rng) for i in (iA, iB, iC)]
    g = po.canonicalize(lA, sA, iA, lB, sB, iB, lC, sC, iC)
    assert g is not None
    assert _canonical_triples(g) == _element_triples(lA, sA, iA, lB, sB, iB, lC, sC, iC)


@pytest.mark.parametrize("seed", range(20))
def test_canonical_invariant_under_dim_permutation(seed):
    rng = np.random.default_rng(400 + seed)
    iA, iB, iC, ext = random_contraction(rng)
    (lA, sA), (lB, sB), (lC, sC) = [layout_for(i, ext, rng) for i in (iA, iB, iC)]
    g1 = po.canonicalize(lA, sA, iA, lB, sB, iB, lC, sC, iC).as_dict()
    p = rng.permutation(len(iA))
    g2 = po.canonicalize([lA[i] for i in p], [sA[i] for i in p], "".join(iA[i] for i in p),
                         lB, sB, iB, lC, sC, iC).as_dict()
    assert g1 == g2


# ---------------------------------------------------------------- contraction
@pytest.mark.parametrize("seed", range(30))
def test_tensor_mult_matches_einsum(seed):
    rng = np.random.default_rng(500 + seed)
    iA, iB, iC, ext = random_contraction(rng, allow_diag=True)
    (lA, sA), (lB, sB), (lC, sC) = [layout_for(i, ext, rng) for i in (iA, iB, iC)]
    A = rng.uniform(-1, 1, storage_size(lA, sA))
    B = rng.uniform(-1, 1, storage_size(lB, sB))
    C0 = rng.uniform(-1, 1, storage_size(lC, sC))
    alpha, beta = 0.75, [0.0, 1.0, -0.5][seed % 3]
    view = lambda X, l, s: np.lib.stride_tricks.as_strided(X, shape=l, strides=[8 * x for x in s])
    ref = alpha * np.einsum("%s,%s->%s" % (iA, iB, iC), view(A, lA, sA), view(B, lB, sB))
    ref = ref + beta * view(C0, lC, sC)
    C = C0.copy()
    po.tensor_mult(alpha, A, lA, sA, iA, B, lB, sB, iB, beta, C, lC, sC, iC)
    np.testing.assert_allclose(view(C, lC, sC), ref, rtol=1e-13, atol=1e-13)


def test_tensor_mult_special_rules():
    e = dict(a=3, b=4, k=5)
    lA, sA = [3, 5], [1, 3]
    lB, sB = [5, 4], [1, 5]
    lC, sC = [3, 4], [1, 3]
    rng = np.random.default_rng(7)
    A, B = rng.uniform(-1, 1, 15), rng.uniform(-1, 1, 20)
    # beta == 0 overwrites without reading C
    C = np.full(12, np.nan)
    po.tensor_mult(1.0, A, lA, sA, "ak", B, lB, sB, "kb", 0.0, C, lC, sC, "ab")
    assert np.isfinite(C).all()
    # alpha == 0 reads neither A nor B
    An = np.full(15, np.nan)
    C = np.ones(12)
    po.tensor_mult(0.0, An, lA, sA, "ak", B, lB, sB, "kb", 2.0, C, lC, sC, "ab")
    assert (C == 2.0).all()
    # k == 0: C <- beta*C
    C = np.ones(12)
    po.tensor_mult(1.0, A, [3, 0], sA, "ak", B, [0, 4], sB, "kb", 3.0, C, lC, sC, "ab")
    assert (C == 3.0).all()
    # empty m: C untouched
    C = np.ones(12)
    po.tensor_mult(1.0, A, [0, 5], sA, "ak", B, lB, sB, "kb", 0.0, C, [0, 4], sC, "ab")
    assert (C == 1.0).all()


def _strided_view(X, lens, strides):
    return np.lib.stride_tricks.as_strided(X, shape=lens, strides=[X.itemsize * s for s in strides])


def _entry_offsets(d, ent_l, ent_m, ent_n):
    """C element offsets of flat (l, m, n) logical indices of a canonical form
    (first dimension of each group fastest)."""
    off = np.zeros(len(ent_l), np.int64)
    for arr, lens, strides in ((ent_l, d["len_abc"], d["sC_abc"]), (ent_m, d["len_ac"], d["sC_ac"]),
                               (ent_n, d["len_bc"], d["sC_bc"])):
        f = np.array(arr, np.int64)
        for ln, s in zip(lens, strides):
            off += (f % ln) * s
            f //= ln
    return off


@pytest.mark.parametrize("seed", range(24))
def test_mult_entries_matches_einsum(seed):
    """orc_mult_entries_d (the checker of the full-size GPU tests) against a
    textbook numpy.einsum of the same strided operands, on random canonical
    forms with batch labels, diagonals in A, and FP32 as well as FP64 data:
    every (l, m, n) entry, mapped to C through the canonical form's C
    strides.  A dropped beta, a swapped stride slot or a transposed operand
    fails here."""
    rng = np.random.default_rng(800 + seed)
    iA, iB, iC, ext = random_contraction(rng, max_dims=2, max_len=4, allow_diag=seed % 4 == 0)
    (lA, sA), (lB, sB), (lC, sC) = [layout_for(i, ext, rng) for i in (iA, iB, iC)]
    dt = np.float32 if seed % 2 else np.float64
    A = rng.uniform(-1, 1, storage_size(lA, sA)).astype(dt)
    B = rng.uniform(-1, 1, storage_size(lB, sB)).astype(dt)
    C0 = rng.uniform(-1, 1, storage_size(lC, sC)).astype(dt)
    alpha, beta = [(0.75, 0.0), (1.0, 1.0), (-0.5, 2.5), (1.5, -1.0)][seed % 4]
    g = po.canonicalize(lA, sA, iA, lB, sB, iB, lC, sC, iC)
    d = g.as_dict()
    L = int(np.prod(d["len_abc"])) if d["len_abc"] else 1
    M = int(np.prod(d["len_ac"])) if d["len_ac"] else 1
    N = int(np.prod(d["len_bc"])) if d["len_bc"] else 1
    el, em, en = (x.ravel() for x in np.meshgrid(np.arange(L), np.arange(M), np.arange(N), indexing="ij"))
    got = po.mult_entries(g, alpha, A, B, beta, C0, el, em, en)
    want_st = C0.astype(np.float64)
    v = _strided_view(want_st, lC, sC)
    prod = np.einsum("%s,%s->%s" % (iA, iB, iC), _strided_view(A.astype(np.float64), lA, sA),
                     _strided_view(B.astype(np.float64), lB, sB))
    v[...] = alpha * prod + beta * v
    want = want_st[_entry_offsets(d, el, em, en)]
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)


def test_mult_entries_beta_zero_and_workload_form():
    """beta == 0 does not read C (NaN there is not propagated), and on the
    permuted workload's canonical form at reduced extents every entry
    equals the einsum."""
    rng = np.random.default_rng(870)
    ext = dict(a=6, b=5, e=4, f=3, i=2, j=3)
    lens = lambda s: [ext[c] for c in s]
    lA, lB, lC = lens("aebf"), lens("fjei"), lens("aibj")
    sA, sB, sC = col_major(lA), col_major(lB), col_major(lC)
    A = rng.uniform(-1, 1, int(np.prod(lA)))
    B = rng.uniform(-1, 1, int(np.prod(lB)))
    C0 = np.full(int(np.prod(lC)), np.nan)
    g = po.canonicalize(lA, sA, "aebf", lB, sB, "fjei", lC, sC, "aibj")
    d = g.as_dict()
    M, N = int(np.prod(d["len_ac"])), int(np.prod(d["len_bc"]))
    em, en = (x.ravel() for x in np.meshgrid(np.arange(M), np.arange(N), indexing="ij"))
    got = po.mult_entries(g, 0.5, A, B, 0.0, C0, np.zeros_like(em), em, en)
    want_st = np.zeros(int(np.prod(lC)))
    _strided_view(want_st, lC, sC)[...] = 0.5 * np.einsum(
        "aebf,fjei->aibj", _strided_view(A, lA, sA), _strided_view(B, lB, sB))
    np.testing.assert_allclose(got, want_st[_entry_offsets(d, np.zeros_like(em), em, en)],
                               rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("seed", range(16))
def test_blis_contract_matches_einsum(seed):
    """orc_blis_contract_d (the CPU baseline: Goto loop nest over scatter
    vectors) against numpy.einsum on random strided contractions, including
    batch labels and sizes that span several MC / KC / NR blocks."""
    rng = np.random.default_rng(900 + seed)
    iA, iB, iC, ext = random_contraction(rng, max_dims=2, max_len=9 if seed < 12 else 4)
    if seed >= 12:  # a long k and a wide n: several KC and MC blocks
        iA, iB, iC = "aek", "kbe", "ab"
        ext = dict(a=75 + seed, b=13 + 3 * seed, e=16, k=33)
    (lA, sA), (lB, sB), (lC, sC) = [layout_for(i, ext, rng) for i in (iA, iB, iC)]
    A = rng.uniform(-1, 1, storage_size(lA, sA))
    B = rng.uniform(-1, 1, storage_size(lB, sB))
    C0 = rng.uniform(-1, 1, storage_size(lC, sC))
    alpha, beta = [(0.75, 0.0), (1.0, 1.0), (-0.5, 2.5), (0.5, 1.0)][seed % 4]
    if beta == 0.0:
        C0[:] = np.nan  # beta == 0 must not read C
    want = np.where(np.isnan(C0), 0.0, C0)
    v = _strided_view(want, lC, sC)
    v[...] = alpha * np.einsum("%s,%s->%s" % (iA, iB, iC), _strided_view(A, lA, sA),
                               _strided_view(B, lB, sB)) + (beta * v if beta else 0.0)
    C = C0.copy()
    flops = po.blis_contract(2, alpha, A, lA, sA, iA, B, lB, sB, iB, beta, C, lC, sC, iC)
    written = np.unique(_entry_offsets_all(lC, sC))
    np.testing.assert_allclose(C[written], want[written], rtol=1e-12, atol=1e-12)
    assert flops == 2.0 * np.prod([ext[c] for c in set(iA + iB + iC)], dtype=np.float64)


def _entry_offsets_all(lens, strides):
    off = np.zeros(1, np.int64)
    for n, s in zip(lens, strides):
        off = (off[:, None] + np.arange(n, dtype=np.int64)[None, :] * s).ravel()
    return off


def test_blis_contract_column_limit():
    """n_limit computes only the first n_limit columns of C (the bench's
    bounded CPU sample): those match the full contraction, the rest of C
    is untouched."""
    rng = np.random.default_rng(950)
    ext = dict(a=20, b=18, e=12, f=7, i=6, j=5)
    lens = lambda s: [ext[c] for c in s]
    lA, lB, lC = lens("aebf"), lens("fjei"), lens("aibj")
    sA, sB, sC = col_major(lA), col_major(lB), col_major(lC)
    A, B = rng.uniform(-1, 1, int(np.prod(lA))), rng.uniform(-1, 1, int(np.prod(lB)))
    C0 = rng.uniform(-1, 1, int(np.prod(lC)))
    full, part = C0.copy(), C0.copy()
    po.blis_contract(3, 0.5, A, lA, sA, "aebf", B, lB, sB, "fjei", 1.0, full, lC, sC, "aibj")
    po.blis_contract(3, 0.5, A, lA, sA, "aebf", B, lB, sB, "fjei", 1.0, part, lC, sC, "aibj", n_limit=7)
    g = po.blis_reorder(po.canonicalize(lA, sA, "aebf", lB, sB, "fjei", lC, sC, "aibj")).as_dict()
    # odometer order, first dimension fastest (the scatter vectors' order)
    cols = np.zeros(1, np.int64)
    for n, s in zip(g["len_n"], g["sC_n"]):
        cols = (np.arange(n, dtype=np.int64)[:, None] * s + cols[None, :]).ravel()
    rows = np.zeros(1, np.int64)
    for n, s in zip(g["len_m"], g["sC_m"]):
        rows = (np.arange(n, dtype=np.int64)[:, None] * s + rows[None, :]).ravel()
    done = (rows[:, None] + cols[None, :7]).ravel()
    rest = (rows[:, None] + cols[None, 7:]).ravel()
    np.testing.assert_allclose(part[done], full[done], rtol=1e-13, atol=1e-13)
    assert (part[rest] == C0[rest]).all()


def test_batch_workload_generator():
    """The batch workload's extents: std::mt19937_64 (the C++ standard's
    known answer: the 10000th output for the default seed 5489 is
    9981545732273789042, [rand.predef]) and libstdc++'s
    uniform_int_distribution<int>(4, 40), whose first draws for seed 7 were
    taken from a g++ 11 build of that call."""
    import bench
    g = bench.Mt19937_64(5489)
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042
    e = bench.batch_extents(4096, 7)
    assert e[:3].tolist() == [[31, 39, 8, 37, 9, 6], [34, 37, 13, 30, 31, 26], [18, 15, 34, 15, 40, 40]]
    assert e.min() == 4 and e.max() == 40
    assert float((2.0 * np.prod(e.astype(np.float64), axis=1)).sum()) == 934466025590.0


# ---------------------------------------------------------------- DPD contraction
def _dpd_problem(rng, nirrep, ndims=(2, 2, 2), maxlen=3, layout=5, irreps=None):
    iA, iB, iC = "abef", "efij", "abij"
    if ndims != (2, 2, 2):
        iA, iB, iC = "aef", "efi", "ai"
    labels = sorted(set(iA + iB + iC))
    ext = {c: [int(rng.integers(0, maxlen + 1)) for _ in range(nirrep)] for c in labels}
    irA, irB = (irreps or (int(rng.integers(0, nirrep)), int(rng.integers(0, nirrep))))[:2]
    irC = irA ^ irB
    dA = po.dpd_init(irA, nirrep, [ext[c] for c in iA], layout)
    dB = po.dpd_init(irB, nirrep, [ext[c] for c in iB], layout)
    dC = po.dpd_init(irC, nirrep, [ext[c] for c in iC], layout)
    return iA, iB, iC, ext, dA, dB, dC


@pytest.mark.parametrize("seed", range(16))
def test_dpd_mult_blockwise_matches_full_expansion(seed):
    rng = np.random.default_rng(600 + seed)
    nirrep = [1, 2, 4, 8][seed % 4]
    iA, iB, iC, ext, dA, dB, dC = _dpd_problem(rng, nirrep, layout=seed % 6,
                                               ndims=(2, 2, 2) if seed % 3 else (3, 3, 2))
    A = rng.uniform(-1, 1, max(po.dpd_size(dA), 1))
    B = rng.uniform(-1, 1, max(po.dpd_size(dB), 1))
    C0 = rng.uniform(-1, 1, max(po.dpd_size(dC), 1))
    C1, C2 = C0.copy(), C0.copy()
    po.dpd_mult(0.5, dA, iA, A, dB, iB, B, -1.5, dC, iC, C1, mode=1)
    po.dpd_mult(0.5, dA, iA, A, dB, iB, B, -1.5, dC, iC, C2, mode=2)
    np.testing.assert_allclose(C1, C2, rtol=1e-12, atol=1e-12)


def _dpd_dense(d, data, nirrep, lens):
    """Dense expansion of a DPD tensor: every symmetry-allowed block (from
    dpd_get_block) placed at its irreps' offsets, zeros elsewhere."""
    import itertools
    starts = [np.concatenate([[0], np.cumsum(l)]) for l in lens]
    out = np.zeros([int(sum(l)) for l in lens], np.asarray(data).dtype)
    for irr in itertools.product(range(nirrep), repeat=len(lens)):
        if np.bitwise_xor.reduce(np.array(irr, np.int64)) != d.irrep:
            continue
        bl, off, st = po.dpd_get_block(d, list(irr))
        if any(x == 0 for x in bl):
            continue
        blk = np.lib.stride_tricks.as_strided(np.asarray(data)[off:], bl,
                                              [x * np.asarray(data).itemsize for x in st])
        out[tuple(slice(starts[i][irr[i]], starts[i][irr[i]] + bl[i]) for i in range(len(lens)))] = blk
    return out


@pytest.mark.parametrize("seed", range(8))
def test_complex_dpd_mult_matches_dense_expansion(seed):
    """complex_dpd_mult (four real block-wise contractions) against numpy's
    complex einsum over the dense expansions of A and B, read back at C's
    blocks; conjugated operands and complex alpha, beta."""
    rng = np.random.default_rng(650 + seed)
    nirrep = [1, 2, 4, 8][seed % 4]
    iA, iB, iC, ext, dA, dB, dC = _dpd_problem(rng, nirrep, layout=seed % 6,
                                               ndims=(2, 2, 2) if seed % 3 else (3, 3, 2))
    cx = lambda n: rng.uniform(-1, 1, max(n, 1)) + 1j * rng.uniform(-1, 1, max(n, 1))
    A, B, C0 = cx(po.dpd_size(dA)), cx(po.dpd_size(dB)), cx(po.dpd_size(dC))
    cA, cB = bool(seed & 1), bool(seed & 2)
    alpha, beta = 0.5 - 0.25j, 1.5 + 0.75j
    got = po.complex_dpd_mult(alpha, dA, iA, A, dB, iB, B, beta, dC, iC, C0, cA, cB)
    lens = lambda idx: [ext[c] for c in idx]
    DA = _dpd_dense(dA, A, nirrep, lens(iA))
    DB = _dpd_dense(dB, B, nirrep, lens(iB))
    DA, DB = (np.conj(DA) if cA else DA), (np.conj(DB) if cB else DB)
    P = np.einsum("%s,%s->%s" % (iA, iB, iC), DA, DB)
    # C's storage positions of the dense product: expand a running index
    pos = _dpd_dense(dC, np.arange(len(C0), dtype=np.float64), nirrep, lens(iC))
    mask = _dpd_dense(dC, np.ones(len(C0)), nirrep, lens(iC)) > 0
    n = po.dpd_size(dC)
    want = C0.copy()
    want[:n] *= beta  # an empty C leaves its placeholder storage alone
    if n > 0:
        want[pos[mask].astype(np.int64)] += alpha * P[mask]
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("seed", range(8))
def test_dpd_patch_scatter_addresses(seed):
    """Every patch of the irrep_AB GEMM addresses exactly the elements of
    the allowed blocks with that column irrep."""
    rng = np.random.default_rng(700 + seed)
    nirrep = [2, 4, 8][seed % 3]
    iA, iB, iC, ext, dA, dB, dC = _dpd_problem(rng, nirrep, layout=5, irreps=(0, 0))
    for r in range(nirrep):
        p = po.dpd_patch_scatter(dA, iA, dB, iB, dC, iC, 0, r, 4, 2)
        addr = []
        ro = co = 0
        for n in range(len(p["rows"])):
            rs = p["rscat"][ro:ro + p["rows"][n]]
            cs = p["cscat"][co:co + p["cols"][n]]
            addr += (p["base"][n] + rs[:, None] + cs[None, :]).ravel().tolist()
            ro += p["rows"][n]
            co += p["cols"][n]
        want = []
        for combo in itertools.product(range(nirrep), repeat=4):
            if combo[0] ^ combo[1] ^ combo[2] ^ combo[3]:
                continue
            if combo[2] ^ combo[3] != r:
                continue
            ln, off, st = po.dpd_get_block(dA, list(combo))
            if 0 in ln:
                continue
            for idx in itertools.product(*[range(l) for l in ln]):
                want.append(off + sum(i * s for i, s in zip(idx, st)))
        assert sorted(addr) == sorted(want)


# ---------------------------------------------------------------- level-1
def _l1(mode, alpha, A, lA, sA, iA, beta, B, lB, sB, iB):
    return po.level1(mode, alpha, A, lA, sA, iA, beta, B, lB, sB, iB)


def _colmaj(l):
    s, a = [], 1
    for x in l:
        s.append(a)
        a *= x
    return s


def test_level1_add_against_einsum():
    """The level-1 oracle against numpy on transposes, sums over labels only
    A has, replication over labels only B has, and beta = 0 with NaN in B."""
    rng = np.random.default_rng(60)
    cases = [("abc", "cab", dict(a=3, b=4, c=5)), ("abc", "ca", dict(a=3, b=4, c=5)),
             ("ab", "abc", dict(a=3, b=4, c=2)), ("ab", "ba", dict(a=6, b=7))]
    for iA, iB, ext in cases:
        lA, lB = [ext[c] for c in iA], [ext[c] for c in iB]
        A = rng.uniform(-1, 1, int(np.prod(lA)))
        B = rng.uniform(-1, 1, int(np.prod(lB)))
        Ar = A.reshape(lA[::-1]).transpose(range(len(lA))[::-1])  # column-major view
        Br = B.reshape(lB[::-1]).transpose(range(len(lB))[::-1])
        a_only = "".join(c for c in iA if c not in iB)
        red = np.einsum(iA + "->" + "".join(c for c in iA if c in iB), Ar)
        kept = "".join(c for c in iA if c in iB)
        # broadcast the reduced A over B's labels
        want = 0.5 * np.einsum(kept + "->" + kept, red)
        want = np.broadcast_to(np.einsum(kept + "->" + "".join(c for c in iB if c in kept), want)
                               .reshape([ext[c] if c in kept else 1 for c in iB]), lB) + 2.0 * Br
        got = _l1(0, 0.5, A, lA, _colmaj(lA), iA, 2.0, B, lB, _colmaj(lB), iB)
        gotr = got.reshape(lB[::-1]).transpose(range(len(lB))[::-1])
        np.testing.assert_allclose(gotr, want, rtol=1e-13, atol=1e-13)
    # beta == 0 overwrites (NaN in B does not propagate)
    B = np.full(6, np.nan)
    A = rng.uniform(-1, 1, 6)
    got = _l1(0, 1.0, A, [2, 3], [1, 2], "ab", 0.0, B, [2, 3], [1, 2], "ab")
    np.testing.assert_array_equal(got, A)


def test_level1_diagonals_scale_set_and_unit_lengths():
    rng = np.random.default_rng(61)
    # B[aa] = A[a] + B[aa]: only the diagonal changes
    A = rng.uniform(-1, 1, 4)
    B = rng.uniform(-1, 1, 16)
    got = _l1(0, 1.0, A, [4], [1], "a", 1.0, B, [4, 4], [1, 4], "aa")
    want = B.copy()
    want[[0, 5, 10, 15]] += A
    np.testing.assert_allclose(got, want, rtol=1e-15)
    # A[aab] summed to B[ab]: diagonal of A
    A = rng.uniform(-1, 1, 3 * 3 * 2)
    B = np.zeros(6)
    got = _l1(0, 1.0, A, [3, 3, 2], [1, 3, 9], "aab", 0.0, B, [3, 2], [1, 3], "ab")
    Ar = A.reshape(2, 3, 3).transpose(2, 1, 0)
    np.testing.assert_allclose(got.reshape(2, 3).T, np.einsum("aab->ab", Ar), rtol=1e-15)
    # a label of length 1 in A only is dropped: A[ax] with x=1 equals A[a]
    A = rng.uniform(-1, 1, 5)
    B = np.zeros(5)
    got = _l1(0, 2.0, A, [5, 1], [1, 5], "ax", 0.0, B, [5], [1], "a")
    np.testing.assert_allclose(got, 2.0 * A, rtol=1e-15)
    # scale and set over a diagonal
    B = rng.uniform(-1, 1, 9)
    got = _l1(1, 0.0, None, [], [], "", -3.0, B, [3, 3], [1, 3], "aa")
    want = B.copy()
    want[[0, 4, 8]] *= -3.0
    np.testing.assert_allclose(got, want, rtol=1e-15)
    got = _l1(2, 7.0, None, [], [], "", 0.0, B, [3, 3], [1, 3], "ab")
    np.testing.assert_array_equal(got, np.full(9, 7.0))


# ---------------------------------------------------------------- indexed
INDEXED_CASES = [
    # (idx_A, dense_A, idx_B, dense_B, idx_C, dense_C)
    ("abef", 4, "efij", 4, "abij", 4),   # all dense: one entry each
    ("abef", 3, "eijf", 3, "abij", 4),   # f indexed in A and in B
    ("abef", 3, "efij", 4, "abij", 4),   # f indexed in A, dense in B (B sliced)
    ("abef", 4, "efij", 4, "abij", 3),   # j indexed in C only (A, B dense)
    ("aek", 3, "ebk", 3, "abk", 3),      # k dense everywhere: a batch dimension
    ("aek", 2, "ebk", 2, "abk", 2),      # k indexed everywhere
    ("aebk", 2, "ekb", 1, "abk", 1),     # mixed: b dense in A, indexed in B and C; k indexed in all
]


@pytest.mark.parametrize("case", range(len(INDEXED_CASES)))
def test_indexed_oracle_against_brute_force(case):
    """The oracle's FULL algorithm against the definition, entry triple by
    entry triple (the reference's own indexed test compares its BLOCKED and
    FULL algorithms, test/3t/mult.cxx:143-174)."""
    from tests.helpers import indexed_brute_force, random_indexed
    iA, nA, iB, nB, iC, nC = INDEXED_CASES[case]
    rng = np.random.default_rng(62000 + case)
    for trial in range(3):
        ext = {c: int(rng.integers(1, 4)) for c in sorted(set(iA + iB + iC))}
        A = random_indexed(rng, iA, ext, nA)
        B = random_indexed(rng, iB, ext, nB)
        C0 = random_indexed(rng, iC, ext, nC)
        alpha, beta = [(0.75, 0.5), (1.0, 0.0), (-1.5, 1.0)][trial]
        got = po.indexed_mult(alpha, A, iA, B, iB, beta, C0, iC)
        want = indexed_brute_force(alpha, A, iA, B, iB, beta, C0, iC)
        assert len(got) == len(want)
        for g, w in zip(got, want):
            np.testing.assert_allclose(g, w, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("case", range(4))
def test_indexed_oracle_complex_matches_brute_force(case):
    """The indexed oracle with complex entries, factors and scalars and
    conjugated operands against the per-triple definition."""
    from tests.helpers import indexed_brute_force, random_indexed
    iA, nA, iB, nB, iC, nC = INDEXED_CASES[case]
    rng = np.random.default_rng(62500 + case)

    def cplx(X):
        X["entries"] = [e + 1j * rng.uniform(-1, 1, np.shape(e)) for e in X["entries"]]
        X["factors"] = [f * (0.5 + 1j * rng.uniform(-1, 1)) for f in X["factors"]]
        return X

    for trial in range(3):
        ext = {c: int(rng.integers(1, 4)) for c in sorted(set(iA + iB + iC))}
        A = cplx(random_indexed(rng, iA, ext, nA))
        B = cplx(random_indexed(rng, iB, ext, nB))
        C0 = cplx(random_indexed(rng, iC, ext, nC))
        alpha, beta = [(0.75 - 0.5j, 0.5 + 1j), (1.0, 0.0), (1j, 1.0)][trial]
        cA, cB = bool(trial & 1), trial == 2
        got = po.indexed_mult(alpha, A, iA, B, iB, beta, C0, iC, cA, cB)
        want = indexed_brute_force(alpha, A, iA, B, iB, beta, C0, iC, cA, cB)
        assert len(got) == len(want)
        for g, w in zip(got, want):
            assert np.iscomplexobj(g)
            np.testing.assert_allclose(g, w, rtol=1e-12, atol=1e-12)


def test_indexed_oracle_fully_populated_is_dense_contraction():
    """Every index value present with factor 1: the indexed contraction is
    the dense one (checked against the dense oracle, oracle.c)."""
    from tests.helpers import random_indexed
    rng = np.random.default_rng(62100)
    ext = dict(a=3, b=2, e=4, f=3, i=2, j=3)
    A = random_indexed(rng, "abef", ext, 2, fill=1.0, factors=(1.0,))
    B = random_indexed(rng, "efij", ext, 3, fill=1.0, factors=(1.0,))
    C0 = random_indexed(rng, "abij", ext, 1, fill=1.0, factors=(1.0,))
    got = po.indexed_mult(0.5, A, "abef", B, "efij", -1.0, C0, "abij")
    full = {k: po.indexed_full(X["entries"], X["indices"], X["factors"], X["dense_len"], X["idx_len"])
            for k, X in (("A", A), ("B", B), ("C", C0))}
    lens = {k: list(v.shape) for k, v in full.items()}
    store = {k: np.asfortranarray(v).ravel(order="F") for k, v in full.items()}
    cm = lambda l: [int(np.prod(l[:i])) for i in range(len(l))]
    want = po.tensor_mult(0.5, store["A"], lens["A"], cm(lens["A"]), "abef", store["B"], lens["B"],
                          cm(lens["B"]), "efij", -1.0, store["C"].copy(), lens["C"], cm(lens["C"]), "abij")
    want = np.asarray(want).reshape(lens["C"], order="F")
    for e, ix in zip(got, C0["indices"]):
        np.testing.assert_allclose(e, want[(slice(None),) + tuple(ix)], rtol=1e-13, atol=1e-13)


# ---------------------------------------------------------------- complex
@pytest.mark.parametrize("seed", range(6))
def test_complex_oracle_against_numpy(seed):
    """The complex restatement (four real oracle contractions combined)
    against numpy's complex einsum, with conjugation and complex scalars."""
    from tests.helpers import layout_for, random_contraction, storage_size
    rng = np.random.default_rng(64000 + seed)
    iA, iB, iC, ext = random_contraction(rng, allow_diag=False)
    (lA, sA), (lB, sB), (lC, sC) = [layout_for(i, ext, rng, "random") for i in (iA, iB, iC)]
    mk = lambda l, s: rng.uniform(-1, 1, storage_size(l, s)) + 1j * rng.uniform(-1, 1, storage_size(l, s))
    A, B, C0 = mk(lA, sA), mk(lB, sB), mk(lC, sC)
    alpha, beta = [(0.5 - 0.25j, 1.5 + 0.5j), (2.0, 0.0), (1j, -1.0)][seed % 3]
    cA, cB = bool(seed & 1), bool(seed & 2)
    got = po.complex_tensor_mult(alpha, A, lA, sA, iA, B, lB, sB, iB, beta, C0, lC, sC, iC, cA, cB)
    view = lambda X, l, s: np.lib.stride_tricks.as_strided(X, l, [16 * x for x in s])
    a, b = view(A, lA, sA), view(B, lB, sB)
    a = np.conj(a) if cA else a
    b = np.conj(b) if cB else b
    want_v = alpha * np.einsum("%s,%s->%s" % (iA, iB, iC), a, b) + beta * view(C0.copy(), lC, sC)
    np.testing.assert_allclose(view(got, lC, sC), want_v, rtol=1e-12, atol=1e-12)


def test_complex_oracle_diagonal_c():
    """C[aa]: only C's diagonal is written (alpha P + beta C there); the
    off-diagonal entries keep their values whatever alpha and beta are."""
    rng = np.random.default_rng(64100)
    A = rng.uniform(-1, 1, 30) + 1j * rng.uniform(-1, 1, 30)
    B = rng.uniform(-1, 1, 30) + 1j * rng.uniform(-1, 1, 30)
    C0 = rng.uniform(-1, 1, 36) + 1j * rng.uniform(-1, 1, 36)
    for alpha, beta in [(1 + 1j, 0.0), (0.0, 0.5 + 0.5j), (2.0, 0.5 - 1j)]:
        got = po.complex_tensor_mult(alpha, A, [6, 5], [1, 6], "ae", B, [5, 6], [1, 5], "ea", beta, C0,
                                     [6, 6], [1, 6], "aa")
        want = C0.copy()
        d = np.arange(6) * 7
        want[d] = alpha * np.einsum("ae,ea->a", A.reshape(5, 6).T, B.reshape(6, 5).T) + beta * C0[d]
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)


# ---------------------------------------------------------------- indexed DPD
def test_indexed_dpd_oracle_one_irrep_is_indexed():
    """With one irrep the DPD parts are plain dense arrays (row-major for the
    PREFIX_ROW_MAJOR layout): the indexed DPD oracle equals the indexed one."""
    from tests.helpers import random_indexed_dpd
    rng = np.random.default_rng(66000)
    for trial in range(4):
        ext = {c: [int(rng.integers(1, 4))] for c in "abefij"}
        A = random_indexed_dpd(rng, "abef", ext, 3, 0, 1)
        B = random_indexed_dpd(rng, "eijf", ext, 3, 0, 1)
        C0 = random_indexed_dpd(rng, "abij", ext, 2, 0, 1)
        got = po.indexed_dpd_mult(0.75, A, "abef", B, "eijf", 0.5, C0, "abij")

        def plain(X):
            nd = X["dense_ndim"]
            dl = [l[0] for l in X["lens"][:nd]]
            return dict(entries=[np.asarray(e).reshape(dl) for e in X["entries"]], indices=X["indices"],
                        factors=X["factors"], dense_len=dl, idx_len=[l[0] for l in X["lens"][nd:]])
        want = po.indexed_mult(0.75, plain(A), "abef", plain(B), "eijf", 0.5, plain(C0), "abij")
        for g, w in zip(got, want):
            np.testing.assert_allclose(g, w.ravel(), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("layout", [0, 3, 5])
def test_indexed_dpd_oracle_no_indices_is_dpd(layout):
    """Without indexed dimensions (one entry, factor 1) the indexed DPD oracle
    equals the DPD oracle (oracle.c, pinned by the MArray layout table)."""
    from tests.helpers import random_indexed_dpd
    rng = np.random.default_rng(66100 + layout)
    nir = 4
    ext = {c: [int(rng.integers(0, 4)) for _ in range(nir)] for c in "abef"}
    A = random_indexed_dpd(rng, "abe", ext, 3, 1, nir, layout, factors=(1.0,))
    B = random_indexed_dpd(rng, "ef", ext, 2, 2, nir, layout, factors=(1.0,))
    C0 = random_indexed_dpd(rng, "abf", ext, 3, 3, nir, layout, factors=(1.0,))
    got = po.indexed_dpd_mult(1.5, A, "abe", B, "ef", -0.5, C0, "abf")
    dA = po.dpd_init(1, nir, A["lens"], layout)
    dB = po.dpd_init(2, nir, B["lens"], layout)
    dC = po.dpd_init(3, nir, C0["lens"], layout)
    want = po.dpd_mult(1.5, dA, "abe", A["entries"][0], dB, "ef", B["entries"][0], -0.5, dC, "abf",
                       np.array(C0["entries"][0]))
    np.testing.assert_allclose(got[0], want, rtol=1e-13, atol=1e-13)


# ------------------------------------------------------------ FP64-accurate mode
def test_ozaki_digits_reconstruct():
    """Digits rebuild each value to within 2^-7S of its row scale, exactly
    when the value needs no more bits than the digits hold."""
    rng = np.random.default_rng(900)
    X = rng.uniform(-1, 1, (20, 50)) * 10.0 ** rng.uniform(-5, 5, (20, 1))
    for S in (6, 7, 8):
        D, e = po.ozaki_digits(X, S)
        assert np.all(np.abs(D) <= 127)
        w = np.array([2.0 ** (-7 * (t + 1)) for t in range(S)])
        rec = np.ldexp(np.tensordot(w, D.astype(np.float64), axes=1), e[:, None])
        scale = np.ldexp(1.0, e)[:, None]
        assert np.all(np.abs(rec - X) < 2.0 ** (-7 * S) * scale)
        assert np.all(np.abs(rec) <= np.abs(X))  # truncation toward zero
    # 10-bit integers need two digits below a 2^10 row scale: exact
    Y = rng.integers(-1023, 1024, (8, 30)).astype(np.float64)
    D, e = po.ozaki_digits(Y, 6)
    rec = np.ldexp(np.tensordot([2.0 ** (-7 * (t + 1)) for t in range(6)], D.astype(np.float64), axes=1), e[:, None])
    assert np.array_equal(rec, Y)
    assert list(po.ozaki_exponents(np.array([[0.0, 0.0], [0.5, -1.0], [3.0, 1.0]]))) == [0, 1, 2]


def test_ozaki_gemm_exact_on_small_integers():
    """Integer matrices whose digits are exact and whose pairs all survive the
    truncation: the scheme returns the exact product."""
    rng = np.random.default_rng(901)
    A = rng.integers(-100, 101, (13, 40)).astype(np.float64)
    B = rng.integers(-100, 101, (40, 9)).astype(np.float64)
    exact = np.array([[sum(int(A[i, k]) * int(B[k, j]) for k in range(40)) for j in range(9)]
                      for i in range(13)], dtype=np.float64)
    for S in (6, 7, 8):
        assert np.array_equal(po.ozaki_gemm(A, B, S), exact)


def test_ozaki_gemm_error_and_flush():
    """Within the scheme's bound of float64 matmul on uniform data, and the
    blocked integer sums (flush_k) change nothing when no block overflows."""
    rng = np.random.default_rng(902)
    A = rng.uniform(-1, 1, (30, 700))
    B = rng.uniform(-1, 1, (700, 25))
    ref = A @ B
    err = np.linalg.norm(po.ozaki_gemm(A, B, 8) - ref) / np.linalg.norm(ref)
    assert err < 1e-14
    assert np.linalg.norm(po.ozaki_gemm(A, B, 6) - ref) / np.linalg.norm(ref) < 1e-11
    a = po.ozaki_gemm(A, B, 7, flush_k=128)
    b = po.ozaki_gemm(A, B, 7)
    assert np.linalg.norm(a - b) <= 1e-15 * np.linalg.norm(ref)
