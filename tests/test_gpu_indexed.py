"""GPU parity of indexed contractions (tblis_indexed_tensor_mult; the
reference's tblis::mult over indexed_varray_view, src/iface/3t/mult.cxx:
299-402) against the oracle's FULL algorithm (oracle/pyoracle.py
indexed_mult, itself pinned against the per-triple definition): labels
indexed in both operands, in one operand only (the other sliced), in C only,
dense batch labels, indexed batch labels, zero factors, the beta rules and
alpha = 0.  FP64 within 1e-12 relative Frobenius error over all C entries,
FP32 (split-BF16) within 1e-6."""
import numpy as np
import pytest
import torch

import tblis_amd as T
from oracle import pyoracle as po
from tests.helpers import random_indexed

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
TOL = {np.float64: 1e-12, np.float32: 1e-6, np.complex128: 1e-12, np.complex64: 1e-6}

CASES = [
    # (idx_A, dense_A, idx_B, dense_B, idx_C, dense_C)
    ("abef", 4, "efij", 4, "abij", 4),   # all dense: one entry each
    ("abef", 3, "eijf", 3, "abij", 4),   # f indexed in A and in B
    ("abef", 3, "efij", 4, "abij", 4),   # f indexed in A, dense in B (B sliced)
    ("abef", 4, "efij", 4, "abij", 3),   # j indexed in C only
    ("aek", 3, "ebk", 3, "abk", 3),      # k dense everywhere: batch
    ("aek", 2, "ebk", 2, "abk", 2),      # k indexed everywhere
    ("aebk", 2, "ekb", 1, "abk", 1),     # mixed: b dense in A, indexed in B and C
    ("aebf", 2, "fjei", 2, "aibj", 2),   # permuted dense parts, two indexed labels each
]


def _device(X, dtype, layout, rng, separate):
    """Entries of a host description on the device: one allocation with the
    entries one after another (as MArray lays them out), or one allocation
    per entry; layout "col" or "row" for the dense part."""
    dl = X["dense_len"]
    n = int(np.prod(dl)) if dl else 1
    if layout == "col":
        st = [int(np.prod(dl[:i])) for i in range(len(dl))]
    else:
        st = [int(np.prod(dl[i + 1:])) for i in range(len(dl))]
    ents = []
    if separate:
        for e in X["entries"]:
            flat = np.asarray(e, dtype).ravel(order="F" if layout == "col" else "C")
            d = torch.from_numpy(flat.copy()).to(DEV)
            ents.append(torch.as_strided(d, dl, st))
        keep = ents
    else:
        pad = int(rng.integers(0, 3))
        store = torch.zeros(max(len(X["entries"]), 1) * (n + pad), dtype=_torch(dtype), device=DEV)
        for i, e in enumerate(X["entries"]):
            flat = np.asarray(e, dtype).ravel(order="F" if layout == "col" else "C")
            store[i * (n + pad): i * (n + pad) + n] = torch.from_numpy(flat.copy()).to(DEV)
            ents.append(torch.as_strided(store, dl, st, i * (n + pad)))
        keep = [store]
    return T.Indexed(ents, X["indices"], X["idx_len"], X["factors"], dtype=_torch(dtype)), keep


def _torch(dtype):
    return {np.float64: torch.float64, np.float32: torch.float32, np.complex128: torch.complex128,
            np.complex64: torch.complex64}[dtype]


FACTORS = (1.0, 0.5, -0.75, 2.0)


def _run(case, dtype, seed, alpha, beta, ext_range=(2, 9), fill=0.8, separate=False, factors=FACTORS,
         config=None):
    """One contraction on the device against the oracle; returns whether any
    entry triple contributed (so that a case cannot pass vacuously)."""
    iA, nA, iB, nB, iC, nC = case
    rng = np.random.default_rng(seed)
    ext = {c: int(rng.integers(*ext_range)) for c in sorted(set(iA + iB + iC))}
    # indexed dimensions short, so that every case has a few entries
    for idx, nd in ((iA, nA), (iB, nB), (iC, nC)):
        for c in idx[nd:]:
            ext[c] = min(ext[c], 4)
    A = random_indexed(rng, iA, ext, nA, fill, factors)
    B = random_indexed(rng, iB, ext, nB, fill, factors)
    C0 = random_indexed(rng, iC, ext, nC, fill, factors)
    for X in (A, B, C0):
        X["entries"] = [np.asarray(e, dtype) for e in X["entries"]]
    want = po.indexed_mult(alpha, A, iA, B, iB, beta, C0, iC)
    zero = dict(C0, entries=[np.zeros_like(e) for e in C0["entries"]])
    worked = any(np.any(w != 0) for w in po.indexed_mult(1.0, A, iA, B, iB, 1.0, zero, iC))
    layouts = ["col", "row"]
    dA, kA = _device(A, dtype, layouts[seed % 2], rng, separate)
    dB, kB = _device(B, dtype, layouts[(seed + 1) % 2], rng, separate)
    dC, kC = _device(C0, dtype, "col", rng, separate)
    T.indexed_mult(alpha, dA, iA, dB, iB, beta, dC, iC, config=config)
    torch.cuda.synchronize()
    if not want:
        return worked
    got = np.concatenate([e.cpu().double().numpy().ravel() for e in dC.entries])
    ref = np.concatenate([w.ravel() for w in want])
    nrm = np.linalg.norm(ref)
    err = np.linalg.norm(got - ref) / (nrm if nrm > 0 else 1.0)
    assert err <= TOL[dtype], (case, dtype, seed, alpha, beta, err)
    return worked


@pytest.mark.parametrize("ci", range(len(CASES)))
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_indexed_cases(ci, dtype):
    worked = [_run(CASES[ci], dtype, 63000 + 10 * ci + t, alpha, beta, separate=(t == 1))
              for t, (alpha, beta) in enumerate([(1.0, 0.0), (0.75, 0.5), (-1.25, 1.0)])]
    assert any(worked)


def test_indexed_alpha_zero_and_zero_factors():
    case = CASES[1]
    for dtype in (np.float64, np.float32):
        _run(case, dtype, 63500, 0.0, 0.5)              # only the beta pass
        _run(case, dtype, 63501, 0.0, 0.0)              # C set to zero
        assert _run(case, dtype, 63502, 2.0, 1.0, fill=1.0)  # every index value present
        # factors of 0 in every operand (the reference's random factors)
        _run(case, dtype, 63503, 1.0, 0.5, factors=(1.0, 0.5, 0.0))
        _run(case, dtype, 63504, 1.0, 0.0, factors=(0.0,))


def test_indexed_larger_blocks():
    """Dense parts large enough for several GEMM tiles, many k blocks."""
    case = ("abef", 3, "eijf", 3, "abij", 3)
    for dtype in (np.float64, np.float32):
        assert _run(case, dtype, 63600, 1.0, 0.0, ext_range=(24, 40), fill=0.7)


@pytest.mark.parametrize("ci", [1, 2, 3, 5, 6, 7])
@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_indexed_complex(ci, dtype):
    """Complex indexed contractions (the reference instantiates indexed mult
    for its complex types, src/iface/3t/mult.cxx:405-409): complex entries
    and factors, real and complex alpha / beta, beta = 0 and 1, alpha = 0,
    conjugated operands, against the oracle's FULL algorithm in complex
    arithmetic (pinned against the per-triple definition)."""
    case = CASES[ci]
    iA, nA, iB, nB, iC, nC = case
    for t, (alpha, beta, cA, cB) in enumerate([(1.0, 0.0, False, False), (0.75 - 0.5j, 0.5 + 1j, True, False),
                                               (1j, 1.0, False, True), (0.0, 0.5 - 0.5j, False, False)]):
        rng = np.random.default_rng(64000 + 10 * ci + t)
        ext = {c: int(rng.integers(2, 8)) for c in sorted(set(iA + iB + iC))}
        for idx, nd in ((iA, nA), (iB, nB), (iC, nC)):
            for c in idx[nd:]:
                ext[c] = min(ext[c], 4)
        X = [random_indexed(rng, i, ext, n, 0.8, FACTORS) for i, n in ((iA, nA), (iB, nB), (iC, nC))]
        for Y in X:
            Y["entries"] = [(np.asarray(e) + 1j * rng.uniform(-1, 1, np.shape(e))).astype(dtype) for e in Y["entries"]]
            Y["factors"] = [f * complex(1.0, float(rng.uniform(-1, 1))) for f in Y["factors"]]
        A, B, C0 = X
        want = po.indexed_mult(alpha, A, iA, B, iB, beta, C0, iC, cA, cB)
        dA, kA = _device(A, dtype, "col", rng, False)
        dB, kB = _device(B, dtype, "row", rng, t == 1)
        dC, kC = _device(C0, dtype, "col", rng, False)
        T.indexed_mult(alpha, dA, iA, dB, iB, beta, dC, iC, conj_A=cA, conj_B=cB)
        torch.cuda.synchronize()
        got = np.concatenate([e.cpu().to(torch.complex128).numpy().ravel() for e in dC.entries])
        ref = np.concatenate([w.ravel() for w in want])
        nrm = np.linalg.norm(ref)
        err = np.linalg.norm(got - ref) / (nrm if nrm > 0 else 1.0)
        assert err <= TOL[dtype], (case, dtype, t, alpha, beta, err)
