"""GPU parity of indexed DPD contractions (tblis_indexed_dpd_tensor_mult; the
reference's tblis::mult over indexed_dpd_varray_view,
src/iface/3t/mult.cxx:413-540), run by the blocked algorithm (the default)
and by the FULL one, against the oracle's FULL algorithm
(oracle/pyoracle.py indexed_dpd_mult, pinned against the DPD and the
indexed oracles): 1, 2, 4 and 8 irreps, every layout, labels indexed in one
or both operands or in C only, nonzero tensor irreps, zero-length irreps,
the beta rules and alpha = 0.  FP64 within 1e-12 relative Frobenius error
over all C entries, FP32 (split-BF16) within 1e-6."""
import numpy as np
import pytest
import torch

import tblis_amd as T
from oracle import pyoracle as po
from tests.helpers import random_indexed_dpd

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
TOL = {np.float64: 1e-12, np.float32: 1e-6, np.complex128: 1e-12, np.complex64: 1e-6}

CASES = [
    # (idx_A, dense_A, idx_B, dense_B, idx_C, dense_C)
    ("abef", 4, "efij", 4, "abij", 4),   # no indexed dimensions: plain DPD
    ("abef", 3, "eijf", 3, "abij", 4),   # f indexed in A and B
    ("abef", 3, "efij", 4, "abij", 4),   # f indexed in A only (B sliced at its irrep and value)
    ("abef", 4, "efij", 4, "abij", 3),   # j indexed in C only
    ("aek", 2, "ebk", 2, "abk", 2),      # k indexed everywhere
    ("aek", 3, "ebk", 3, "abk", 3),      # k a dense batch label (in all three)
    ("akef", 3, "ebkf", 3, "abk", 3),    # dense batch label k, f indexed in A and B
]


def _device(X, dtype, separate):
    ents = []
    if separate:
        ents = [torch.from_numpy(np.asarray(e, dtype)).to(DEV) for e in X["entries"]]
        keep = ents
    else:
        n = len(X["entries"][0]) if X["entries"] else 1
        store = torch.zeros(max(len(X["entries"]), 1) * n, dtype=_torch(dtype), device=DEV)
        for i, e in enumerate(X["entries"]):
            store[i * n:(i + 1) * n] = torch.from_numpy(np.asarray(e, dtype)).to(DEV)
            ents.append(store[i * n:(i + 1) * n])
        keep = [store]
    return T.IndexedDpd(ents, X["indices"], X["idx_irrep"], X["irrep"], X["nirrep"], X["lens"],
                        X["dense_ndim"], X["factors"], X["layout"], dtype=_torch(dtype)), keep


def _torch(dtype):
    return {np.float64: torch.float64, np.float32: torch.float32, np.complex128: torch.complex128,
            np.complex64: torch.complex64}[dtype]


def _run(case, dtype, seed, alpha, beta, nirrep, layout, separate=False, maxlen=4, conj=(False, False)):
    iA, nA, iB, nB, iC, nC = case
    rng = np.random.default_rng(seed)
    labels = set(iA + iB + iC)
    lens = {c: [int(rng.integers(0, maxlen)) for _ in range(nirrep)] for c in labels}
    for c in labels:
        if sum(lens[c]) == 0:
            lens[c][0] = 1
    irA, irB = int(rng.integers(nirrep)), int(rng.integers(nirrep))
    A = random_indexed_dpd(rng, iA, lens, nA, irA, nirrep, layout)
    B = random_indexed_dpd(rng, iB, lens, nB, irB, nirrep, layout)
    C0 = random_indexed_dpd(rng, iC, lens, nC, irA ^ irB, nirrep, layout)
    cplx = np.iscomplexobj(np.zeros(1, dtype))
    for X in (A, B, C0):
        if cplx:
            X["entries"] = [np.asarray(e) + 1j * rng.uniform(-1, 1, np.shape(e)) for e in X["entries"]]
            X["factors"] = [f * complex(1.0, float(rng.uniform(-1, 1))) for f in X["factors"]]
        X["entries"] = [np.asarray(e, dtype) for e in X["entries"]]
    want = po.indexed_dpd_mult(alpha, A, iA, B, iB, beta, C0, iC, conj[0], conj[1])
    dA, kA = _device(A, dtype, separate)
    dB, kB = _device(B, dtype, separate)
    dC, kC = _device(C0, dtype, separate)
    T.indexed_dpd_mult(alpha, dA, iA, dB, iB, beta, dC, iC, conj_A=conj[0], conj_B=conj[1])
    torch.cuda.synchronize()
    if not want:
        return
    got = np.concatenate([e.cpu().to(torch.complex128 if cplx else torch.float64).numpy().ravel()
                          for e in dC.entries])
    ref = np.concatenate([w.ravel() for w in want])
    nrm = np.linalg.norm(ref)
    err = np.linalg.norm(got - ref) / (nrm if nrm > 0 else 1.0)
    assert err <= TOL[dtype], (case, dtype, seed, nirrep, layout, alpha, beta, err)


@pytest.mark.parametrize("ci", range(len(CASES)))
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_indexed_dpd_cases(ci, dtype):
    for t, (nirrep, layout, alpha, beta) in enumerate([(1, 5, 1.0, 0.0), (2, 0, 0.75, 0.5),
                                                        (4, 3, -1.25, 1.0), (8, 5, 1.0, -0.5)]):
        _run(CASES[ci], dtype, 67000 + 10 * ci + t, alpha, beta, nirrep, layout, separate=(t % 2 == 1))


def test_indexed_dpd_alpha_zero_and_layouts():
    for layout in range(6):
        _run(CASES[1], np.float64, 67500 + layout, 1.0, 0.25, 4, layout)
    _run(CASES[1], np.float64, 67510, 0.0, 0.5, 2, 5)     # only the beta rule
    _run(CASES[3], np.float32, 67511, 0.0, 0.0, 2, 5)     # C set to zero


@pytest.mark.parametrize("ci", range(len(CASES)))
@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_indexed_dpd_complex(ci, dtype):
    """Complex indexed DPD contractions (the reference instantiates them for
    its complex types, src/iface/3t/mult.cxx:539-543): complex entries and
    factors, real and complex alpha / beta, conjugated operands, 1 to 8
    irreps, against the oracle's FULL algorithm in complex arithmetic."""
    for t, (nirrep, layout, alpha, beta, conj) in enumerate([
            (1, 5, 1.0, 0.0, (False, False)), (2, 0, 0.75 - 0.5j, 0.5 + 1j, (True, False)),
            (4, 3, 1j, 1.0, (False, True)), (8, 5, -1.0, 0.25, (True, True)),
            (2, 5, 0.0, 0.5 - 0.5j, (False, False))]):
        _run(CASES[ci], dtype, 67700 + 10 * ci + t, alpha, beta, nirrep, layout, separate=(t % 2 == 1),
             conj=conj)


def test_indexed_dpd_full_algorithm(monkeypatch):
    """The FULL algorithm (TBLIS_HIP_INDEXED_DPD=full, the reference's
    dpd_impl == FULL, src/internal/3t/indexed_dpd/mult.cxx:1282-1287) on
    every case, real and complex; the tests above run the blocked one."""
    monkeypatch.setenv("TBLIS_HIP_INDEXED_DPD", "full")
    for ci in range(len(CASES)):
        _run(CASES[ci], np.float64, 67800 + ci, 0.75, 0.5, 4, 5, separate=True)
        _run(CASES[ci], np.float32, 67820 + ci, 1.0, 0.0, 8, 0)
        _run(CASES[ci], np.complex128, 67840 + ci, 0.5 - 0.25j, 1.0, 2, 3, conj=(True, False))
