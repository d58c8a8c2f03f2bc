"""GPU parity of complex contractions (tblis_tensor_mult with TYPE_SCOMPLEX /
TYPE_DCOMPLEX, src/iface/3t/mult.cxx:24-165) against the oracle's complex
restatement (four real oracle contractions, itself pinned against numpy's
complex einsum): random labels and strided layouts, real and complex alpha
and beta, conjugated operands, beta = 0 over NaN, alpha = 0, empty k.
complex128 within 1e-12 relative Frobenius error (FP64 MFMA), complex64
within 1e-6 (split-BF16)."""
import numpy as np
import pytest
import torch

import tblis_amd as T
from oracle import pyoracle as po
from tests.helpers import layout_for, random_contraction, storage_size

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
TOL = {torch.complex128: 1e-12, torch.complex64: 1e-6}


def _case(rng, iA, iB, iC, ext, dtype, alpha, beta, conj_A=False, conj_B=False, kinds="random",
          nan_c=False, config=None, tol=None):
    (lA, sA), (lB, sB), (lC, sC) = [layout_for(i, ext, rng, kinds) for i in (iA, iB, iC)]
    mk = lambda l, s: (rng.uniform(-1, 1, storage_size(l, s)) + 1j * rng.uniform(-1, 1, storage_size(l, s)))
    A, B, C0 = mk(lA, sA), mk(lB, sB), mk(lC, sC)
    if dtype == torch.complex64:
        A, B, C0 = (x.astype(np.complex64).astype(np.complex128) for x in (A, B, C0))
    if nan_c:
        C0[:] = np.nan
    want = po.complex_tensor_mult(alpha, A, lA, sA, iA, B, lB, sB, iB, beta,
                                  np.nan_to_num(C0) if nan_c else C0, lC, sC, iC, conj_A, conj_B)
    dA = torch.from_numpy(A).to(dtype).to(DEV)
    dB = torch.from_numpy(B).to(dtype).to(DEV)
    dC = torch.from_numpy(C0).to(dtype).to(DEV)
    T.mult(alpha, torch.as_strided(dA, lA, sA), iA, torch.as_strided(dB, lB, sB), iB, beta,
           torch.as_strided(dC, lC, sC), iC, conj_A=conj_A, conj_B=conj_B, config=config)
    torch.cuda.synchronize()
    got = dC.cpu().to(torch.complex128).numpy()
    if nan_c:
        # only the addressed elements are overwritten
        off = np.zeros(1, np.int64)
        for n, st in zip(lC, sC):
            off = (off[:, None] + np.arange(n, dtype=np.int64)[None, :] * st).ravel()
        got, want = got[np.unique(off)], want[np.unique(off)]
    nrm = np.linalg.norm(want)
    err = np.linalg.norm(got - want) / (nrm if nrm > 0 else 1.0)
    assert err <= (tol or TOL[dtype]), (iA, iB, iC, ext, dtype, alpha, beta, conj_A, conj_B, err)


SCALARS = [(1.0, 0.0), (0.5 - 0.25j, 1.5 + 0.5j), (2.0, -0.5), (1j, 0.0)]


@pytest.mark.parametrize("seed", range(10))
@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_complex_random(seed, dtype):
    rng = np.random.default_rng(65000 + seed)
    iA, iB, iC, ext = random_contraction(rng, max_dims=3, max_len=9)
    alpha, beta = SCALARS[seed % len(SCALARS)]
    _case(rng, iA, iB, iC, ext, dtype, alpha, beta, conj_A=bool(seed & 1), conj_B=bool(seed & 2))


@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_complex_gemm_sized(dtype):
    """Several panel tiles: C[abij] = A[aebf] B[fjei] with extents up to 24."""
    rng = np.random.default_rng(65100)
    ext = dict(a=24, b=17, e=20, f=9, i=13, j=22)
    _case(rng, "aebf", "fjei", "aibj", ext, dtype, 0.75 + 0.5j, -0.5, kinds="col")
    _case(rng, "aebf", "fjei", "aibj", ext, dtype, -1.0, 0.25j, conj_A=True, kinds="row")


@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_complex_special_cases(dtype):
    rng = np.random.default_rng(65200)
    ext = dict(a=7, b=5, k=6)
    _case(rng, "ak", "kb", "ab", ext, dtype, 1.0 + 1.0j, 0.0, nan_c=True)   # beta = 0 over NaN
    _case(rng, "ak", "kb", "ab", ext, dtype, 0.0, 0.5 - 2.0j)               # alpha = 0: C = beta C
    _case(rng, "ak", "kb", "ab", dict(a=7, b=5, k=0), dtype, 1.0, 2.0j)     # k = 0: C = beta C
    _case(rng, "ak", "kb", "ab", ext, dtype, 1.0, 1.0)                      # beta = 1


@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_complex_diagonal_c(dtype):
    """C[aa] = A[ae] B[ea]: a repeated C label addresses C's diagonal
    (src/iface/3t/mult.cxx:44-48), so only the diagonal changes, also when
    complex beta and alpha are applied by passes over C; the off-diagonal
    entries keep their values."""
    rng = np.random.default_rng(65300)
    ext = dict(a=6, e=5)
    for alpha, beta in [(1.0 + 1.0j, 0.0), (0.0, 0.5 + 0.5j), (0.0, 0.5), (1.0 + 1.0j, 0.5 + 0.5j),
                        (2.0, 0.5 + 0.5j), (0.5 - 1.0j, 1.0)]:
        _case(rng, "ae", "ea", "aa", ext, dtype, alpha, beta, kinds="col")


def test_complex_scalar_reset():
    """After the call C's scalar is 1 + 0i (src/iface/3t/mult.cxx:162-163)."""
    A = torch.ones(3, 4, dtype=torch.complex128, device=DEV)
    B = torch.ones(4, 2, dtype=torch.complex128, device=DEV)
    C = torch.zeros(3, 2, dtype=torch.complex128, device=DEV)
    d = T._Desc(C, 2.0 - 1.0j)
    a, b = T._Desc(A, 1.0), T._Desc(B, 1.0)
    import ctypes
    T.lib().tblis_tensor_mult(None, None, ctypes.byref(a.d), b"ak", ctypes.byref(b.d), b"kb",
                              ctypes.byref(d.d), b"ab")
    torch.cuda.synchronize()
    assert (d.d.scalar.data.z[0], d.d.scalar.data.z[1], d.d.conj) == (1.0, 0.0, 0)


# ---------------------------------------------------------------- DPD
def _dpd_complex_case(rng, nirrep, iA, iB, iC, ext, irA, irB, layout, dtype, alpha, beta, conj_A=False,
                      conj_B=False):
    irC = irA ^ irB
    lens = lambda idx: [ext[c] for c in idx]
    dA, dB, dC = (po.dpd_init(irA, nirrep, lens(iA), layout), po.dpd_init(irB, nirrep, lens(iB), layout),
                  po.dpd_init(irC, nirrep, lens(iC), layout))
    cx = lambda d: (rng.uniform(-1, 1, max(po.dpd_size(d), 1)) + 1j * rng.uniform(-1, 1, max(po.dpd_size(d), 1)))
    A, B, C0 = cx(dA), cx(dB), cx(dC)
    if dtype == torch.complex64:
        A, B, C0 = (x.astype(np.complex64).astype(np.complex128) for x in (A, B, C0))
    want = po.complex_dpd_mult(alpha, dA, iA, A, dB, iB, B, beta, dC, iC, C0, conj_A, conj_B)
    At, Bt, Ct = (torch.from_numpy(x).to(dtype).to(DEV) for x in (A, B, C0))
    T.dpd_mult(alpha, At, irA, iA, Bt, irB, iB, beta, Ct, irC, iC, nirrep, lens(iA), lens(iB), lens(iC),
               layout=layout, conj_A=conj_A, conj_B=conj_B)
    torch.cuda.synchronize()
    got = Ct.cpu().to(torch.complex128).numpy()
    nrm = np.linalg.norm(want)
    err = np.linalg.norm(got - want) / (nrm if nrm > 0 else 1.0)
    assert err <= TOL[dtype], (nirrep, iA, iB, iC, dtype, alpha, beta, conj_A, conj_B, err)
    return err


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_complex_dpd_random(seed, dtype):
    """Complex DPD contractions (the reference instantiates DPD mult for its
    complex types, src/iface/3t/mult.cxx:292-297) against the oracle's
    complex DPD restatement: random irreps, all six layouts, real and
    complex alpha and beta, conjugated operands."""
    rng = np.random.default_rng(66000 + seed)
    nirrep = [1, 2, 4, 8][seed % 4]
    cases = [("abef", "efij", "abij"), ("aebf", "fjei", "aibj"), ("abe", "eij", "abij"),
             ("aeb", "ebcd", "acbd")]
    iA, iB, iC = cases[seed % 4]
    ext = {c: [int(rng.integers(0, 6)) for _ in range(nirrep)] for c in sorted(set(iA + iB + iC))}
    irA, irB = int(rng.integers(0, nirrep)), int(rng.integers(0, nirrep))
    alpha, beta = SCALARS[seed % len(SCALARS)]
    _dpd_complex_case(rng, nirrep, iA, iB, iC, ext, irA, irB, seed % 6, dtype, alpha, beta,
                      conj_A=bool(seed & 1), conj_B=bool(seed & 2))


@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_complex_dpd_8_irreps(dtype):
    """The 8-irrep DPD structure (XOR product table, PREFIX_ROW_MAJOR) at
    per-irrep extents [6,5,4,3,6,4,3,2], complex alpha and beta."""
    rng = np.random.default_rng(66100)
    ext = {c: [6, 5, 4, 3, 6, 4, 3, 2] for c in "abefij"}
    _dpd_complex_case(rng, 8, "abef", "efij", "abij", ext, 0, 0, 5, dtype, 0.75 + 0.5j, -0.5 + 1j)


# ---------------------------------------------------------------- batch
@pytest.mark.parametrize("scalars", [(1.0, 0.0), (0.5 - 0.25j, 1.5 + 0.5j), (0.0, 0.5j)])
@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_complex_batch(dtype, scalars):
    """A batch of uneven complex contractions (tblis_tensor_mult_batch):
    every instance against the oracle's complex restatement, with real and
    complex alpha / beta, alpha = 0, and conjugated operands."""
    alpha, beta = scalars
    rng = np.random.default_rng(67000 + int(abs(alpha) * 10))
    cA, cB = alpha != 1.0, alpha == 0.0
    As, Bs, Cs, host = [], [], [], []
    for _ in range(24):
        e = {c: int(rng.integers(1, 9)) for c in "abefij"}
        lA, lB, lC = [e[c] for c in "abef"], [e[c] for c in "efij"], [e[c] for c in "abij"]
        mk = lambda l: (rng.uniform(-1, 1, int(np.prod(l))) + 1j * rng.uniform(-1, 1, int(np.prod(l))))
        A, B, C0 = mk(lA), mk(lB), mk(lC)
        if dtype == torch.complex64:
            A, B, C0 = (x.astype(np.complex64).astype(np.complex128) for x in (A, B, C0))
        cm = lambda l: [int(np.prod(l[:i])) for i in range(len(l))]
        host.append((A, B, C0, lA, lB, lC, cm(lA), cm(lB), cm(lC)))
        As.append(torch.as_strided(torch.from_numpy(A).to(dtype).to(DEV), lA, cm(lA)))
        Bs.append(torch.as_strided(torch.from_numpy(B).to(dtype).to(DEV), lB, cm(lB)))
        Cs.append(torch.from_numpy(C0.copy()).to(dtype).to(DEV))
    Cv = [torch.as_strided(c, h[5], h[8]) for c, h in zip(Cs, host)]
    T.mult_batch(alpha, As, ["abef"] * 24, Bs, ["efij"] * 24, beta, Cv, ["abij"] * 24, conj_A=cA, conj_B=cB)
    torch.cuda.synchronize()
    for c, (A, B, C0, lA, lB, lC, sA, sB, sC) in zip(Cs, host):
        want = po.complex_tensor_mult(alpha, A, lA, sA, "abef", B, lB, sB, "efij", beta, C0, lC, sC, "abij",
                                      cA, cB)
        got = c.cpu().to(torch.complex128).numpy()
        err = np.linalg.norm(got - want) / max(np.linalg.norm(want), 1e-300)
        assert err <= TOL[dtype], (lA, lB, alpha, beta, err)
