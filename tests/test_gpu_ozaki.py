"""GPU parity of the FP64-accurate mode (config "ozaki": int8 MFMA over 7-bit
digit planes of power-of-two-scaled rows, 8 digits unless
TBLIS_HIP_OZAKI_SLICES picks 6 or 7) against the oracle's FP64 result
on the same seeded inputs.  The reference has no such mode; the bar is the
FP64 drop-in's, relative Frobenius error <= 1e-12, here on dense, DPD and
batched contractions, FP64 and FP32 data, rows and columns whose magnitudes
differ by up to 1e16 (checked row by row, which a single global scale would
fail), k long enough to flush the int32 accumulators, and the alpha/beta
rules."""
import numpy as np
import pytest
import torch

import tblis_amd as T
from oracle import pyoracle as po
from tests.helpers import col_major, random_contraction
from tests.test_gpu_parity import FP64_TOL, _dense_case, _dpd_case, rel_fro

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


@pytest.mark.parametrize("seed", range(16))
def test_ozaki_dense_random(seed):
    rng = np.random.default_rng(71000 + seed)
    iA, iB, iC, ext = random_contraction(rng, max_dims=3, max_len=7, allow_diag=seed % 5 == 0)
    kinds = [["col", "row", "random"][(seed + j) % 3] for j in range(3)]
    beta = [0.0, 1.0, -0.5, 2.0][seed % 4]
    _dense_case(rng, iA, iB, iC, ext, kinds, beta=beta, config="ozaki")
    assert T.last_algorithm() == "ozaki"


def test_ozaki_several_tiles():
    """C[aibj] = A[aebf] B[fjei] over several 64 x 64 tiles and partial edges."""
    rng = np.random.default_rng(71100)
    ext = dict(a=24, b=9, e=20, f=13, i=10, j=22)
    _dense_case(rng, "aebf", "fjei", "aibj", ext, ["col", "row", "col"], alpha=0.5, beta=1.0, config="ozaki")


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_ozaki_dynamic_range_rows(dtype):
    """Rows of A and columns of B scaled by 10^u, u uniform in [-8, 8]: every
    row of C keeps the FP64 bar relative to its own size."""
    rng = np.random.default_rng(71200)
    m, n, k = 96, 80, 300
    A = rng.uniform(-1, 1, (m, k)) * 10.0 ** rng.uniform(-8, 8, (m, 1))
    B = rng.uniform(-1, 1, (k, n)) * 10.0 ** rng.uniform(-8, 8, (1, n))
    A, B = A.astype(dtype), B.astype(dtype)
    ref = A.astype(np.float64) @ B.astype(np.float64)
    At = torch.from_numpy(A).to(DEV)
    Bt = torch.from_numpy(B).to(DEV)
    Ct = torch.zeros(m, n, dtype=At.dtype, device=DEV)
    T.mult(1.0, At, "ak", Bt, "kb", 0.0, Ct, "ab", config="ozaki")
    torch.cuda.synchronize()
    got = Ct.cpu().numpy().astype(np.float64)
    if dtype == np.float64:
        for i in range(m):
            assert rel_fro(got[i], ref[i]) <= FP64_TOL, i
        for j in range(n):
            assert rel_fro(got[:, j], ref[:, j]) <= FP64_TOL, j
    else:
        # FP32 inputs: the products are exact, the FP32 result is the FP64
        # sum rounded once
        np.testing.assert_allclose(got, ref.astype(np.float32).astype(np.float64), rtol=2 ** -23, atol=0)


def test_ozaki_fp32_inputs_fp64_accurate():
    """FP32 tensors: the result equals the float64 product of the same FP32
    values rounded to FP32 (no split-BF16 error)."""
    rng = np.random.default_rng(71300)
    ext = dict(a=20, b=18, e=33, f=7)
    lA, lB, lC = [ext[c] for c in "aef"], [ext[c] for c in "efb"], [ext[c] for c in "ab"]
    A = rng.standard_normal(lA).astype(np.float32)
    A[rng.random(lA) < 0.01] *= 1e4
    B = rng.standard_normal(lB).astype(np.float32)
    ref = np.einsum("aef,efb->ab", A.astype(np.float64), B.astype(np.float64))
    Ct = torch.zeros(lC, dtype=torch.float32, device=DEV)
    T.mult(1.0, torch.from_numpy(A).to(DEV), "aef", torch.from_numpy(B).to(DEV), "efb", 0.0, Ct, "ab",
           config="ozaki")
    torch.cuda.synchronize()
    got = Ct.cpu().numpy().astype(np.float64)
    np.testing.assert_allclose(got, ref.astype(np.float32).astype(np.float64), rtol=2 ** -23, atol=0)


def test_ozaki_long_k_flush():
    """k = 9000 > 8160: the int32 digit accumulators are flushed mid-sum."""
    rng = np.random.default_rng(71400)
    m, n, k = 40, 72, 9000
    A = rng.uniform(-1, 1, (m, k))
    B = rng.uniform(-1, 1, (k, n))
    Ct = torch.zeros(m, n, dtype=torch.float64, device=DEV)
    T.mult(1.0, torch.from_numpy(A).to(DEV), "ak", torch.from_numpy(B).to(DEV), "kb", 0.0, Ct, "ab",
           config="ozaki")
    torch.cuda.synchronize()
    assert rel_fro(Ct.cpu().numpy(), A @ B) <= FP64_TOL


def test_ozaki_special_rules():
    rng = np.random.default_rng(71500)
    A = torch.from_numpy(rng.uniform(-1, 1, (7, 6))).to(DEV)
    B = torch.from_numpy(rng.uniform(-1, 1, (6, 5))).to(DEV)
    ref = A.cpu().numpy() @ B.cpu().numpy()
    # beta = 0 overwrites C without reading it (NaN stays out)
    C = torch.full((7, 5), float("nan"), dtype=torch.float64, device=DEV)
    T.mult(2.0, A, "ak", B, "kb", 0.0, C, "ab", config="ozaki")
    assert rel_fro(C.cpu().numpy(), 2.0 * ref) <= FP64_TOL
    # alpha = 0: C = beta C
    C0 = rng.uniform(-1, 1, (7, 5))
    C = torch.from_numpy(C0.copy()).to(DEV)
    T.mult(0.0, A, "ak", B, "kb", -0.5, C, "ab", config="ozaki")
    assert np.array_equal(C.cpu().numpy(), -0.5 * C0)
    # k = 0: C = beta C
    C = torch.from_numpy(C0.copy()).to(DEV)
    T.mult(1.0, A[:, :0], "ak", B[:0, :], "kb", 3.0, C, "ab", config="ozaki")
    assert np.array_equal(C.cpu().numpy(), 3.0 * C0)
    # zero rows and columns
    Az = A.clone()
    Az[2] = 0
    Bz = B.clone()
    Bz[:, 3] = 0
    C = torch.zeros(7, 5, dtype=torch.float64, device=DEV)
    T.mult(1.0, Az, "ak", Bz, "kb", 0.0, C, "ab", config="ozaki")
    assert rel_fro(C.cpu().numpy(), Az.cpu().numpy() @ Bz.cpu().numpy()) <= FP64_TOL
    assert not C[2].any() and not C[:, 3].any()


@pytest.mark.parametrize("seed", range(6))
def test_ozaki_dpd(seed):
    rng = np.random.default_rng(71600 + seed)
    nirrep = [1, 2, 4, 8][seed % 4]
    iA, iB, iC = [("abef", "efij", "abij"), ("aebf", "fjei", "aibj"), ("abe", "eij", "abij")][seed % 3]
    ext = {c: [int(rng.integers(0, 7)) for _ in range(nirrep)] for c in sorted(set(iA + iB + iC))}
    irA, irB = int(rng.integers(0, nirrep)), int(rng.integers(0, nirrep))
    _dpd_case(rng, nirrep, iA, iB, iC, ext, irA, irB, seed % 6, alpha=0.5, beta=[0.0, 1.0, -2.0][seed % 3],
              config="ozaki")


def test_ozaki_dpd_workload_8_irreps():
    rng = np.random.default_rng(71700)
    ext = {c: [12, 9, 7, 5, 11, 8, 6, 4] for c in "abefij"}
    _dpd_case(rng, 8, "abef", "efij", "abij", ext, 0, 0, 5, mode=0, config="ozaki")


def test_ozaki_batch():
    rng = np.random.default_rng(71800)
    n = 24
    As, Bs, Cs, refs = [], [], [], []
    for i in range(n):
        e = {c: int(rng.integers(1, 12)) for c in "abefij"}
        lA, lB, lC = ([e[c] for c in s] for s in ("abef", "efij", "abij"))
        A = rng.uniform(-1, 1, int(np.prod(lA)))
        B = rng.uniform(-1, 1, int(np.prod(lB)))
        C0 = rng.uniform(-1, 1, int(np.prod(lC)))
        refs.append(po.tensor_mult(1.0, A, lA, col_major(lA), "abef", B, lB, col_major(lB), "efij",
                                   0.5, C0.copy(), lC, col_major(lC), "abij"))
        As.append(torch.as_strided(torch.from_numpy(A).to(DEV), lA, col_major(lA)))
        Bs.append(torch.as_strided(torch.from_numpy(B).to(DEV), lB, col_major(lB)))
        Cs.append(torch.as_strided(torch.from_numpy(C0).to(DEV), lC, col_major(lC)))
    T.mult_batch(1.0, As, ["abef"] * n, Bs, ["efij"] * n, 0.5, Cs, ["abij"] * n, config="ozaki")
    torch.cuda.synchronize()
    for i in range(n):
        got = torch.as_strided(Cs[i], (Cs[i].numel(),), (1,)).cpu().numpy()
        assert rel_fro(got, refs[i]) <= FP64_TOL, i


def test_ozaki_digit_truncation_is_the_int8_path():
    """The mode's own rounding, which FP64 arithmetic would not produce: in
    the row [1, x], x = 2^-4 (1 + 2^-52), x sits 5 bits below the row's
    scale and needs 57 bits there, more than 6, 7 or 8 digits hold (42, 49,
    56), so the digits keep 2^-4 and [1, x] . [0, 1] is exactly 2^-4
    (FP64 MFMA gives x)."""
    x = 2.0 ** -4 * (1.0 + 2.0 ** -52)
    A = torch.tensor([[1.0, x]], dtype=torch.float64, device=DEV)
    B = torch.tensor([[0.0], [1.0]], dtype=torch.float64, device=DEV)
    C = torch.zeros(1, 1, dtype=torch.float64, device=DEV)
    T.mult(1.0, A, "ak", B, "kb", 0.0, C, "ab", config="ozaki")
    assert C.item() == 2.0 ** -4
    T.mult(1.0, A, "ak", B, "kb", 0.0, C, "ab")
    assert C.item() == x


@pytest.mark.parametrize("shape", [(37, 45, 1000), (64, 64, 64), (70, 9, 9000), (5, 130, 33), (16, 20, 20000)])
def test_ozaki_bit_exact_vs_restatement(shape):
    """Bit for bit against the oracle's restatement of the scheme (same
    digits, exact integer diagonal sums, the same FP64 order; k = 9000 spans
    two accumulator flushes of 8160 k), with rows and columns of very different
    magnitude and an all-zero row."""
    import os
    S = int(os.environ.get("TBLIS_HIP_OZAKI_SLICES", "8"))
    m, n, k = shape
    rng = np.random.default_rng(72000 + m)
    A = rng.standard_normal((m, k)) * 10.0 ** rng.uniform(-6, 6, (m, 1))
    B = rng.standard_normal((k, n)) * 10.0 ** rng.uniform(-6, 6, (1, n))
    A[m // 2] = 0.0
    want = po.ozaki_gemm(A, B, S)
    C = torch.zeros(m, n, dtype=torch.float64, device=DEV)
    T.mult(1.0, torch.from_numpy(A).to(DEV), "ak", torch.from_numpy(B).to(DEV), "kb", 0.0, C, "ab",
           config="ozaki")
    torch.cuda.synchronize()
    assert np.array_equal(C.cpu().numpy(), want)
    # FP32 data: the same digits of the widened values, rounded once to FP32
    A32, B32 = A.astype(np.float32), B.astype(np.float32)
    want32 = po.ozaki_gemm(A32.astype(np.float64), B32.astype(np.float64), S).astype(np.float32)
    C32 = torch.zeros(m, n, dtype=torch.float32, device=DEV)
    T.mult(1.0, torch.from_numpy(A32).to(DEV), "ak", torch.from_numpy(B32).to(DEV), "kb", 0.0, C32, "ab",
           config="ozaki")
    torch.cuda.synchronize()
    assert np.array_equal(C32.cpu().numpy(), want32)


@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_ozaki_complex(dtype):
    """Complex contractions run as real GEMMs of the same mode: both complex
    types meet the FP64 bar."""
    from tests.test_gpu_complex import _case
    rng = np.random.default_rng(72100)
    ext = dict(a=12, b=7, e=9, f=5, i=6, j=10)
    _case(rng, "aebf", "fjei", "aibj", ext, dtype, 0.75 + 0.5j, -0.5, kinds="col", config="ozaki",
          tol=FP64_TOL if dtype == torch.complex128 else 1e-7)
    assert T.last_algorithm() == "ozaki"


def test_ozaki_indexed():
    """Indexed contractions (entry GEMMs scaled by the entry factors while
    packed) in the same mode, at the FP64 bar."""
    from tests.test_gpu_indexed import CASES, _run
    for t, (alpha, beta) in enumerate([(1.0, 0.0), (0.75, 0.5)]):
        assert _run(CASES[1], np.float64, 72200 + t, alpha, beta, fill=1.0, config="ozaki") or t
