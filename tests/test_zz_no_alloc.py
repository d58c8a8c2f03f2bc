"""The no-allocation entry point tblis_hip_tensor_mult_noalloc against
tblis_tensor_mult (the entry point it is a variant of): two calls in a row,
the second on a smaller contraction, FP64 (1e-12 relative Frobenius, the
FP64 bar) and FP32 through split-BF16 (1e-6); both entry points also
against the oracle on the small case.  The library's device allocation
count must not move across the no-allocation calls, and a contraction
whose workspace exceeds the 64 MiB scratch is refused with an error."""
import numpy as np
import pytest
import torch

import tblis_amd as T
from oracle import pyoracle as po
from tests.helpers import col_major

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def rel_fro(x, ref):
    return float(np.linalg.norm(x - ref) / max(np.linalg.norm(ref), 1e-300))


def _views(ext, dtype, rng):
    lens = lambda s: [ext[c] for c in s]
    lA, lB, lC = lens("aebf"), lens("fjei"), lens("aibj")
    A = rng.uniform(-1, 1, int(np.prod(lA))).astype(dtype)
    B = rng.uniform(-1, 1, int(np.prod(lB))).astype(dtype)
    C0 = rng.uniform(-1, 1, int(np.prod(lC))).astype(dtype)
    return (lA, lB, lC), (A, B, C0)


def _both(ext, dtype, alpha, beta, seed):
    rng = np.random.default_rng(seed)
    (lA, lB, lC), (A, B, C0) = _views(ext, dtype, rng)
    a = torch.as_strided(torch.from_numpy(A).to(DEV), lA, col_major(lA))
    b = torch.as_strided(torch.from_numpy(B).to(DEV), lB, col_major(lB))
    out = []
    for fn in (T.mult, T.mult_noalloc):
        cs = torch.from_numpy(C0.copy()).to(DEV)
        fn(alpha, a, "aebf", b, "fjei", beta, torch.as_strided(cs, lC, col_major(lC)), "aibj")
        torch.cuda.synchronize()
        out.append(cs.cpu().numpy().astype(np.float64))
    return out, (lA, lB, lC), (A, B, C0)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 1e-6)])
def test_noalloc_matches_tensor_mult(dtype, tol):
    T.lib().tblis_hip_clear_plan_cache()
    big = dict(a=48, b=40, e=36, f=24, i=20, j=18)
    small = dict(a=13, b=9, e=7, f=11, i=5, j=6)
    # warm the library (module load, first plan) through the existing entry
    _both(small, dtype, 0.5, 1.0, 1)
    n0 = T.device_alloc_count()
    for ext, seed in ((big, 2), (small, 3)):
        (ref, got), (lA, lB, lC), (A, B, C0) = _both(ext, dtype, 0.5, 1.0, seed)
        err = rel_fro(got, ref)
        print("noalloc vs tensor_mult", dtype.__name__, ext, "%.3e" % err, flush=True)
        assert err <= tol, (ext, err)
    # the oracle on the small case
    want = po.tensor_mult(0.5, A.astype(np.float64), lA, col_major(lA), "aebf", B.astype(np.float64), lB,
                          col_major(lB), "fjei", 1.0, C0.astype(np.float64), lC, col_major(lC), "aibj")
    assert rel_fro(got, want) <= tol
    assert rel_fro(ref, want) <= tol
    # the existing entry point allocated its plans; the no-allocation calls did not
    calls_alloc = T.device_alloc_count() - n0
    (ref, got), _, _ = _both(small, dtype, -1.5, 0.0, 4)
    assert rel_fro(got, ref) <= tol
    n1 = T.device_alloc_count()
    for seed in (5, 6):
        rng = np.random.default_rng(seed)
        (lA, lB, lC), (A, B, C0) = _views(small, dtype, rng)
        a = torch.as_strided(torch.from_numpy(A).to(DEV), lA, col_major(lA))
        b = torch.as_strided(torch.from_numpy(B).to(DEV), lB, col_major(lB))
        cs = torch.from_numpy(C0).to(DEV)
        T.mult_noalloc(1.0, a, "aebf", b, "fjei", 0.0, torch.as_strided(cs, lC, col_major(lC)), "aibj")
    torch.cuda.synchronize()
    assert T.device_alloc_count() == n1, "the no-allocation entry point allocated device memory"
    assert calls_alloc >= 0


def test_noalloc_refuses_oversized(monkeypatch):
    # direct reads off: both operands are packed, 2 x 72 MiB of FP64 panels
    monkeypatch.setenv("TBLIS_HIP_DIRECT", "0")
    m = n = k = 3072
    A = torch.zeros(2 * m * 2 * k, dtype=torch.float64, device=DEV)
    B = torch.zeros(2 * k * 2 * n, dtype=torch.float64, device=DEV)
    Cm = torch.zeros(m * n, dtype=torch.float64, device=DEV)
    a = torch.as_strided(A, (m, k), (2, 4 * m))
    b = torch.as_strided(B, (k, n), (2, 4 * k))
    c = torch.as_strided(Cm, (m, n), (1, m))
    n0 = T.device_alloc_count()
    with pytest.raises(T.TblisError, match="scratch"):
        T.mult_noalloc(1.0, a, "ak", b, "kb", 0.0, c, "ab")
    assert T.device_alloc_count() == n0
    monkeypatch.delenv("TBLIS_HIP_DIRECT")
    # a contraction that fits still runs after the refusal
    (ref, got), _, _ = _both(dict(a=9, b=7, e=5, f=6, i=4, j=3), np.float64, 1.0, 0.5, 7)
    assert rel_fro(got, ref) <= 1e-12
