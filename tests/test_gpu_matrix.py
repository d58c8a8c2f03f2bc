"""GPU parity of tblis_matrix_mult (src/iface/3m/mult.cxx:19-191): random
shapes and strides (column-major, row-major, transposed and padded views),
the reference's special cases (empty C, k = 0, alpha = 0, beta = 0, m or n
= 1), FP64 at 1e-12 and FP32 split-BF16 at 1e-6 relative Frobenius error
against a float64 numpy product of the same inputs."""
import numpy as np
import pytest
import torch

import tblis_amd as T

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _view(rng, rows, cols, kind, dtype):
    """A rows x cols device view of the given layout kind over fresh storage."""
    if kind == "col":
        st = torch.from_numpy(rng.uniform(-1, 1, max(rows * cols, 1)).astype(dtype)).to(DEV)
        return torch.as_strided(st, (rows, cols), (1, max(rows, 1)))
    if kind == "row":
        st = torch.from_numpy(rng.uniform(-1, 1, max(rows * cols, 1)).astype(dtype)).to(DEV)
        return torch.as_strided(st, (rows, cols), (max(cols, 1), 1))
    # padded column-major: leading dimension rows + 3
    ld = rows + 3
    st = torch.from_numpy(rng.uniform(-1, 1, max(ld * cols, 1)).astype(dtype)).to(DEV)
    return torch.as_strided(st, (rows, cols), (1, ld))


def _case(rng, m, n, k, kinds, dtype, alpha, beta, tol):
    A = _view(rng, m, k, kinds[0], dtype)
    B = _view(rng, k, n, kinds[1], dtype)
    Cm = _view(rng, m, n, kinds[2], dtype)
    a64, b64, c64 = (x.cpu().double().numpy() for x in (A, B, Cm))
    want = alpha * (a64 @ b64) + (beta * c64 if beta != 0 else 0)
    T.matrix_mult(alpha, A, B, beta, Cm)
    torch.cuda.synchronize()
    got = Cm.cpu().double().numpy()
    if m * n == 0:
        return
    n_ref = np.linalg.norm(want)
    err = np.linalg.norm(got - want) / (n_ref if n_ref > 0 else 1.0)
    assert err <= tol, (m, n, k, kinds, alpha, beta, err)


@pytest.mark.parametrize("seed", range(12))
def test_matrix_mult_random(seed):
    rng = np.random.default_rng(35000 + seed)
    kinds = ["col", "row", "pad"]
    m, n, k = (int(x) for x in rng.integers(1, 300, 3))
    ks = [kinds[(seed + i) % 3] for i in range(3)]
    _case(rng, m, n, k, ks, np.float64, 0.75, [0.0, 1.0, -0.5][seed % 3], 1e-12)
    _case(rng, m, n, k, ks, np.float32, 1.25, [0.0, 1.0, -0.5][seed % 3], 1e-6)


def test_matrix_mult_special_cases():
    rng = np.random.default_rng(35100)
    for dtype, tol in ((np.float64, 1e-12), (np.float32, 1e-6)):
        _case(rng, 17, 9, 0, ["col"] * 3, dtype, 1.0, 0.5, tol)     # k = 0: C = beta C
        _case(rng, 17, 9, 0, ["col"] * 3, dtype, 1.0, 0.0, tol)     # k = 0, beta = 0: zeros
        _case(rng, 17, 9, 13, ["col"] * 3, dtype, 0.0, 2.0, tol)    # alpha = 0
        _case(rng, 1, 40, 23, ["row"] * 3, dtype, 1.0, 1.0, tol)    # m = 1
        _case(rng, 40, 1, 23, ["col"] * 3, dtype, 1.0, 1.0, tol)    # n = 1
        _case(rng, 1, 1, 64, ["col"] * 3, dtype, 1.0, 0.0, tol)     # dot product
        _case(rng, 31, 29, 1, ["pad", "row", "col"], dtype, 1.0, 0.0, tol)  # outer product
    # empty C: nothing is written, and C's scalar is still reset
    A = torch.zeros((0, 5), dtype=torch.float64, device=DEV)
    B = torch.zeros((5, 3), dtype=torch.float64, device=DEV)
    Cm = torch.zeros((0, 3), dtype=torch.float64, device=DEV)
    T.matrix_mult(1.0, A, B, 0.5, Cm)
