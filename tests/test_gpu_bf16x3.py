"""GPU parity of the FP32 path: 3-term split-BF16 MFMA against an
extended-precision oracle.  Bar: relative Frobenius error <= 1e-6 (the
split-BF16 accuracy target), stated per case."""
import ctypes as C

import numpy as np
import pytest
import torch

import tblis_amd as T
from oracle import pyoracle as po
from tests.helpers import col_major, layout_for, random_contraction, storage_size
from tests.test_gpu_parity import rel_fro, dev_view, _dpd_case

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
SPLIT_TOL = 1e-6


def outlier_normal(rng, n):
    """normal(0,1) with 1% of the entries scaled by 1e4 (dynamic-range stress)."""
    x = rng.standard_normal(n)
    x[rng.random(n) < 0.01] *= 1e4
    return x.astype(np.float32)


def test_bf16_mfma_lane_layout():
    rng = np.random.default_rng(1)
    A = rng.integers(-8, 9, (32, 16)).astype(np.float32)
    B = rng.integers(-8, 9, (16, 32)).astype(np.float32)
    a = torch.from_numpy(A).to(DEV)
    b = torch.from_numpy(B).to(DEV)
    d = torch.zeros((32, 32), dtype=torch.float32, device=DEV)
    assert T.lib().tblis_hip_debug_mfma_probe_bf16(C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()),
                                                   C.c_void_p(d.data_ptr())) == 0
    np.testing.assert_array_equal(d.cpu().numpy(), A @ B)


def _float_case(rng, iA, iB, iC, ext, kinds, gen, alpha=1.0, beta=0.0, config=None, tol=SPLIT_TOL):
    (lA, sA), (lB, sB), (lC, sC) = [layout_for(i, ext, rng, k) for i, k in zip((iA, iB, iC), kinds)]
    A = gen(rng, storage_size(lA, sA))
    B = gen(rng, storage_size(lB, sB))
    C0 = gen(rng, storage_size(lC, sC))
    ref = po.tensor_mult(alpha, A, lA, sA, iA, B, lB, sB, iB, beta, C0.copy(), lC, sC, iC)
    _, a = dev_view(A, lA, sA)
    _, b = dev_view(B, lB, sB)
    cs, c = dev_view(C0.copy(), lC, sC)
    T.mult(alpha, a, iA, b, iB, beta, c, iC, config=config)
    torch.cuda.synchronize()
    err = rel_fro(cs.cpu().numpy(), ref)
    assert err <= tol, (iA, iB, iC, ext, err)
    return err


uniform = lambda rng, n: rng.uniform(-1, 1, n).astype(np.float32)


@pytest.mark.parametrize("seed", range(16))
def test_float_random_split_bf16(seed):
    rng = np.random.default_rng(40000 + seed)
    iA, iB, iC, ext = random_contraction(rng, max_dims=3, max_len=7, allow_diag=seed % 5 == 0)
    kinds = [["col", "row", "random"][(seed + j) % 3] for j in range(3)]
    _float_case(rng, iA, iB, iC, ext, kinds, uniform, alpha=0.75, beta=[0.0, 1.0, -0.5][seed % 3])
    assert T.last_algorithm() == "bf16x3"


def test_float_outlier_workload_reduced():
    """C[abcd] = A[aebf] B[ecfd], extents 24, normal(0,1) with 1% x 1e4."""
    rng = np.random.default_rng(47)
    ext = {c: 24 for c in "abcdef"}
    err = _float_case(rng, "aebf", "ecfd", "abcd", ext, ["col"] * 3, outlier_normal)
    print("split-bf16 rel. Frobenius error (extents 24, outliers): %.3e" % err)


def test_float_outlier_workload_full_sampled():
    """Full size (extents 96, k = 9216): 4096 random entries of C against
    extended-precision dot products of the same FP32 inputs."""
    ext = {c: 96 for c in "abcdef"}
    lens = lambda s: [ext[c] for c in s]
    lA, lB, lC = lens("aebf"), lens("ecfd"), lens("abcd")
    rng = np.random.default_rng(48)
    A = outlier_normal(rng, int(np.prod(lA)))
    B = outlier_normal(rng, int(np.prod(lB)))
    At, Bt = torch.from_numpy(A).to(DEV), torch.from_numpy(B).to(DEV)
    Ct = torch.zeros(int(np.prod(lC)), dtype=torch.float32, device=DEV)
    a = torch.as_strided(At, lA, col_major(lA))
    b = torch.as_strided(Bt, lB, col_major(lB))
    c = torch.as_strided(Ct, lC, col_major(lC))
    T.mult(1.0, a, "aebf", b, "ecfd", 0.0, c, "abcd")
    torch.cuda.synchronize()
    g = po.canonicalize(lA, col_major(lA), "aebf", lB, col_major(lB), "ecfd", lC, col_major(lC), "abcd")
    d = g.as_dict()
    m, n = int(np.prod(d["len_ac"])), int(np.prod(d["len_bc"]))
    em, en = rng.integers(0, m, 4096), rng.integers(0, n, 4096)
    want = po.mult_entries(g, 1.0, A, B, 0.0, np.zeros(1, np.float32), np.zeros(4096, np.int64), em, en)
    off = np.zeros(4096, np.int64)
    for arr, lens_, strides in ((em, d["len_ac"], d["sC_ac"]), (en, d["len_bc"], d["sC_bc"])):
        f = arr.copy()
        for l, s in zip(lens_, strides):
            off += (f % l) * s
            f //= l
    got = Ct.cpu().numpy()[off].astype(np.float64)
    err = rel_fro(got, want)
    print("split-bf16 rel. Frobenius error (extents 96, k=9216, 4096 samples): %.3e" % err)
    assert err <= SPLIT_TOL


def test_fast_variant_small_k():
    rng = np.random.default_rng(49)
    ext = dict(a=20, b=9, e=12, f=10, i=7, j=11)
    _float_case(rng, "aebf", "fjei", "aibj", ext, ["col"] * 3, uniform, config="bf16x3_fast")
    assert T.last_algorithm() == "bf16x3_fast"


def test_float_dpd_and_batch():
    rng = np.random.default_rng(50)
    ext = {c: [5, 3, 4, 2] for c in "abefij"}
    _dpd_case(rng, 4, "abef", "efij", "abij", ext, 0, 0, 5, alpha=1.0, beta=-1.0, dtype=np.float32,
              tol=SPLIT_TOL)
    As, Bs, Cs, refs = [], [], [], []
    for i in range(20):
        e = {c: int(rng.integers(1, 10)) for c in "abefij"}
        lA, lB, lC = ([e[c] for c in s] for s in ("abef", "efij", "abij"))
        A, B, C0 = (uniform(rng, int(np.prod(l))) for l in (lA, lB, lC))
        refs.append(po.tensor_mult(1.0, A, lA, col_major(lA), "abef", B, lB, col_major(lB), "efij",
                                   2.0, C0.copy(), lC, col_major(lC), "abij"))
        As.append(torch.as_strided(torch.from_numpy(A).to(DEV), lA, col_major(lA)))
        Bs.append(torch.as_strided(torch.from_numpy(B).to(DEV), lB, col_major(lB)))
        Cs.append(torch.as_strided(torch.from_numpy(C0).to(DEV), lC, col_major(lC)))
    T.mult_batch(1.0, As, ["abef"] * 20, Bs, ["efij"] * 20, 2.0, Cs, ["abij"] * 20)
    torch.cuda.synchronize()
    for i in range(20):
        got = torch.as_strided(Cs[i], (Cs[i].numel(),), (1,)).cpu().numpy()
        assert rel_fro(got, refs[i]) <= SPLIT_TOL, i


@pytest.mark.parametrize("config", ["bf16x3", "bf16x3_fast"])
@pytest.mark.parametrize("tile", range(6))
def test_every_bf16_panel_tile(tile, config, monkeypatch):
    """Each tile of the split-BF16 panel GEMM (tiles 4 and 5 split k
    between two and four sets of waves), chunked and single accumulator:
    ragged dense cases, scattered k (incl. k longer than one 256-k
    accumulation chunk per set), DPD (merged irrep GEMMs with C panels)."""
    monkeypatch.setenv("TBLIS_HIP_PANEL_TILE", str(tile))
    T.lib().tblis_hip_clear_plan_cache()
    try:
        rng = np.random.default_rng(44000 + tile)
        _float_case(rng, "aebf", "fjei", "aibj", dict(a=21, b=13, e=9, f=7, i=25, j=14), ["col"] * 3,
                    uniform, alpha=0.5, beta=1.0, config=config)
        _float_case(rng, "abef", "efij", "abij", dict(a=37, b=3, e=70, f=1, i=29, j=2),
                    ["row", "col", "random"], outlier_normal, config=config)
        _float_case(rng, "aek", "kbe", "ab", dict(a=70, b=45, e=3, k=700), ["col", "random", "col"],
                    uniform, alpha=-0.75, beta=0.25, config=config)
        ext = {c: [4, 3, 2, 1, 5, 3, 2, 1] for c in "abefij"}
        _dpd_case(rng, 8, "abef", "efij", "abij", ext, 0, 0, 5, alpha=0.5, beta=1.0,
                  dtype=np.float32, tol=SPLIT_TOL, config=config)
        ext = {c: [3, 5, 0, 2] for c in "abefij"}
        _dpd_case(rng, 4, "abef", "efij", "abij", ext, 1, 3, 4, alpha=1.0, beta=0.5,
                  dtype=np.float32, tol=SPLIT_TOL, config=config)
    finally:
        monkeypatch.delenv("TBLIS_HIP_PANEL_TILE")
        T.lib().tblis_hip_clear_plan_cache()


def test_packed_alpha_zero_and_beta():
    rng = np.random.default_rng(43)
    A = torch.from_numpy(uniform(rng, 300 * 40)).to(DEV).view(300, 40)
    B = torch.from_numpy(uniform(rng, 40 * 280)).to(DEV).view(40, 280)
    C0 = torch.from_numpy(uniform(rng, 300 * 280)).to(DEV).view(300, 280)
    Cm = C0.clone()
    T.mult(0.0, A, "ak", B, "kb", 2.0, Cm, "ab")
    assert torch.equal(Cm, C0 * 2.0)
    Cm = C0.clone()
    T.mult(1.0, A, "ak", B, "kb", 0.5, Cm, "ab")
    ref = (A.double() @ B.double() + 0.5 * C0.double()).cpu().numpy()
    assert rel_fro(Cm.cpu().numpy(), ref) <= SPLIT_TOL
