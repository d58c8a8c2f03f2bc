"""GPU parity: the HIP path (through the C ABI) against the oracle on the same
seeded inputs.  FP64 bar: relative Frobenius error <= 1e-12."""
import ctypes as C
import itertools

import numpy as np
import pytest
import torch

import tblis_amd as T
from oracle import pyoracle as po
from tests.helpers import col_major, layout_for, random_contraction, storage_size

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
FP64_TOL = 1e-12


def rel_fro(x, ref):
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    n = np.linalg.norm(ref)
    return np.linalg.norm(x - ref) / (n if n > 0 else 1.0)


def dev_view(storage_np, lens, strides):
    st = torch.from_numpy(storage_np).to(DEV)
    return st, torch.as_strided(st, lens, strides)


def _dense_case(rng, iA, iB, iC, ext, kinds, dtype=np.float64, alpha=0.75, beta=-0.5,
                config=None, tol=FP64_TOL):
    (lA, sA), (lB, sB), (lC, sC) = [layout_for(i, ext, rng, k) for i, k in zip((iA, iB, iC), kinds)]
    A = rng.uniform(-1, 1, storage_size(lA, sA)).astype(dtype)
    B = rng.uniform(-1, 1, storage_size(lB, sB)).astype(dtype)
    C0 = rng.uniform(-1, 1, storage_size(lC, sC)).astype(dtype)
    ref = po.tensor_mult(alpha, A, lA, sA, iA, B, lB, sB, iB, beta, C0.copy(), lC, sC, iC)
    _, a = dev_view(A, lA, sA)
    _, b = dev_view(B, lB, sB)
    cs, c = dev_view(C0.copy(), lC, sC)
    T.mult(alpha, a, iA, b, iB, beta, c, iC, config=config)
    torch.cuda.synchronize()
    got = cs.cpu().numpy()
    err = rel_fro(got, ref)
    assert err <= tol, (iA, iB, iC, ext, err)
    return err


def test_mfma_lane_layout():
    out = torch.zeros(256, dtype=torch.float64, device=DEV)
    assert T.lib().tblis_hip_debug_mfma_probe(C.c_void_p(out.data_ptr())) == 0
    got = out.cpu().numpy().reshape(16, 16)
    i = np.arange(16)[:, None]
    k = np.arange(4)[None, :]
    Am = 1000.0 * i + k + 1
    Bm = (np.arange(4)[:, None] + 1.0) * (np.arange(16)[None, :] + 1.0)
    np.testing.assert_array_equal(got, Am @ Bm)


@pytest.mark.parametrize("seed", range(40))
def test_dense_random_fp64(seed):
    rng = np.random.default_rng(10000 + seed)
    iA, iB, iC, ext = random_contraction(rng, max_dims=3, max_len=6, allow_diag=seed % 5 == 0)
    kinds = [["col", "row", "random"][(seed + j) % 3] for j in range(3)]
    beta = [0.0, 1.0, -0.5, 2.0][seed % 4]
    _dense_case(rng, iA, iB, iC, ext, kinds, beta=beta)


@pytest.mark.parametrize("seed", range(8))
def test_dense_random_reference_config(seed):
    """The reference's portable configuration name ("reference") selects the
    same MFMA kernels and agrees with the oracle."""
    rng = np.random.default_rng(11000 + seed)
    iA, iB, iC, ext = random_contraction(rng, max_dims=3, max_len=6)
    _dense_case(rng, iA, iB, iC, ext, ["random"] * 3, config="reference")


@pytest.mark.parametrize("shape", [(24, 24), (32, 16), (40, 24)])
def test_large_tiles_fp64(shape):
    """Sizes that select the 128x128 tile (and its edge tiles)."""
    n1, n2 = shape
    ext = dict(a=n1, b=n2, e=n1, f=n2, i=n1, j=n2)
    rng = np.random.default_rng(n1 * 100 + n2)
    _dense_case(rng, "abef", "efij", "abij", ext, ["col"] * 3, alpha=1.0, beta=0.0)
    _dense_case(rng, "aebf", "fjei", "aibj", ext, ["col"] * 3, alpha=0.5, beta=1.0)


def test_workload_dense_24():
    """C[abij] = A[abef] B[efij], extents 24, column major, alpha 1, beta 0."""
    rng = np.random.default_rng(42)
    ext = {c: 24 for c in "abefij"}
    _dense_case(rng, "abef", "efij", "abij", ext, ["col"] * 3, alpha=1.0, beta=0.0)


def test_workload_permuted_reduced():
    """C[aibj] += 0.5 A[aebf] B[fjei] with the full permutation structure at
    a=b=e=f=16, i=j=8."""
    rng = np.random.default_rng(43)
    ext = dict(a=16, b=16, e=16, f=16, i=8, j=8)
    _dense_case(rng, "aebf", "fjei", "aibj", ext, ["col"] * 3, alpha=0.5, beta=1.0)


def test_workload_permuted_full_sampled():
    """Full-size C[aibj] += 0.5 A[aebf] B[fjei] (a=b=e=f=64, i=j=32): 4096
    random entries of C against extended-precision dot products."""
    ext = dict(a=64, b=64, e=64, f=64, i=32, j=32)
    lens = lambda s: [ext[c] for c in s]
    lA, lB, lC = lens("aebf"), lens("fjei"), lens("aibj")
    gen = torch.Generator(device=DEV).manual_seed(5)
    A = torch.rand(int(np.prod(lA)), dtype=torch.float64, device=DEV, generator=gen) * 2 - 1
    B = torch.rand(int(np.prod(lB)), dtype=torch.float64, device=DEV, generator=gen) * 2 - 1
    C0 = torch.rand(int(np.prod(lC)), dtype=torch.float64, device=DEV, generator=gen) * 2 - 1
    Cd = C0.clone()
    a = torch.as_strided(A, lA, col_major(lA))
    b = torch.as_strided(B, lB, col_major(lB))
    c = torch.as_strided(Cd, lC, col_major(lC))
    T.mult(0.5, a, "aebf", b, "fjei", 1.0, c, "aibj")
    torch.cuda.synchronize()
    g = po.canonicalize(lA, col_major(lA), "aebf", lB, col_major(lB), "fjei", lC, col_major(lC),
                        "aibj")
    rng = np.random.default_rng(6)
    m = int(np.prod(g.as_dict()["len_ac"]))
    n = int(np.prod(g.as_dict()["len_bc"]))
    em = rng.integers(0, m, 4096)
    en = rng.integers(0, n, 4096)
    An, Bn, Cn = A.cpu().numpy(), B.cpu().numpy(), C0.cpu().numpy()
    want = po.mult_entries(g, 0.5, An, Bn, 1.0, Cn, np.zeros(4096, np.int64), em, en)
    d = g.as_dict()
    off = np.zeros(4096, np.int64)
    for arr, lens_, strides in ((em, d["len_ac"], d["sC_ac"]), (en, d["len_bc"], d["sC_bc"])):
        f = arr.copy()
        for l, s in zip(lens_, strides):
            off += (f % l) * s
            f //= l
    got = Cd.cpu().numpy()[off]
    assert rel_fro(got, want) <= FP64_TOL


def test_special_rules():
    rng = np.random.default_rng(3)
    lA, lB, lC = [30, 20], [20, 40], [30, 40]
    A = torch.from_numpy(rng.uniform(-1, 1, 600)).to(DEV).view(30, 20)
    B = torch.from_numpy(rng.uniform(-1, 1, 800)).to(DEV).view(20, 40)
    ref = (A @ B).cpu().numpy()
    # beta == 0 overwrites C without reading it (NaN in C)
    Cm = torch.full((30, 40), float("nan"), dtype=torch.float64, device=DEV)
    T.mult(1.0, A, "ak", B, "kb", 0.0, Cm, "ab")
    assert np.allclose(Cm.cpu().numpy(), ref, rtol=1e-13, atol=1e-13)
    # alpha == 0 reads neither A nor B (NaN in A)
    An = torch.full((30, 20), float("nan"), dtype=torch.float64, device=DEV)
    Cm = torch.ones((30, 40), dtype=torch.float64, device=DEV)
    T.mult(0.0, An, "ak", B, "kb", 2.0, Cm, "ab")
    assert (Cm.cpu().numpy() == 2.0).all()
    T.mult(0.0, An, "ak", B, "kb", 0.0, Cm, "ab")
    assert (Cm.cpu().numpy() == 0.0).all()
    # k == 0: C <- beta*C
    A0 = torch.zeros((30, 0), dtype=torch.float64, device=DEV)
    B0 = torch.zeros((0, 40), dtype=torch.float64, device=DEV)
    Cm = torch.ones((30, 40), dtype=torch.float64, device=DEV)
    T.mult(1.0, A0, "ak", B0, "kb", 3.0, Cm, "ab")
    assert (Cm.cpu().numpy() == 3.0).all()
    # empty m: nothing happens
    T.mult(1.0, torch.zeros((0, 20), dtype=torch.float64, device=DEV), "ak", B, "kb", 0.0,
           torch.zeros((0, 40), dtype=torch.float64, device=DEV), "ab")
    # outer product (no contracted index), dot product (scalar C), batch only
    a = torch.from_numpy(rng.uniform(-1, 1, 7)).to(DEV)
    b = torch.from_numpy(rng.uniform(-1, 1, 9)).to(DEV)
    c = torch.zeros((7, 9), dtype=torch.float64, device=DEV)
    T.mult(1.0, a, "a", b, "b", 0.0, c, "ab")
    assert np.allclose(c.cpu().numpy(), np.outer(a.cpu(), b.cpu()), rtol=1e-14)
    s = torch.zeros((), dtype=torch.float64, device=DEV)
    a2 = torch.from_numpy(rng.uniform(-1, 1, 9)).to(DEV)
    T.mult(1.0, a2, "k", b, "k", 0.0, s, "")
    assert abs(s.item() - float(a2.cpu() @ b.cpu())) < 1e-13
    h = torch.zeros(9, dtype=torch.float64, device=DEV)
    T.mult(1.0, a2, "k", b, "k", 0.0, h, "k")
    assert np.allclose(h.cpu().numpy(), (a2 * b).cpu().numpy(), rtol=1e-15)


def test_float_through_fp64_mfma():
    rng = np.random.default_rng(9)
    ext = dict(a=12, b=10, e=9, f=11, i=7, j=13)
    _dense_case(rng, "aebf", "fjei", "aibj", ext, ["col"] * 3, dtype=np.float32,
                config="fp64_mfma", tol=2e-7)


@pytest.mark.parametrize("tile", [9, 10, 11])
def test_lds_staged_tiles(tile, monkeypatch):
    """The FP64 tiles that read both operands from the tensors through LDS
    (forced with TBLIS_HIP_PANEL_TILE): ragged edges, k not a multiple of
    the 32-k stage, operands with lanes along rows and along k, scattered
    k groups, broadcast dimensions, FP32 data, beta = 0 over NaN and the
    alpha = 0 rule."""
    monkeypatch.setenv("TBLIS_HIP_PANEL_TILE", str(tile))
    T.lib().tblis_hip_clear_plan_cache()
    try:
        rng = np.random.default_rng(12000 + tile)
        _dense_case(rng, "aebf", "fjei", "aibj", dict(a=21, b=13, e=9, f=7, i=25, j=14), ["col"] * 3,
                    alpha=0.5, beta=1.0)
        assert T.last_tile() == tile
        _dense_case(rng, "abef", "efij", "abij", dict(a=37, b=3, e=70, f=1, i=29, j=2),
                    ["row", "col", "random"], beta=0.0)
        _dense_case(rng, "aek", "kbe", "ab", dict(a=130, b=67, e=5, k=41), ["random"] * 3, beta=-2.0)
        _dense_case(rng, "aebf", "fjei", "aibj", dict(a=12, b=10, e=9, f=11, i=7, j=13), ["col"] * 3,
                    dtype=np.float32, config="fp64_mfma", tol=2e-7)
        # beta = 0 over NaN, alpha = 0
        A = torch.rand(70, 33, dtype=torch.float64, device=DEV)
        B = torch.rand(33, 50, dtype=torch.float64, device=DEV)
        Cm = torch.full((70, 50), float("nan"), dtype=torch.float64, device=DEV)
        T.mult(1.0, A, "ak", B, "kb", 0.0, Cm, "ab")
        assert torch.allclose(Cm, A @ B, rtol=1e-13, atol=1e-13)
        C0 = torch.rand(70, 50, dtype=torch.float64, device=DEV)
        Cm = C0.clone()
        T.mult(0.0, torch.full_like(A, float("nan")), "ak", B, "kb", 3.0, Cm, "ab")
        assert torch.equal(Cm, 3.0 * C0)
    finally:
        monkeypatch.delenv("TBLIS_HIP_PANEL_TILE")
        T.lib().tblis_hip_clear_plan_cache()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_broadcast_stride0_dims(dtype):
    """Operands with a stride-0 (broadcast) dimension, which every element of
    it reads at one address: k of stride 0 under rows of stride 2 (packed
    through the general, not the k-fast, gather), broadcast rows, and a
    broadcast B; FP64 and split-BF16."""
    rng = np.random.default_rng(60 + (dtype == np.float32))
    tol = FP64_TOL if dtype == np.float64 else 1e-6
    cases = [([40, 24], [2, 0], [24, 16], [1, 24]),
             ([40, 24], [0, 1], [24, 16], [16, 1]),
             ([40, 24], [1, 40], [24, 16], [0, 1]),
             ([40, 24], [24, 1], [24, 16], [1, 0]),
             ([70, 33], [3, 0], [33, 50], [0, 33])]
    for lA, sA, lB, sB in cases:
        lC = [lA[0], lB[1]]
        sC = col_major(lC)
        A = rng.uniform(-1, 1, storage_size(lA, sA)).astype(dtype)
        B = rng.uniform(-1, 1, storage_size(lB, sB)).astype(dtype)
        C0 = rng.uniform(-1, 1, storage_size(lC, sC)).astype(dtype)
        ref = po.tensor_mult(0.5, A, lA, sA, "ae", B, lB, sB, "eb", 1.0, C0.copy(), lC, sC, "ab")
        _, a = dev_view(A, lA, sA)
        _, b = dev_view(B, lB, sB)
        cs, c = dev_view(C0.copy(), lC, sC)
        T.mult(0.5, a, "ae", b, "eb", 1.0, c, "ab")
        torch.cuda.synchronize()
        err = rel_fro(cs.cpu().numpy(), ref)
        assert err <= tol, (lA, sA, lB, sB, dtype, err)


def test_plan_cache_reuse_and_host_pointer_error():
    rng = np.random.default_rng(12)
    for it in range(3):
        A = torch.from_numpy(rng.uniform(-1, 1, (33, 17))).to(DEV)
        B = torch.from_numpy(rng.uniform(-1, 1, (17, 29))).to(DEV)
        Cm = torch.empty((33, 29), dtype=torch.float64, device=DEV)
        T.mult(1.0, A, "ak", B, "kb", 0.0, Cm, "ab")
        assert np.allclose(Cm.cpu().numpy(), (A @ B).cpu().numpy(), rtol=1e-13, atol=1e-13)
    with pytest.raises(T.TblisError, match="device"):
        T.mult(1.0, A.cpu(), "ak", B.cpu(), "kb", 0.0, Cm.cpu(), "ab")


# ---------------------------------------------------------------- DPD
def _dpd_case(rng, nirrep, iA, iB, iC, ext, irA, irB, layout, alpha=1.0, beta=0.0, mode=1,
              config=None, dtype=np.float64, tol=FP64_TOL):
    irC = irA ^ irB
    lens = lambda idx: [ext[c] for c in idx]
    dA, dB, dC = (po.dpd_init(irA, nirrep, lens(iA), layout), po.dpd_init(irB, nirrep, lens(iB), layout),
                  po.dpd_init(irC, nirrep, lens(iC), layout))
    nA, nB, nC = (max(po.dpd_size(d), 1) for d in (dA, dB, dC))
    A = rng.uniform(-1, 1, nA).astype(dtype)
    B = rng.uniform(-1, 1, nB).astype(dtype)
    C0 = rng.uniform(-1, 1, nC).astype(dtype)
    ref = po.dpd_mult(alpha, dA, iA, A, dB, iB, B, beta, dC, iC, C0.copy(), mode=mode)
    At, Bt, Ct = (torch.from_numpy(x).to(DEV) for x in (A, B, C0.copy()))
    T.dpd_mult(alpha, At, irA, iA, Bt, irB, iB, beta, Ct, irC, iC, nirrep, lens(iA), lens(iB),
               lens(iC), layout=layout, config=config)
    torch.cuda.synchronize()
    err = rel_fro(Ct.cpu().numpy(), ref)
    assert err <= tol, (nirrep, iA, iB, iC, err)


@pytest.mark.parametrize("seed", range(24))
def test_dpd_random(seed):
    rng = np.random.default_rng(20000 + seed)
    nirrep = [1, 2, 4, 8][seed % 4]
    cases = [("abef", "efij", "abij"), ("aebf", "fjei", "aibj"), ("abe", "eij", "abij"),
             ("aeb", "ebcd", "acbd")]
    iA, iB, iC = cases[seed % 4]
    ext = {c: [int(rng.integers(0, 7)) for _ in range(nirrep)] for c in sorted(set(iA + iB + iC))}
    irA, irB = int(rng.integers(0, nirrep)), int(rng.integers(0, nirrep))
    _dpd_case(rng, nirrep, iA, iB, iC, ext, irA, irB, seed % 6, alpha=0.5,
              beta=[0.0, 1.0, -2.0][seed % 3])


def test_dpd_workload_8_irreps():
    """C[abij] = A[abef] B[efij], 8 irreps, per-irrep extents
    [12,9,7,5,11,8,6,4] on every index, PREFIX_ROW_MAJOR, all irrep 0."""
    rng = np.random.default_rng(44)
    ext = {c: [12, 9, 7, 5, 11, 8, 6, 4] for c in "abefij"}
    _dpd_case(rng, 8, "abef", "efij", "abij", ext, 0, 0, 5, mode=0)


def test_dpd_with_batch_label():
    rng = np.random.default_rng(45)
    ext = {c: [int(rng.integers(1, 4)) for _ in range(4)] for c in "abez"}
    _dpd_case(rng, 4, "aez", "ebz", "abz", ext, 1, 2, 5, alpha=1.5, beta=0.5, mode=2)


def test_dpd_reference_irrep_check_compat(monkeypatch):
    """With the compatibility switch, the reference front end's A^(B!=C) test
    is reproduced: valid contraction with irrep_A = irrep_B = 2 only scales C."""
    rng = np.random.default_rng(46)
    ext = {c: [2, 1, 3, 2] for c in "abef"}
    lens = lambda idx: [ext[c] for c in idx]
    dA, dB, dC = po.dpd_init(2, 4, lens("ae"), 5), po.dpd_init(2, 4, lens("eb"), 5), po.dpd_init(0, 4, lens("ab"), 5)
    A = torch.from_numpy(rng.uniform(-1, 1, po.dpd_size(dA))).to(DEV)
    B = torch.from_numpy(rng.uniform(-1, 1, po.dpd_size(dB))).to(DEV)
    C0 = rng.uniform(-1, 1, po.dpd_size(dC))
    Ct = torch.from_numpy(C0.copy()).to(DEV)
    monkeypatch.setenv("TBLIS_HIP_DPD_REFERENCE_IRREP_CHECK", "1")
    T.dpd_mult(1.0, A, 2, "ae", B, 2, "eb", 0.5, Ct, 0, "ab", 4, lens("ae"), lens("eb"), lens("ab"))
    assert np.allclose(Ct.cpu().numpy(), 0.5 * C0)
    monkeypatch.delenv("TBLIS_HIP_DPD_REFERENCE_IRREP_CHECK")
    Ct = torch.from_numpy(C0.copy()).to(DEV)
    T.dpd_mult(1.0, A, 2, "ae", B, 2, "eb", 0.5, Ct, 0, "ab", 4, lens("ae"), lens("eb"), lens("ab"))
    ref = po.dpd_mult(1.0, dA, "ae", A.cpu().numpy(), dB, "eb", B.cpu().numpy(), 0.5, dC, "ab",
                      C0.copy(), mode=2)
    assert rel_fro(Ct.cpu().numpy(), ref) <= FP64_TOL


# ---------------------------------------------------------------- batch
def test_batch_uneven():
    rng = np.random.default_rng(50)
    n = 48
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
    T.mult_batch(1.0, As, ["abef"] * n, Bs, ["efij"] * n, 0.5, Cs, ["abij"] * n)
    torch.cuda.synchronize()
    for i in range(n):
        got = torch.as_strided(Cs[i], (Cs[i].numel(),), (1,)).cpu().numpy()
        assert rel_fro(got, refs[i]) <= FP64_TOL, i


@pytest.mark.parametrize("direct", ["1", "0"])
@pytest.mark.parametrize("tile", range(9))
def test_every_panel_tile(tile, direct, monkeypatch):
    """Each tile of the FP64 panel GEMM (forced through TBLIS_HIP_PANEL_TILE;
    tiles 6 and 7 split k between two sets of waves inside the workgroup,
    tile 8 between four),
    with operands read straight from the tensors where their layout allows
    (direct) and with every operand packed: dense contractions with ragged
    edges and scattered k, FP32 data through FP64 MFMA, the 8-irrep DPD
    structure at small extents (merged irrep GEMMs with a C panel), a DPD
    problem with a batch label, and a batch."""
    monkeypatch.setenv("TBLIS_HIP_PANEL_TILE", str(tile))
    monkeypatch.setenv("TBLIS_HIP_DIRECT", direct)
    T.lib().tblis_hip_clear_plan_cache()
    try:
        rng = np.random.default_rng(31000 + tile)
        _dense_case(rng, "aebf", "fjei", "aibj", dict(a=13, b=11, e=9, f=7, i=5, j=14),
                    ["col", "col", "col"])
        _dense_case(rng, "abef", "efij", "abij", dict(a=24, b=7, e=19, f=5, i=33, j=3),
                    ["random", "row", "random"], alpha=1.0, beta=0.0)
        _dense_case(rng, "aebf", "fjei", "aibj", dict(a=9, b=13, e=7, f=6, i=11, j=5),
                    ["col", "col", "col"], dtype=np.float32, config="fp64_mfma", tol=1e-6)
        ext = {c: [4, 3, 2, 1, 5, 3, 2, 1] for c in "abefij"}
        _dpd_case(rng, 8, "abef", "efij", "abij", ext, 0, 0, 5, alpha=0.5, beta=1.0)
        ext = {c: [3, 0, 2, 4] for c in "aebfij"}
        _dpd_case(rng, 4, "aebf", "fjei", "aibj", ext, 1, 3, 4, alpha=2.0, beta=-1.0)
        ext = {c: [int(rng.integers(1, 4)) for _ in range(4)] for c in "abez"}
        _dpd_case(rng, 4, "aez", "ebz", "abz", ext, 1, 2, 5, alpha=1.5, beta=0.5, mode=2)
        test_batch_uneven()
        # permuted workload shape at reduced size: A's rows and k are affine
        _dense_case(rng, "aebf", "fjei", "aibj", dict(a=32, b=5, e=16, f=3, i=7, j=9),
                    ["col"] * 3, alpha=0.5, beta=1.0)
        # k = 3 (two k groups): the k-split tiles have wave sets with no k
        # group, which add zeros in the reduction
        _dense_case(rng, "ae", "eb", "ab", dict(a=70, b=45, e=3), ["col", "row", "col"], alpha=1.5, beta=0.5)
    finally:
        monkeypatch.delenv("TBLIS_HIP_PANEL_TILE")
        monkeypatch.delenv("TBLIS_HIP_DIRECT")
        T.lib().tblis_hip_clear_plan_cache()


@pytest.mark.parametrize("unpack", ["0", "1"])
def test_dpd_c_epilogue_modes(unpack, monkeypatch):
    """DPD irrep GEMMs span several C blocks: their C tiles are written from
    the GEMM epilogue through per-row/column block codes (default) or stored
    as C panels and scattered by the unpack kernel (TBLIS_HIP_C_UNPACK=1);
    both agree with the oracle, FP64 and split-BF16, with beta != 0."""
    monkeypatch.setenv("TBLIS_HIP_C_UNPACK", unpack)
    T.lib().tblis_hip_clear_plan_cache()
    try:
        rng = np.random.default_rng(34000)
        ext = {c: [5, 4, 3, 2, 6, 3, 2, 1] for c in "abefij"}
        _dpd_case(rng, 8, "abef", "efij", "abij", ext, 0, 0, 5, alpha=0.5, beta=-1.0)
        ext = {c: [3, 0, 2, 4] for c in "aebfij"}
        _dpd_case(rng, 4, "aebf", "fjei", "aibj", ext, 1, 3, 2, alpha=2.0, beta=0.5)
        _dpd_case(rng, 4, "aebf", "fjei", "aibj", ext, 2, 1, 0, alpha=1.0, beta=1.0,
                  dtype=np.float32, tol=1e-6)
    finally:
        monkeypatch.delenv("TBLIS_HIP_C_UNPACK")
        T.lib().tblis_hip_clear_plan_cache()
