"""GPU parity of the benchmark workloads at their full sizes, through the C
ABI, against the oracle:

  * the 4096-instance uneven batch C[abij] = A[abef] B[efij] (extents drawn
    from 4..40 by std::mt19937_64(7)), FP64 and split-BF16, replayed from a
    hipGraph: every instance at sampled entries of C against
    extended-precision dot products, and the largest instances in full
    against the oracle's blocked contraction;
  * the 8-irrep DPD contraction (per-irrep extents [12,9,7,5,11,8,6,4]) in
    split-BF16, against the oracle's block-wise DPD contraction;
  * the 96^4 contraction C[abcd] = A[aebf] B[ecfd] on the FP64 MFMA path
    (FP32 outlier data held as FP64), sampled entries.

Bars: FP64 relative Frobenius error <= 1e-12, split-BF16 <= 1e-6."""
import sys

import numpy as np
import pytest
import torch

import bench
import tblis_amd as T
from oracle import pyoracle as po
from tests.helpers import col_major
from tests.test_gpu_bf16x3 import outlier_normal
from tests.test_gpu_graph import _capture
from tests.test_gpu_parity import _dpd_case, rel_fro

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
FP64_TOL = 1e-12
SPLIT_TOL = 1e-6


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _sample_offsets(d, em, en):
    off = np.zeros(len(em), np.int64)
    for arr, lens, strides in ((em, d["len_ac"], d["sC_ac"]), (en, d["len_bc"], d["sC_bc"])):
        f = np.array(arr, np.int64)
        for ln, s in zip(lens, strides):
            off += (f % ln) * s
            f //= ln
    return off


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_batch_4096_uneven_graph(dtype):
    count, seed = 4096, 7
    exts = bench.batch_extents(count, seed)
    views, dims = bench.batch_views(exts)
    say("batch 4096 uneven, %s: A/B/C buffers of %d/%d/%d elements" % (dtype, *dims))
    g = torch.Generator(device=DEV).manual_seed(seed)
    bufs = [(torch.rand(dims[t], dtype=torch.float64, device=DEV, generator=g) * 2 - 1).to(dtype)
            for t in range(3)]
    inst = [[torch.as_strided(bufs[t], sh, col_major(sh), off) for sh, off in views[t]] for t in range(3)]
    As, Bs, Cs = inst
    fn = T.prepare_batch(1.0, As, ["abef"] * count, Bs, ["efij"] * count, 0.0, Cs, ["abij"] * count)
    graph = _capture(fn)
    say("captured; replaying twice")
    bufs[2].fill_(float("nan"))  # beta == 0: C is not read
    graph.replay()
    graph.replay()
    torch.cuda.synchronize()
    say("replays done (%s, tile %d); checking" % (T.last_algorithm(), T.last_tile()))
    host = [b.cpu().numpy() for b in bufs]
    rng = np.random.default_rng(70)
    got_all, want_all, worst = [], [], 0.0
    for i in range(count):
        (lA, oA), (lB, oB), (lC, oC) = views[0][i], views[1][i], views[2][i]
        A = host[0][oA:oA + int(np.prod(lA))]
        B = host[1][oB:oB + int(np.prod(lB))]
        Cg = host[2][oC:oC + int(np.prod(lC))]
        gform = po.canonicalize(lA, col_major(lA), "abef", lB, col_major(lB), "efij", lC, col_major(lC),
                                "abij")
        d = gform.as_dict()
        m, n = int(np.prod(d["len_ac"])), int(np.prod(d["len_bc"]))
        em, en = rng.integers(0, m, 8), rng.integers(0, n, 8)
        want = po.mult_entries(gform, 1.0, A, B, 0.0, np.zeros(1, A.dtype), np.zeros(8, np.int64), em, en)
        got = Cg[_sample_offsets(d, em, en)].astype(np.float64)
        got_all.append(got)
        want_all.append(want)
        worst = max(worst, rel_fro(got, want))
    err = rel_fro(np.concatenate(got_all), np.concatenate(want_all))
    say("sampled entries of all 4096 instances: rel. Frobenius %.3e, worst instance %.3e" % (err, worst))
    tol = FP64_TOL if dtype == torch.float64 else SPLIT_TOL
    assert err <= tol
    assert worst <= (FP64_TOL if dtype == torch.float64 else 10 * SPLIT_TOL)
    # the two largest instances in full (FP64 data: the oracle's blocked
    # contraction; FP32 data: that contraction on the values widened)
    flops = np.prod(exts.astype(np.float64), axis=1)
    for i in np.argsort(flops)[-2:]:
        (lA, oA), (lB, oB), (lC, oC) = views[0][i], views[1][i], views[2][i]
        A = host[0][oA:oA + int(np.prod(lA))].astype(np.float64)
        B = host[1][oB:oB + int(np.prod(lB))].astype(np.float64)
        want = np.zeros(int(np.prod(lC)))
        po.blis_contract(0, 1.0, A, lA, col_major(lA), "abef", B, lB, col_major(lB), "efij", 0.0, want, lC,
                         col_major(lC), "abij")
        got = host[2][oC:oC + int(np.prod(lC))].astype(np.float64)
        e = rel_fro(got, want)
        say("instance %d extents %s: full rel. Frobenius %.3e" % (i, exts[i].tolist(), e))
        assert e <= tol


def test_dpd_8_irreps_split_bf16():
    """C[abij] = A[abef] B[efij], 8 irreps, per-irrep extents
    [12,9,7,5,11,8,6,4] on every index, PREFIX_ROW_MAJOR, irrep 0, FP32 data
    on the split-BF16 path."""
    rng = np.random.default_rng(144)
    ext = {c: [12, 9, 7, 5, 11, 8, 6, 4] for c in "abefij"}
    say("DPD 8 irreps split-BF16 at the workload's extents")
    _dpd_case(rng, 8, "abef", "efij", "abij", ext, 0, 0, 5, mode=0, dtype=np.float32, tol=SPLIT_TOL)
    assert T.last_algorithm() == "bf16x3"


def test_fp32_96_through_fp64_mfma_sampled():
    """C[abcd] = A[aebf] B[ecfd], extents 96 (k = 9216), normal(0,1) with 1%
    of entries x 1e4, the FP32 values held as FP64, on the FP64 MFMA path:
    4096 random entries of C against extended-precision dot products."""
    ext = {c: 96 for c in "abcdef"}
    lens = lambda s: [ext[c] for c in s]
    lA, lB, lC = lens("aebf"), lens("ecfd"), lens("abcd")
    rng = np.random.default_rng(148)
    A = outlier_normal(rng, int(np.prod(lA))).astype(np.float64)
    B = outlier_normal(rng, int(np.prod(lB))).astype(np.float64)
    say("96^4 FP64 MFMA: operands of %d elements" % A.size)
    At, Bt = torch.from_numpy(A).to(DEV), torch.from_numpy(B).to(DEV)
    Ct = torch.full((int(np.prod(lC)),), float("nan"), dtype=torch.float64, device=DEV)
    a = torch.as_strided(At, lA, col_major(lA))
    b = torch.as_strided(Bt, lB, col_major(lB))
    c = torch.as_strided(Ct, lC, col_major(lC))
    T.mult(1.0, a, "aebf", b, "ecfd", 0.0, c, "abcd", config="fp64_mfma")
    torch.cuda.synchronize()
    assert T.last_algorithm() == "fp64_mfma"
    g = po.canonicalize(lA, col_major(lA), "aebf", lB, col_major(lB), "ecfd", lC, col_major(lC), "abcd")
    d = g.as_dict()
    m, n = int(np.prod(d["len_ac"])), int(np.prod(d["len_bc"]))
    em, en = rng.integers(0, m, 4096), rng.integers(0, n, 4096)
    want = po.mult_entries(g, 1.0, A, B, 0.0, np.zeros(1), np.zeros(4096, np.int64), em, en)
    got = Ct.cpu().numpy()[_sample_offsets(d, em, en)]
    err = rel_fro(got, want)
    say("96^4 FP64 MFMA, 4096 samples: rel. Frobenius %.3e" % err)
    assert err <= FP64_TOL
