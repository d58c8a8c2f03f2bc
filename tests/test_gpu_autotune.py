"""The panel-tile autotuner: with tuning on, a new GEMM shape is measured
once (the tuning counter moves, the tuning file gets a line), later plans of
the same shapes reuse the result, and the contraction it then runs matches
the oracle (FP64: 1e-12, split-BF16: 1e-6 relative Frobenius); the
operands a plan reads straight from the tensors are reported per caller
operand."""
import json

import numpy as np
import pytest
import torch

import tblis_amd as T
from tests.test_gpu_bf16x3 import _float_case, uniform
from tests.test_gpu_parity import _dense_case, _dpd_case

pytestmark = pytest.mark.gpu


def test_autotune_picks_and_runs(tmp_path, monkeypatch):
    path = tmp_path / "tune.txt"
    monkeypatch.setenv("TBLIS_HIP_TUNE_FILE", str(path))
    T.set_autotune(True)
    T.lib().tblis_hip_clear_plan_cache()
    try:
        rng = np.random.default_rng(33000)
        r0 = T.autotune_runs()
        ext = dict(a=40, b=33, e=21, f=17, i=29, j=31)
        _dense_case(rng, "aebf", "fjei", "aibj", ext, ["col"] * 3)
        r1 = T.autotune_runs()
        assert r1 == r0 + 1
        assert T.last_tile() >= 0
        # same shapes, new plan (cache cleared): no new measurement
        T.lib().tblis_hip_clear_plan_cache()
        _dense_case(rng, "aebf", "fjei", "aibj", ext, ["col"] * 3)
        assert T.autotune_runs() == r1
        dext = {c: [4, 3, 2, 1, 5, 3, 2, 1] for c in "abefij"}
        _dpd_case(rng, 8, "abef", "efij", "abij", dext, 0, 0, 5, alpha=0.5, beta=1.0)
        _float_case(rng, "aebf", "fjei", "aibj", dict(a=30, b=17, e=12, f=10, i=23, j=19), ["col"] * 3,
                    uniform)
        assert T.autotune_runs() >= r1 + 2
        lines = [json.loads(l) for l in path.read_text().splitlines()]
        assert len(lines) == T.autotune_runs() - r0
        assert all(l["format"] == "tblis-tune-v3" and l["tile"] >= 0 and l["packed"] in (0, 1) for l in lines)
    finally:
        T.set_autotune(False)
        T.lib().tblis_hip_clear_plan_cache()


def test_packed_variant_recorded_and_reused(tmp_path, monkeypatch):
    """Without the tuner the col-major 24^4 contraction reads both operands
    straight from the tensors (last_direct 3), and TBLIS_HIP_DIRECT=0 packs
    both (0).  With the tuner, a tile that reads an operand directly is also
    timed with every operand packed: the choice lands in the tuning file as
    (tile, packed), the plan follows it (packed 1: last_direct 0), a new
    plan of the same shapes reuses it, and every form matches the oracle."""
    from tests.test_gpu_parity import _dense_case
    ext = dict(a=24, b=24, e=24, f=24, i=24, j=24)
    rng = np.random.default_rng(33100)
    T.lib().tblis_hip_clear_plan_cache()
    _dense_case(rng, "abef", "efij", "abij", ext, ["col"] * 3, beta=0.0)
    assert T.last_direct() == 3
    monkeypatch.setenv("TBLIS_HIP_DIRECT", "0")
    T.lib().tblis_hip_clear_plan_cache()
    _dense_case(rng, "abef", "efij", "abij", ext, ["col"] * 3, beta=0.0)
    assert T.last_direct() == 0
    monkeypatch.delenv("TBLIS_HIP_DIRECT")
    path = tmp_path / "tune.txt"
    monkeypatch.setenv("TBLIS_HIP_TUNE_FILE", str(path))
    T.set_autotune(True)
    T.lib().tblis_hip_clear_plan_cache()
    try:
        r0 = T.autotune_runs()
        _dense_case(rng, "abef", "efij", "abij", ext, ["col"] * 3, beta=0.0)
        assert T.autotune_runs() == r0 + 1
        line = json.loads(path.read_text().splitlines()[-1])
        assert line["format"] == "tblis-tune-v3" and line["tile"] == T.last_tile()
        if line["packed"] == 1:
            assert T.last_direct() == 0
        tile, direct = T.last_tile(), T.last_direct()
        T.lib().tblis_hip_clear_plan_cache()
        _dense_case(rng, "abef", "efij", "abij", ext, ["col"] * 3, alpha=-1.25, beta=0.5)
        assert T.autotune_runs() == r0 + 1
        assert (T.last_tile(), T.last_direct()) == (tile, direct)
    finally:
        T.set_autotune(False)
        T.lib().tblis_hip_clear_plan_cache()


def test_last_direct_names_the_callers_operands():
    """last_direct() reports the caller's A and B, whichever way round the
    GEMM runs: A with no unit stride (strides (2e, 2)) is packed, B row-major
    (unit stride along its columns) is read from the tensor, with C row-major
    (the GEMM computes C) and column-major (it computes C^T = B^T A^T)."""
    torch.manual_seed(33200)
    a, e, b = 64, 40, 48
    A = torch.rand(a, 2 * e, dtype=torch.float64, device="cuda")[:, ::2]
    B = torch.rand(e, b, dtype=torch.float64, device="cuda")
    want = A.cpu() @ B.cpu()
    for C in (torch.zeros(a, b, dtype=torch.float64, device="cuda"),
              torch.zeros(b, a, dtype=torch.float64, device="cuda").t()):
        T.lib().tblis_hip_clear_plan_cache()
        T.mult(1.0, A, "ae", B, "eb", 0.0, C, "ab")
        torch.cuda.synchronize()
        assert torch.allclose(C.cpu(), want, rtol=1e-13, atol=1e-13)
        assert T.last_direct() == 2, (C.stride(), T.last_direct())
