"""Pack chunk shapes.  The planner packs each job in chunks of 4 row groups
by at most 8 k groups, and switches to wide chunks (a job's k groups split
evenly, up to 8 row groups within a 32-group budget) when the narrow ones
outnumber the resident four-wave workgroups; when they fit in one round it
splits them further (down to 8 groups, split-BF16 4) while they still fit.
TBLIS_HIP_PACK_CHUNKS=wide or =narrow forces one shape, unset ("auto") the
small problems here take the finest one-round chunks, for each kernel family
that reads pack chunks (FP64 pack, split-BF16 pack, the FP64-accurate mode's
row-maximum and digit passes), against the oracle."""

import numpy as np
import pytest

import tblis_amd as T
from tests.test_gpu_parity import FP64_TOL, _dense_case, _dpd_case

pytestmark = pytest.mark.gpu

SPLIT_TOL = 1e-6
FAMILIES = [(None, np.float64, FP64_TOL), ("bf16x3", np.float32, SPLIT_TOL), ("ozaki", np.float64, FP64_TOL)]


@pytest.mark.parametrize("shape", ["wide", "narrow", "auto"])
@pytest.mark.parametrize("config,dtype,tol", FAMILIES)
def test_forced_pack_chunks(shape, config, dtype, tol, monkeypatch):
    if shape != "auto":
        monkeypatch.setenv("TBLIS_HIP_PACK_CHUNKS", shape)
    monkeypatch.setenv("TBLIS_HIP_DIRECT", "0")  # every operand through the pack
    T.lib().tblis_hip_clear_plan_cache()
    try:
        rng = np.random.default_rng(41000)
        # k = 7 * 9 = 63 (uneven k groups), ragged rows
        _dense_case(rng, "aebf", "fjei", "aibj", dict(a=13, b=11, e=9, f=7, i=5, j=14),
                    ["col", "row", "col"], dtype=dtype, config=config, tol=tol)
        # long k (several chunks along k per job), row-major B
        _dense_case(rng, "abef", "efij", "abij", dict(a=24, b=7, e=19, f=11, i=33, j=3),
                    ["col", "row", "random"], dtype=dtype, config=config, tol=tol)
        # DPD: many short-k blocks, the case the wide shape is for
        ext = {c: [6, 4, 3, 5, 2, 7, 1, 4] for c in "abefij"}
        _dpd_case(rng, 8, "abef", "efij", "abij", ext, 0, 0, 5, alpha=0.5, beta=1.0, mode=0,
                  config=config, dtype=dtype, tol=tol)
    finally:
        monkeypatch.delenv("TBLIS_HIP_PACK_CHUNKS", raising=False)
        monkeypatch.delenv("TBLIS_HIP_DIRECT")
        T.lib().tblis_hip_clear_plan_cache()
