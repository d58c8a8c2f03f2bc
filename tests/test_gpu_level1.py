"""GPU parity of the level-1 operations tblis_tensor_add / _scale / _set
(src/iface/1t/add.cxx:24-110, scale.cxx:20-55, set.cxx:20-51) against the
oracle's restatement (oracle/oracle.c orc_level1_*): transposes, sums over
labels only A has, replication over labels only B has, diagonals on either
side, padded strides, and the front ends' special cases (alpha = 0 reads no
A, beta = 0 reads no B, empty tensors).  FP64 within 1e-13 relative
Frobenius error (the A-only sums are the only rounding that differs), FP32
within 1e-6; elements outside the views must be left untouched."""
import numpy as np
import pytest
import torch

import tblis_amd as T
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
TOL = {np.float64: 1e-13, np.float32: 1e-6}


def _strides(lens, pad, perm):
    """Strides of a padded layout whose dimension order (fastest first) is perm."""
    s = [0] * len(lens)
    acc = 1
    for d in perm:
        s[d] = acc
        acc *= lens[d] + pad
    return s, max(acc, 1)


def _tensor(rng, lens, dtype, pad=0, perm=None, fill=None):
    perm = list(range(len(lens))) if perm is None else perm
    s, size = _strides(lens, pad, perm)
    host = rng.uniform(-1, 1, size).astype(dtype) if fill is None else np.full(size, fill, dtype)
    dev = torch.from_numpy(host.copy()).to(DEV)
    view = torch.as_strided(dev, tuple(lens), tuple(s)) if lens else torch.as_strided(dev, (), ())
    return host, dev, view, s


def _addressed(lens, s, size):
    """Mask of the storage elements a view of these lengths and strides reaches."""
    mask = np.zeros(size, bool)
    off = np.zeros(1, np.int64)
    for n, st in zip(lens, s):
        off = (off[:, None] + np.arange(n, dtype=np.int64)[None, :] * st).ravel()
    mask[off] = True
    return mask


def _compare(got, want, dtype, before=None, lens=None, s=None):
    got = got.astype(np.float64)
    nrm = np.linalg.norm(want)
    err = np.linalg.norm(got - want) / (nrm if nrm > 0 else 1.0)
    assert err <= TOL[dtype], err
    if before is not None:
        # elements outside the view are left bit-identical
        out = ~_addressed(lens, s, before.size)
        np.testing.assert_array_equal(got[out], before.astype(np.float64)[out])


def _add_case(rng, dtype, lA, iA, lB, iB, alpha, beta, padA=0, padB=0, permA=None, permB=None,
              fillB=None):
    hA, _, vA, sA = _tensor(rng, lA, dtype, padA, permA)
    hB, dB, vB, sB = _tensor(rng, lB, dtype, padB, permB, fill=fillB)
    want = po.level1(0, alpha, hA, lA, sA, iA, beta, hB, lB, sB, iB)
    T.add(alpha, vA, iA, beta, vB, iB)
    torch.cuda.synchronize()
    _compare(dB.cpu().numpy(), want, dtype, hB, lB, sB)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_add_transpose_reduce_replicate(dtype):
    rng = np.random.default_rng(61000)
    # transpose with padding on both sides
    _add_case(rng, dtype, [13, 7, 9], "abc", [9, 13, 7], "cab", 0.5, -1.25, padA=1, padB=2,
              permB=[2, 0, 1])
    # sum over two labels only A has
    _add_case(rng, dtype, [20, 31, 17, 5], "abcd", [17, 20], "ca", 1.5, 1.0, padA=1)
    # replicate over a label only B has
    _add_case(rng, dtype, [11, 6], "ab", [11, 6, 9], "abc", -0.75, 2.0, permB=[2, 1, 0])
    # transpose through LDS with partial tiles and an outer dimension
    _add_case(rng, dtype, [100, 70, 3], "abc", [70, 3, 100], "bca", 1.0, 1.0, padA=1)
    # one wave per B element over three A-only dimensions
    _add_case(rng, dtype, [40, 6, 9, 5], "abcd", [9], "c", 0.5, -1.0, padB=2)
    # long strided A-only sums over few B elements: split into chunks
    _add_case(rng, dtype, [10, 9, 40, 30], "abcd", [9, 10], "ba", 1.0, 0.5)
    _add_case(rng, dtype, [50, 3000], "ab", [50], "a", 2.0, 1.0, padA=3)
    # everything summed into a scalar
    _add_case(rng, dtype, [64, 33, 5], "abc", [], "", 1.0, 0.5)
    # a large plain copy-add (many blocks, grid stride)
    _add_case(rng, dtype, [300, 257, 13], "abc", [300, 257, 13], "abc", 2.0, 1.0)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_add_diagonals_and_unit_lengths(dtype):
    rng = np.random.default_rng(61100)
    # only B's diagonal is written
    _add_case(rng, dtype, [8, 5], "ab", [8, 8, 5], "aab", 1.0, 1.0, padB=1)
    # A's diagonal summed: B[ab] = sum_c A[aacb]
    _add_case(rng, dtype, [6, 6, 4, 3], "aacb", [6, 3], "ab", 0.25, 0.0)
    # a length-1 dimension only in A is ignored
    _add_case(rng, dtype, [17, 1, 9], "axb", [17, 9], "ab", 3.0, -1.0)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_add_special_cases(dtype):
    rng = np.random.default_rng(61200)
    # beta == 0: B is overwritten without being read (NaN does not survive)
    _add_case(rng, dtype, [40, 30], "ab", [30, 40], "ba", 1.0, 0.0, fillB=np.nan)
    # alpha == 0: A is not read (NaN in A does not propagate), B <- beta B
    hA, _, vA, sA = _tensor(rng, [10, 12], dtype, fill=np.nan)
    hB, dB, vB, sB = _tensor(rng, [10, 12], dtype)
    T.add(0.0, vA, "ab", -2.0, vB, "ab")
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dB.cpu().numpy(), (hB * dtype(-2.0)).astype(dtype))
    # an empty A-only sum: B <- beta B
    hA, _, vA, sA = _tensor(rng, [10, 0], dtype)
    hB, dB, vB, sB = _tensor(rng, [10], dtype)
    T.add(1.0, vA, "ab", 3.0, vB, "a")
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dB.cpu().numpy(), (hB * dtype(3.0)).astype(dtype))
    # an empty B (a label only B has, of length 0): nothing is written
    hA, _, vA, sA = _tensor(rng, [4, 5], dtype)
    hB, dB, vB, _ = _tensor(rng, [4, 0, 5], dtype)
    T.add(1.0, vA, "ab", 1.0, vB, "acb")
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dB.cpu().numpy(), hB)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_scale_and_set(dtype):
    rng = np.random.default_rng(61300)
    empty = np.zeros(1, dtype)
    for lens, idx, pad, perm in (([30, 20, 7], "abc", 1, [1, 2, 0]), ([9, 9, 4], "aab", 0, None),
                                 ([513, 129], "ab", 3, None)):
        for alpha in (0.0, 1.0, -1.5):
            h, d, v, s = _tensor(rng, lens, dtype, pad, perm)
            want = po.level1(1, 0.0, empty, [], [], "", alpha, h, lens, s, idx)
            T.scale(alpha, v, idx)
            torch.cuda.synchronize()
            _compare(d.cpu().numpy(), want, dtype, h, lens, s)
        h, d, v, s = _tensor(rng, lens, dtype, pad, perm, fill=np.nan)
        want = po.level1(2, 0.125, empty, [], [], "", 0.0, h, lens, s, idx)
        T.set(0.125, v, idx)
        torch.cuda.synchronize()
        got = d.cpu().numpy().astype(np.float64)
        # set writes exactly the addressed elements; padding keeps its NaN
        np.testing.assert_array_equal(np.isnan(got), np.isnan(want))
        np.testing.assert_array_equal(got[~np.isnan(got)], want[~np.isnan(want)])


def test_add_on_caller_stream():
    """The call runs on the stream set by tblis_hip_set_stream."""
    rng = np.random.default_rng(61400)
    hA, _, vA, sA = _tensor(rng, [64, 64], np.float64)
    hB, dB, vB, sB = _tensor(rng, [64, 64], np.float64)
    want = po.level1(0, 1.0, hA, [64, 64], sA, "ab", 1.0, hB, [64, 64], sB, "ba")
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    T.set_stream(stream.cuda_stream)
    try:
        T.add(1.0, vA, "ab", 1.0, vB, "ba")
    finally:
        T.set_stream(None)
    stream.synchronize()
    _compare(dB.cpu().numpy(), want, np.float64, hB, [64, 64], sB)
