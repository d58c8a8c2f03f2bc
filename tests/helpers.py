"""Shared helpers for the tests (shape generators, strided layouts)."""
import numpy as np


def col_major(lens):
    s, acc = [], 1
    for l in lens:
        s.append(acc)
        acc *= l
    return s


def row_major(lens):
    s, acc = [0] * len(lens), 1
    for i in range(len(lens) - 1, -1, -1):
        s[i] = acc
        acc *= lens[i]
    return s


def permuted_layout(lens, perm):
    """Strides of a dense tensor whose dims are stored in the order perm
    (perm[0] fastest)."""
    s, acc = [0] * len(lens), 1
    for d in perm:
        s[d] = acc
        acc *= lens[d]
    return s


def random_contraction(rng, max_dims=3, max_len=4, allow_batch=True, allow_diag=False):
    """Random A/B/C label strings over groups AB, AC, BC (and ABC)."""
    letters = list("abcdefghijklmnop")
    rng.shuffle(letters)
    n_ab = int(rng.integers(0, max_dims + 1))
    n_ac = int(rng.integers(0, max_dims + 1))
    n_bc = int(rng.integers(0, max_dims + 1))
    n_abc = int(rng.integers(0, 2)) if allow_batch else 0
    it = iter(letters)
    ab = [next(it) for _ in range(n_ab)]
    ac = [next(it) for _ in range(n_ac)]
    bc = [next(it) for _ in range(n_bc)]
    abc = [next(it) for _ in range(n_abc)]
    ext = {c: int(rng.integers(1, max_len + 1)) for c in ab + ac + bc + abc}
    A = ab + ac + abc
    B = ab + bc + abc
    Cl = ac + bc + abc
    rng.shuffle(A)
    rng.shuffle(B)
    rng.shuffle(Cl)
    if allow_diag and A and rng.random() < 0.3:
        A.append(A[int(rng.integers(0, len(A)))])
    return "".join(A), "".join(B), "".join(Cl), ext


def layout_for(idx, ext, rng, kind="random"):
    lens = [ext[c] for c in idx]
    if kind == "col":
        return lens, col_major(lens)
    if kind == "row":
        return lens, row_major(lens)
    perm = list(range(len(lens)))
    rng.shuffle(perm)
    strides = permuted_layout(lens, perm)
    # occasionally pad the outermost stride so dims are not contiguous
    if lens and rng.random() < 0.3:
        strides = [s * 2 if s == max(strides) and s > 0 else s for s in strides]
    return lens, strides


def storage_size(lens, strides):
    if any(l == 0 for l in lens):
        return 1
    return 1 + sum((l - 1) * s for l, s in zip(lens, strides))
