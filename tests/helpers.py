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


def random_indexed(rng, labels, ext, dense_ndim, fill=0.5, factors=(1.0, 0.5, 0.0)):
    """Host description of a random indexed tensor over `labels` (the first
    dense_ndim are dense), like the reference's random_tensor for
    indexed_varray (test/test.cxx:294-306): a random subset of the index
    values (each kept with probability `fill`, at least one), entries
    uniform in [-1, 1], factors drawn from `factors`."""
    import itertools
    dense_len = [ext[c] for c in labels[:dense_ndim]]
    idx_len = [ext[c] for c in labels[dense_ndim:]]
    every = list(itertools.product(*[range(n) for n in idx_len]))
    pick = [ix for ix in every if rng.random() < fill]
    if not pick and every:
        pick = [every[int(rng.integers(len(every)))]]
    entries = [rng.uniform(-1, 1, dense_len) for _ in pick]
    facs = [float(rng.choice(factors)) for _ in pick]
    return dict(entries=entries, indices=pick, factors=facs, dense_len=dense_len, idx_len=idx_len)


def indexed_brute_force(alpha, A, idx_A, B, idx_B, beta, C_, idx_C, conj_A=False, conj_B=False):
    """The indexed contraction by its definition, one entry triple and one
    assignment of every label at a time (pure Python, tiny sizes only):
    C entry c gains alpha f_a f_b f_c a[..] b[..] wherever the labels' values
    agree with the three entries' index values; C entries are first scaled
    by the reference's rule (beta == 0: 0; beta != 1: f_c beta)."""
    import itertools
    labels = sorted(set(idx_A) | set(idx_B) | set(idx_C))
    ext = {}
    for X, idx in ((A, idx_A), (B, idx_B), (C_, idx_C)):
        for c, n in zip(idx, list(X["dense_len"]) + list(X["idx_len"])):
            ext[c] = n
    cx = any(np.iscomplexobj(x) for X in (A, B, C_) for x in list(X["entries"]) + list(X["factors"]))
    cx = cx or np.iscomplexobj(alpha) or np.iscomplexobj(beta)
    dt = np.complex128 if cx else np.float64
    out = []
    for e, f in zip(C_["entries"], C_["factors"]):
        e = np.array(e, dt)
        if beta == 0.0 or (beta != 1.0 and f == 0.0):
            e[...] = 0.0
        elif beta != 1.0:
            e *= f * beta
        out.append(e)
    if alpha == 0.0:
        return out

    def split(X, idx, vals):
        nd = len(X["dense_len"])
        return tuple(vals[c] for c in idx[:nd]), tuple(vals[c] for c in idx[nd:])

    for vs in itertools.product(*[range(ext[c]) for c in labels]):
        vals = dict(zip(labels, vs))
        da, ia = split(A, idx_A, vals)
        db, ib = split(B, idx_B, vals)
        dc, ic = split(C_, idx_C, vals)
        for ea, xa, fa in zip(A["entries"], A["indices"], A["factors"]):
            if tuple(xa) != ia:
                continue
            for eb, xb, fb in zip(B["entries"], B["indices"], B["factors"]):
                if tuple(xb) != ib:
                    continue
                for k, (xc, fc) in enumerate(zip(C_["indices"], C_["factors"])):
                    if tuple(xc) != ic:
                        continue
                    va = np.conj(ea[da]) if conj_A else ea[da]
                    vb = np.conj(eb[db]) if conj_B else eb[db]
                    out[k][dc] += alpha * fa * fb * fc * va * vb
    return out


def random_indexed_dpd(rng, labels, lens, dense_ndim, irrep, nirrep, layout=5, fill=0.8,
                       factors=(1.0, 0.5, -0.75, 2.0)):
    """Host description of a random indexed DPD tensor (the reference's
    random_tensor for indexed_dpd_varray, test/test.cxx:312-351): lens maps a
    label to its nirrep per-irrep lengths; each indexed dimension gets a
    random irrep; entries are DPD storage of the dense part."""
    import itertools
    from oracle import pyoracle as po
    nd = dense_ndim
    L = [list(lens[c]) for c in labels]
    idx_irrep = [int(rng.integers(nirrep)) for _ in labels[nd:]]
    dirrep = irrep
    for r in idx_irrep:
        dirrep ^= r
    size = po.dpd_size(po.dpd_init(dirrep, nirrep, L[:nd], layout)) if nd else 1
    ranges = [range(L[nd + j][r]) for j, r in enumerate(idx_irrep)]
    every = list(itertools.product(*ranges))
    pick = [ix for ix in every if rng.random() < fill]
    if not pick and every:
        pick = [every[int(rng.integers(len(every)))]]
    entries = [rng.uniform(-1, 1, max(size, 1)) for _ in pick]
    facs = [float(rng.choice(factors)) for _ in pick]
    return dict(entries=entries, indices=pick, factors=facs, irrep=irrep, nirrep=nirrep, lens=L,
                dense_ndim=nd, idx_irrep=idx_irrep, layout=layout)
