"""hipGraph capture of library calls: a prepared dense, packed split-BF16,
DPD, batched, complex, indexed, indexed DPD and FP64-accurate (int8 digit)
contraction is
captured once and replayed
several times,
with the inputs changed by eager work between replays; every replay must
match an eager call on the same inputs (FP64: 1e-12, split-BF16: 1e-6
relative Frobenius).  The plan cache is cleared after capture to show that
captured plans stay alive."""
import numpy as np
import pytest
import torch

import tblis_amd as T

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def colmaj(lens):
    s, a = [], 1
    for l in lens:
        s.append(a)
        a *= l
    return s


def rel(x, ref):
    x, ref = x.double(), ref.double()
    return float(torch.linalg.vector_norm(x - ref) / torch.linalg.vector_norm(ref))


def _capture(fn):
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    T.set_stream(side.cuda_stream)
    T.set_async(True)
    try:
        with torch.cuda.stream(side):
            fn()
        side.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=side):
            fn()
    finally:
        T.set_async(False)
        T.set_stream(torch.cuda.current_stream().cuda_stream)
    return graph


def _replay_check(graph, inputs, outputs, eager, tol):
    """Replay three times; before each, scale the inputs eagerly, then
    compare every output against an eager call into fresh buffers."""
    T.lib().tblis_hip_clear_plan_cache()
    for rep in range(3):
        for x in inputs:
            x.mul_(1.5 if rep % 2 == 0 else -0.75)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        got = [o.clone() for o in outputs]
        want = eager()
        torch.cuda.synchronize()
        for g, w in zip(got, want):
            err = rel(g, w)
            assert err <= tol, (rep, err)


def _dense(dtype, ext, ia, ib, ic, config=None, seed=0):
    g = torch.Generator(device=DEV).manual_seed(seed)
    L = lambda s: [ext[c] for c in s]
    mk = lambda s: torch.rand(int(np.prod(L(s))), dtype=dtype, device=DEV, generator=g) * 2 - 1
    A, B, C = mk(ia), mk(ib), mk(ic)
    v = lambda x, s: torch.as_strided(x, L(s), colmaj(L(s)))

    def eager():
        Ce = torch.zeros_like(C)
        T.mult(0.5, v(A, ia), ia, v(B, ib), ib, 0.0, v(Ce, ic), ic, config=config)
        return [Ce]

    fn = T.prepare_mult(0.5, v(A, ia), ia, v(B, ib), ib, 0.0, v(C, ic), ic, config=config)
    return fn, [A, B], [C], eager


def test_graph_dense_fp64():
    fn, ins, outs, eager = _dense(torch.float64, dict(a=24, b=20, e=16, f=18, i=12, j=10),
                                  "aebf", "fjei", "aibj")
    _replay_check(_capture(fn), ins, outs, eager, 1e-12)


def test_graph_packed_split_bf16():
    # large enough for the packed panel path (pack kernels + workspace)
    ext = dict(a=24, b=16, e=20, f=16, i=24, j=16)
    fn, ins, outs, eager = _dense(torch.float32, ext, "aebf", "fjei", "aibj", seed=1)
    _replay_check(_capture(fn), ins, outs, eager, 1e-6)


def test_graph_ozaki():
    """The FP64-accurate mode (row maxima cleared by a fill kernel, digit
    planes, int8 GEMM): replays equal eager calls exactly (the scheme is
    deterministic), dense and DPD (C panels and the unpack pass)."""
    fn, ins, outs, eager = _dense(torch.float64, dict(a=24, b=20, e=16, f=18, i=12, j=10),
                                  "aebf", "fjei", "aibj", config="ozaki", seed=4)
    _replay_check(_capture(fn), ins, outs, eager, 0.0)
    nirrep, ext = 4, [5, 3, 4, 2]
    lens = [ext] * 4
    n = T.dpd_size(0, nirrep, lens)
    g = torch.Generator(device=DEV).manual_seed(5)
    A = torch.rand(n, dtype=torch.float64, device=DEV, generator=g) * 2 - 1
    B = torch.rand(n, dtype=torch.float64, device=DEV, generator=g) * 2 - 1
    C = torch.rand(n, dtype=torch.float64, device=DEV, generator=g) * 2 - 1
    C0 = C.clone()

    def eager_dpd():
        Ce = C0.clone()
        T.dpd_mult(1.0, A, 0, "abef", B, 0, "efij", 0.5, Ce, 0, "abij", nirrep, lens, lens, lens,
                   config="ozaki")
        return [Ce]

    def fn_dpd():
        C.copy_(C0)
        T.dpd_mult(1.0, A, 0, "abef", B, 0, "efij", 0.5, C, 0, "abij", nirrep, lens, lens, lens,
                   config="ozaki")

    _replay_check(_capture(fn_dpd), [A, B], [C], eager_dpd, 0.0)


def test_graph_dpd():
    nirrep, ext = 4, [5, 3, 4, 2]
    lens = [ext] * 4
    n = T.dpd_size(0, nirrep, lens)
    g = torch.Generator(device=DEV).manual_seed(2)
    A = torch.rand(n, dtype=torch.float64, device=DEV, generator=g) * 2 - 1
    B = torch.rand(n, dtype=torch.float64, device=DEV, generator=g) * 2 - 1
    C = torch.zeros(n, dtype=torch.float64, device=DEV)

    def eager():
        Ce = torch.zeros_like(C)
        T.dpd_mult(1.0, A, 0, "abef", B, 0, "efij", 0.0, Ce, 0, "abij", nirrep, lens, lens, lens)
        return [Ce]

    fn = T.prepare_dpd_mult(1.0, A, 0, "abef", B, 0, "efij", 0.0, C, 0, "abij", nirrep, lens, lens,
                            lens)
    _replay_check(_capture(fn), [A, B], [C], eager, 1e-12)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_graph_batch(dtype):
    rng = np.random.default_rng(3)
    g = torch.Generator(device=DEV).manual_seed(3)
    As, Bs, Cs = [], [], []
    for _ in range(64):
        a, b, e, f, i, j = (int(x) for x in rng.integers(2, 13, 6))
        mk = lambda sh: torch.as_strided(
            torch.rand(int(np.prod(sh)), dtype=dtype, device=DEV, generator=g) * 2 - 1, sh, colmaj(sh))
        As.append(mk([a, b, e, f]))
        Bs.append(mk([e, f, i, j]))
        Cs.append(mk([a, b, i, j]))
    n = len(As)

    def eager():
        Ce = [torch.zeros_like(c) for c in Cs]
        T.mult_batch(1.0, As, ["abef"] * n, Bs, ["efij"] * n, 0.0, Ce, ["abij"] * n)
        return Ce

    fn = T.prepare_batch(1.0, As, ["abef"] * n, Bs, ["efij"] * n, 0.0, Cs, ["abij"] * n)
    tol = 1e-12 if dtype == torch.float64 else 1e-6
    _replay_check(_capture(fn), As + Bs, Cs, eager, tol)


def test_graph_complex_alpha_and_beta():
    """Complex contraction with a complex alpha and beta: the C passes before
    and after the GEMMs are kernels, captured with them."""
    g = torch.Generator(device=DEV).manual_seed(11)
    ext = dict(a=20, b=12, e=16, f=6, i=9, j=14)
    L = lambda s: [ext[c] for c in s]
    mk = lambda s: torch.complex(torch.rand(int(np.prod(L(s))), dtype=torch.float64, device=DEV, generator=g),
                                 torch.rand(int(np.prod(L(s))), dtype=torch.float64, device=DEV, generator=g))
    A, B, C0 = mk("aebf"), mk("fjei"), mk("aibj")
    C = C0.clone()
    v = lambda x, s: torch.as_strided(x, L(s), colmaj(L(s)))
    alpha, beta = 0.5 - 0.25j, 0.75 + 0.5j

    def eager():
        Ce = C0.clone()
        T.mult(alpha, v(A, "aebf"), "aebf", v(B, "fjei"), "fjei", beta, v(Ce, "aibj"), "aibj")
        return [Ce]

    fn = T.prepare_mult(alpha, v(A, "aebf"), "aebf", v(B, "fjei"), "fjei", beta, v(C, "aibj"), "aibj")
    graph = _capture(fn)
    # each replay starts from the same C (beta != 0 reads it)
    T.lib().tblis_hip_clear_plan_cache()
    for rep in range(3):
        A.mul_(1.5 if rep % 2 == 0 else -0.75)
        C.copy_(C0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        got = C.clone()
        want = eager()[0]
        torch.cuda.synchronize()
        assert rel(torch.view_as_real(got), torch.view_as_real(want)) <= 1e-12, rep


def test_graph_indexed_with_beta_pass():
    """Indexed contraction with beta = 0.5 (beta pass kernel, then GEMMs)."""
    g = torch.Generator(device=DEV).manual_seed(12)
    n = 10 * 9 * 8
    sa = torch.rand(3 * n, dtype=torch.float64, device=DEV, generator=g)
    sb = torch.rand(3 * 8 * 7 * 6, dtype=torch.float64, device=DEV, generator=g)
    sc0 = torch.rand(10 * 9 * 7 * 6, dtype=torch.float64, device=DEV, generator=g)
    sc = sc0.clone()
    A = T.Indexed([torch.as_strided(sa, [10, 9, 8], colmaj([10, 9, 8]), i * n) for i in range(3)],
                  [(0,), (2,), (3,)], [4], [1.0, -0.5, 2.0])
    B = T.Indexed([torch.as_strided(sb, [8, 7, 6], colmaj([8, 7, 6]), i * 336) for i in range(3)],
                  [(0,), (1,), (3,)], [4], [0.5, 1.0, 1.0])
    Cx = T.Indexed([torch.as_strided(sc, [10, 9, 7, 6], colmaj([10, 9, 7, 6]))], [()], [], [0.75])

    def eager():
        ce = sc0.clone()
        Ce = T.Indexed([torch.as_strided(ce, [10, 9, 7, 6], colmaj([10, 9, 7, 6]))], [()], [], [0.75])
        T.indexed_mult(1.25, A, "abef", B, "eijf", 0.5, Ce, "abij")
        return [ce]

    fn = T.prepare_indexed_mult(1.25, A, "abef", B, "eijf", 0.5, Cx, "abij")
    graph = _capture(fn)
    T.lib().tblis_hip_clear_plan_cache()
    for rep in range(3):
        sa.mul_(1.5 if rep % 2 == 0 else -0.75)
        sc.copy_(sc0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        got = sc.clone()
        want = eager()[0]
        torch.cuda.synchronize()
        assert rel(got, want) <= 1e-12, rep


def test_graph_indexed_dpd_blocked():
    """Indexed DPD contraction (the blocked algorithm: beta pass, then one
    GEMM launch over the C blocks) captured and replayed."""
    from tests.helpers import random_indexed_dpd
    rng = np.random.default_rng(15)
    lens = {"a": [3, 2, 1, 2], "b": [2, 2, 1, 3], "e": [2, 3, 1, 2], "i": [1, 2, 3, 2], "j": [2, 1, 2, 3],
            "f": [3, 2, 2, 1]}
    A0 = random_indexed_dpd(rng, "abef", lens, 3, 1, 4)
    B0 = random_indexed_dpd(rng, "eijf", lens, 3, 2, 4)
    C0 = random_indexed_dpd(rng, "abij", lens, 4, 3, 4)

    def store(X):
        n = len(X["entries"][0])
        s = torch.zeros(max(len(X["entries"]), 1) * n, dtype=torch.float64, device=DEV)
        for i, e in enumerate(X["entries"]):
            s[i * n:(i + 1) * n] = torch.from_numpy(np.asarray(e)).to(DEV)
        return s, n

    def tensor(X, s, n):
        return T.IndexedDpd([s[i * n:(i + 1) * n] for i in range(len(X["entries"]))], X["indices"],
                            X["idx_irrep"], X["irrep"], X["nirrep"], X["lens"], X["dense_ndim"], X["factors"],
                            X["layout"], dtype=torch.float64)

    (sa, na), (sb, nb), (sc0, nc) = store(A0), store(B0), store(C0)
    sc = sc0.clone()
    A, B, Cx = tensor(A0, sa, na), tensor(B0, sb, nb), tensor(C0, sc, nc)

    def eager():
        ce = sc0.clone()
        T.indexed_dpd_mult(0.75, A, "abef", B, "eijf", 0.5, tensor(C0, ce, nc), "abij")
        return [ce]

    graph = _capture(lambda: T.indexed_dpd_mult(0.75, A, "abef", B, "eijf", 0.5, Cx, "abij"))
    T.lib().tblis_hip_clear_plan_cache()
    for rep in range(3):
        sa.mul_(1.5 if rep % 2 == 0 else -0.75)
        sc.copy_(sc0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        got = sc.clone()
        want = eager()[0]
        torch.cuda.synchronize()
        assert rel(got, want) <= 1e-12, rep
