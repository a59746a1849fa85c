"""GPU: the factorization captured as a HIP graph (umf_set_graph) against direct launches.

A graph replays the captured level loop with the plan's fixed buffers, so every replay must give
bit-for-bit the factor blocks and solves of a direct-launch factorization of the operator the
plan holds at that moment.  The replays are interleaved with the program's own direct work
(Update, Solve) and synchronizations, the operator changes between replays (an Update writes the
new values where the graph reads them), and a precision change recaptures.  Pivot failures are
reported by every replay.
"""
import numpy as np
import pytest

from oracle_py import coeff_field, laplacian, moddofs

pytestmark = pytest.mark.gpu


def _plan(A, dims, maxs):
    import umf
    return (umf.MF2D if len(dims) == 2 else umf.MF3D)(A, *dims, width=1, maxSideLength=maxs)


def _blocks(G):
    out = []
    for ell in range(G.levels):
        for b in range(min(G.nboxes(ell), 3)):
            for blk in (["A22"] if ell == G.levels - 1 else ["A22", "XR", "S"]):
                out.append(G.front_block(ell, b, blk).copy())
    return out


def _same(a, b):
    return len(a) == len(b) and all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("dims,maxs", [((64, 47), 5), ((15, 15, 15), 3), ((23, 23, 23), 5)])
def test_graph_replays_match_direct_launches(dims, maxs):
    co = coeff_field(dims, "lognormal", seed=21)
    A1 = laplacian(dims, co)
    co2 = co.copy()
    sl = tuple(slice(d // 3, d // 3 + 4) for d in dims)
    co2[sl] *= 30.0
    A2 = laplacian(dims, co2)
    mod = moddofs(A1, A2)
    b = np.random.default_rng(4).standard_normal(A1.n)

    ref1 = _plan(A1, dims, maxs)  # direct launches, operator 1
    ref1.Factor()
    blk1, x1 = _blocks(ref1), ref1.Solve(b)
    ref2 = _plan(A2, dims, maxs)  # direct launches, operator 2
    ref2.Factor()
    blk2, x2 = _blocks(ref2), ref2.Solve(b)

    G = _plan(A1, dims, maxs)
    G.use_graph(True)
    G.Factor()  # capture + first launch
    assert _same(_blocks(G), blk1)
    assert np.array_equal(G.Solve(b), x1)
    G.Factor()  # replay, same operator
    assert _same(_blocks(G), blk1)
    G.Update(A2, mod)  # direct work between replays; the values the graph reads change
    G.Solve(b)
    G.Factor()  # replay on operator 2
    assert _same(_blocks(G), blk2)
    assert np.array_equal(G.Solve(b), x2)
    G.Factor()  # and again
    assert np.array_equal(G.Solve(b), x2)
    G.use_graph(False)
    G.Factor()  # direct launches again
    assert _same(_blocks(G), blk2)


def test_graph_recaptures_after_precision_change():
    dims, maxs = (63, 61), 4
    A = laplacian(dims, coeff_field(dims, "lognormal", seed=3))
    b = np.random.default_rng(5).standard_normal(A.n)
    ref = _plan(A, dims, maxs)
    ref.SetPrecision("fp32", tol=1e-12)
    ref.Factor()
    want = ref.Solve(b)
    G = _plan(A, dims, maxs)
    G.use_graph(True)
    G.Factor()  # FP64 graph
    x64 = G.Solve(b)
    G.SetPrecision("fp32", tol=1e-12)
    G.Factor()  # recaptured in FP32
    assert np.array_equal(G.Solve(b), want)
    G.Factor()
    assert np.array_equal(G.Solve(b), want)
    G.SetPrecision("fp64")
    G.Factor()  # recaptured in FP64
    assert np.array_equal(G.Solve(b), x64)


def test_graph_batch_and_error_flags():
    """A batch plan replays like its direct launches; a non-SPD operator's pivot failure is
    reported by the captured factorization and by every replay (the error flags are cleared
    and set inside the graph)."""
    from test_gpu_batch import MIX, _mk as mk_batch
    import umf
    cos, As, B, Os = mk_batch(MIX[:4])
    RHS = np.random.default_rng(3).standard_normal((B.n, 2))
    B.Factor()
    want = B.Solve(RHS)
    B.use_graph(True)
    for _ in range(3):
        B.Factor()
        assert np.array_equal(B.Solve(RHS), want)

    d = (31, 31)
    bad = laplacian(d, coeff_field(d, "uniform", seed=1))
    bad.val = -bad.val  # negative definite
    G = umf.MF2D(bad, *d, width=1, maxSideLength=3)
    G.use_graph(True)
    for _ in range(3):
        with pytest.raises(ArithmeticError, match="not SPD"):
            G.Factor()
