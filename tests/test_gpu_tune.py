"""GPU: the GEMM-stage tuning knobs (umf_set_tune) and the autotuner (umf_autotune).

The tile size of the trailing and Schur GEMMs does not change any result: every entry still
accumulates the same products in the same order, so 64- and 128-tiles give bit-identical
factors.  The outer column-group width reorders the trailing updates, so its results agree to
rounding.  The autotuner picks one of its candidates and leaves a factorization that solves
like the oracle's.
"""
import numpy as np
import pytest

from oracle_py import Factorization, coeff_field, laplacian

pytestmark = pytest.mark.gpu


def _mk(dims=(23, 23, 23), maxs=5, seed=4):
    import umf
    A = laplacian(dims, coeff_field(dims, "lognormal", seed=seed))
    G = umf.MF3D(A, *dims, width=1, maxSideLength=maxs)
    return A, G


def _blocks(G):
    return [G.front_block(ell, 0, blk).copy() for ell in range(G.levels - 1) for blk in ("A22", "XR", "S")]


def relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_tile_size_does_not_change_results(prec):
    A, G = _mk()
    G.SetPrecision(prec, tol=1e-12)
    b = np.random.default_rng(1).standard_normal(A.n)
    out = []
    for tile in (64, 128, 0):
        G.set_tune(1024, tile)
        G.Factor()
        out.append((_blocks(G), G.Solve(b)))
    for blk, x in out[1:]:
        assert all(np.array_equal(p, q) for p, q in zip(blk, out[0][0]))
        assert np.array_equal(x, out[0][1])


def test_group_width_agrees_to_rounding():
    A, G = _mk()
    b = np.random.default_rng(2).standard_normal(A.n)
    G.Factor()
    x0 = G.Solve(b)
    assert G.tune() == {"gw2": 1024, "tile": 0}
    for gw2 in (256, 2048):
        G.set_tune(gw2, 0)
        G.Factor()
        assert relerr(G.Solve(b), x0) <= 1e-12


def test_autotune_picks_a_candidate_and_solves():
    A, G = _mk(dims=(19, 21, 17), maxs=4, seed=7)
    O = Factorization(A, (19, 21, 17), 1, 4, "orc")
    O.factor()
    res = G.autotune(reps=1)
    print("autotune", res)
    assert res["gw2"] in (1024, 2048) and res["tile"] in (0, 64, 128)
    assert res["factor_ms"] > 0
    b = np.random.default_rng(3).standard_normal(A.n)
    assert relerr(G.Solve(b), O.solve(b)) <= 1e-10  # the plan holds the chosen settings' factors
    G.SetPrecision("fp32", tol=1e-12)
    res = G.autotune(reps=1)
    assert res["tile"] in (0, 64, 128)
    x = G.Solve(b)
    assert relerr(x, O.solve(b)) <= 1e-8


def test_autotune_with_graph_and_failures():
    """A captured factorization is recaptured after the autotuner changed the settings (the
    replay equals a direct factorization with them); on an operator that fails every candidate
    the settings are left as they were."""
    import umf
    A, G = _mk(dims=(19, 21, 17), maxs=4, seed=7)
    b = np.random.default_rng(5).standard_normal(A.n)
    G.use_graph(True)
    G.Factor()  # captured with the defaults
    res = G.autotune(reps=1)
    G.Factor()  # a replay of the settings the autotuner chose
    x = G.Solve(b)
    _, H = _mk(dims=(19, 21, 17), maxs=4, seed=7)
    H.set_tune(res["gw2"], res["tile"])
    H.Factor()
    assert np.array_equal(x, H.Solve(b))

    d = (31, 31)
    bad = laplacian(d, coeff_field(d, "uniform", seed=1))
    bad.val = -bad.val  # negative definite
    B = umf.MF2D(bad, *d, width=1, maxSideLength=3)
    B.set_tune(2048, 128)
    with pytest.raises(ArithmeticError):
        B.autotune(reps=1)
    assert B.tune() == {"gw2": 2048, "tile": 128}
