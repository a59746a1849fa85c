"""GPU: batches of independent problems (one plan of the block-diagonal operator) against the
oracle, problem by problem: the solution of each problem's rows within 1e-10 of the oracle's
solve of that problem alone, after Factor and after an Update of one problem; the FP32 form's
refined solve; a non-positive pivot reported with its problem.
"""
import numpy as np
import pytest

from oracle_py import Factorization, coeff_field, laplacian, moddofs

pytestmark = pytest.mark.gpu

SOLVE_TOL = 1e-10
MIX = [((31, 31), 3, "uniform"), ((40, 33), 5, "lognormal"), ((9, 9, 9), 3, "lognormal"),
       ((63, 61), 4, "lognormal"), ((11, 14, 9), 2, "lognormal"), ((127, 127), 8, "lognormal")]


def relerr(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _mk(cases, seed=5):
    import umf
    cos = [coeff_field(d, k, seed=seed + i) for i, (d, m, k) in enumerate(cases)]
    As = [laplacian(d, co) for (d, m, k), co in zip(cases, cos)]
    B = umf.Batch([(A, d, 1, m) for A, (d, m, k) in zip(As, cases)])
    Os = [Factorization(A, d, 1, m, "orc") for A, (d, m, k) in zip(As, cases)]
    return cos, As, B, Os


def test_batch_solve_apply_match_oracle():
    cos, As, B, Os = _mk(MIX)
    B.Factor()
    for O in Os:
        O.factor()
    rng = np.random.default_rng(1)
    RHS = rng.standard_normal((B.n, 3))
    X = B.Solve(RHS)
    Y = B.Apply(RHS)
    for q, (O, Xq, Yq, Rq) in enumerate(zip(Os, B.split(X), B.split(Y), B.split(RHS))):
        for j in range(3):
            assert relerr(Xq[:, j], O.solve(Rq[:, j])) <= SOLVE_TOL, (q, j)
            assert relerr(Yq[:, j], O.apply(Rq[:, j])) <= SOLVE_TOL, (q, j)


def test_batch_update_one_problem():
    cos, As, B, Os = _mk(MIX)
    B.Factor()
    for O in Os:
        O.factor()
    q = 3
    dims = MIX[q][0]
    co2 = cos[q].copy()
    co2[20:25, 30:35] *= 40.0
    A2 = laplacian(dims, co2)
    mod = moddofs(As[q], A2)
    B.Update(q, A2, mod)
    Os[q].update(A2, mod)
    st = B.stats()
    assert 0 < st["n_fronts_update"] < st["n_fronts"]  # only problem q's dirty fronts
    rng = np.random.default_rng(2)
    RHS = rng.standard_normal((B.n, 2))
    X = B.Solve(RHS)
    for p, (O, Xp, Rp) in enumerate(zip(Os, B.split(X), B.split(RHS))):
        for j in range(2):
            assert relerr(Xp[:, j], O.solve(Rp[:, j])) <= SOLVE_TOL, (p, j)


def test_batch_fp32_refined():
    cos, As, B, Os = _mk(MIX[:5])
    B.SetPrecision("fp32", tol=1e-12)
    B.Factor()
    rng = np.random.default_rng(3)
    RHS = rng.standard_normal((B.n, 2))
    X = B.Solve(RHS)
    st = B.stats()
    print("iterations", st["ir_iters"], "berr", st["ir_berr"])
    assert st["ir_berr"] <= 1e-12
    for O in Os:
        O.factor()
    for p, (O, Xp, Rp) in enumerate(zip(Os, B.split(X), B.split(RHS))):
        assert relerr(Xp[:, 0], O.solve(Rp[:, 0])) <= 1e-8, p


def test_batch_non_spd_names_problem():
    import umf
    d = (31, 31)
    A = laplacian(d, coeff_field(d, "uniform", seed=1))
    bad = laplacian(d, coeff_field(d, "uniform", seed=1))
    bad.val = -bad.val  # negative definite
    B = umf.Batch([(A, d, 1, 3), (A, d, 1, 3), (bad, d, 1, 3)])
    with pytest.raises(ArithmeticError, match="problem 2"):
        B.Factor()
