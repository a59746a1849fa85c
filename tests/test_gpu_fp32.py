"""GPU: the FP32-MFMA precision form against the CPU oracle (FP64).

The form stores and factors the fronts in FP32 and refines Solve in FP64 (preconditioned
conjugate gradients on the FP64 operator), so its contract is the FP64 one on the solution:
normwise backward error ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) <= the requested tolerance, and the
solution then agrees with the oracle's FP64 solve to within the problem's conditioning.  The
FP32 factor blocks themselves agree with the oracle's to FP32 accuracy; structure (dirty sets)
is the same as in FP64.
"""
import numpy as np
import pytest

from metrics import berr
from oracle_py import Factorization, coeff_field, laplacian, moddofs

pytestmark = pytest.mark.gpu

IR_TOL = 1e-12      # requested backward error of the refined solve
FWD_TOL = 1e-8      # solution vs the FP64 oracle (small test problems: condition numbers <= 1e4)
BLOCK_TOL = 1e-3    # FP32 factor blocks vs the FP64 oracle (normwise per block)

CASES = [
    ((31, 31), 3, "uniform", 1),
    ((64, 47), 5, "lognormal", 1),
    ((127, 127), 8, "lognormal", 1),
    ((15, 15, 15), 3, "lognormal", 1),
    ((11, 14, 9), 2, "lognormal", 1),
]


def ids(c):
    return "x".join(map(str, c[0])) + f"-m{c[1]}-{c[2]}-w{c[3]}"


def _mk(dims, maxs, kind, w, seed=11):
    import umf
    co = coeff_field(dims, kind, seed=seed, width=w)
    A = laplacian(dims, co, width=w)
    G = (umf.MF2D if len(dims) == 2 else umf.MF3D)(A, *dims, width=w, maxSideLength=maxs)
    O = Factorization(A, dims, w, maxs, "orc")
    return co, A, G, O


def relerr(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.mark.parametrize("case", CASES, ids=ids)
def test_fp32_solve_reaches_fp64_accuracy(case):
    dims, maxs, kind, w = case
    _, A, G, O = _mk(dims, maxs, kind, w)
    G.SetPrecision("fp32", tol=IR_TOL, max_iter=60)
    G.Factor()
    O.factor()
    rng = np.random.default_rng(3)
    B = rng.standard_normal((A.n, 5))
    X = G.Solve(B)
    st = G.stats()
    print(dims, "iterations", st["ir_iters"], "berr", st["ir_berr"])
    assert 1 <= st["ir_iters"] <= 60
    assert st["ir_berr"] <= IR_TOL
    assert berr(A, X, B) <= 4 * IR_TOL  # recomputed on the host
    for j in range(B.shape[1]):
        assert relerr(X[:, j], O.solve(B[:, j])) <= FWD_TOL
    # the FP32 front pool is half the FP64 one
    assert st["bytes_front_pool"] > 0


@pytest.mark.parametrize("case", CASES[:3] + CASES[3:4], ids=ids)
def test_fp32_factor_blocks_close_to_oracle(case):
    dims, maxs, kind, w = case
    _, A, G, O = _mk(dims, maxs, kind, w)
    G.SetPrecision("fp32")
    G.Factor()
    O.factor()
    worst = 0.0
    for ell in range(O.L):
        for b in range(O.nboxes(ell)):
            for blk in (["A22"] if ell == 0 else ["A22", "XR", "S"]):
                worst = max(worst, relerr(G.front_block(ell, b, blk), O.box_block(ell, b, blk)))
    print("worst block", worst)
    assert worst <= BLOCK_TOL
    assert worst > 1e-12  # the blocks really are FP32


def test_fp32_update_then_solve():
    dims, maxs = (64, 47), 5
    co, A, G, O = _mk(dims, maxs, "lognormal", 1)
    G.SetPrecision("fp32", tol=IR_TOL)
    G.Factor()
    O.factor()
    rng = np.random.default_rng(9)
    co2 = co.copy()
    co2[20:26, 10:16] *= 50.0
    A2 = laplacian(dims, co2, width=1)
    mod = moddofs(A, A2)
    G.Update(A2, mod)
    O.update(A2, mod)
    for ell in range(O.L):
        for which in (0, 1, 2):
            assert np.array_equal(G.dirty(ell, which), O.trace(ell, which)), (ell, which)
    b = rng.standard_normal(A.n)
    x = G.Solve(b)
    assert berr(A2, x, b) <= 4 * IR_TOL
    assert relerr(x, O.solve(b)) <= FWD_TOL


def test_fp32_apply_is_fp32_accurate():
    dims, maxs = (63, 61), 4
    _, A, G, O = _mk(dims, maxs, "lognormal", 1)
    G.SetPrecision("fp32")
    G.Factor()
    x = np.random.default_rng(1).uniform(-1, 1, A.n)
    y = G.Apply(x)
    e = relerr(y, A.matvec(x))
    print("apply rel err", e)
    assert 1e-12 < e <= 1e-5


def test_fp32_refinement_limit_reported():
    import umf
    dims, maxs = (63, 61), 4
    _, A, G, O = _mk(dims, maxs, "lognormal", 1)
    G.SetPrecision("fp32", tol=1e-15, max_iter=1)
    G.Factor()
    with pytest.raises(umf.RefinementError):
        G.Solve(np.ones(A.n))


def test_precision_switch_needs_factor():
    dims, maxs = (31, 31), 3
    _, A, G, O = _mk(dims, maxs, "uniform", 1)
    G.Factor()
    G.SetPrecision("fp32")
    with pytest.raises(RuntimeError):
        G.Solve(np.ones(A.n))
    G.Factor()
    G.SetPrecision("fp64")
    with pytest.raises(RuntimeError):
        G.Solve(np.ones(A.n))
    G.Factor()
    O.factor()
    b = np.ones(A.n)
    assert relerr(G.Solve(b), O.solve(b)) <= 1e-10


def test_fp32_large_fronts():
    """Fronts of 1500+ rows: the FP32 form's 128 x 128-tile trailing and Schur updates."""
    dims, maxs = (23, 23, 23), 5
    _, A, G, O = _mk(dims, maxs, "lognormal", 1, seed=4)
    G.SetPrecision("fp32", tol=IR_TOL)
    G.Factor()
    O.factor()
    for ell, b, blk in ((0, 0, "A22"), (1, 0, "S"), (1, 5, "XR")):
        e = relerr(G.front_block(ell, b, blk), O.box_block(ell, b, blk))
        print(ell, b, blk, e)
        assert 1e-12 < e <= BLOCK_TOL
    B = np.random.default_rng(8).standard_normal((A.n, 3))
    X = G.Solve(B)
    assert berr(A, X, B) <= 4 * IR_TOL
    for j in range(3):
        assert relerr(X[:, j], O.solve(B[:, j])) <= FWD_TOL


def test_fp32_matches_fp64_at_scale():
    """3D 63^3 (root front ~12k rows): the FP32 form's 128 x 128-tile trailing and Schur
    updates.  No oracle at this size: FP64 Solve(Ax) = x, the FP32 form reaches the requested
    backward error and agrees with the FP64 solve."""
    import umf
    dims, maxs = (63, 63, 63), 8
    co = coeff_field(dims, "lognormal", seed=6)
    A = laplacian(dims, co)
    G = umf.MF3D(A, *dims, width=1, maxSideLength=maxs)
    G.Factor()
    rng = np.random.default_rng(2)
    x = rng.standard_normal((A.n, 2))
    B = np.stack([A.matvec(x[:, j]) for j in range(2)], axis=1)
    X64 = G.Solve(B)
    assert relerr(X64, x) <= 1e-7
    G.SetPrecision("fp32", tol=IR_TOL)
    G.Factor()
    X32 = G.Solve(B)
    st = G.stats()
    print("iterations", st["ir_iters"], "berr", st["ir_berr"])
    assert berr(A, X32, B) <= 4 * IR_TOL
    assert relerr(X32, X64) <= 1e-7
