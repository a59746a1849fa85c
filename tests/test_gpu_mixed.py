"""GPU: the mixed precision form 'fp32t64' (UMF_PREC_FP32T64): the top three separator levels
stored and factored in FP64, the levels below in FP32, the Solve refined in FP64 as in the FP32
form.  Against the CPU oracle (FP64) and the library's own FP32 and FP64 forms:

- the FP32 levels are computed exactly as in the FP32 form, so their blocks equal that form's
  bit for bit; a plan of at most three levels is FP64 throughout, so its blocks equal the FP64
  form's bit for bit;
- the top levels' blocks (FP64 arithmetic on the FP32 levels' Schur complements) are within the
  FP32 form's bound of the oracle's;
- the refined solve reaches the requested backward error with no more iterations than the FP32
  form, each column within ||A^-1||_2 ||b - A x|| of the oracle's solution;
- Update equals a fresh factorization of the new operator bit for bit; Apply agrees with the
  FP64 form to FP32 accuracy; the captured factorization equals the direct one.
"""
import numpy as np
import pytest

from metrics import berr
from oracle_py import Factorization, coeff_field, laplacian, moddofs

pytestmark = pytest.mark.gpu

IR_TOL = 1e-12
TOP = 3  # levels 0..2 in FP64


def _mk(dims, maxs, seed=11):
    import umf
    co = coeff_field(dims, "lognormal", seed=seed)
    A = laplacian(dims, co)
    G = (umf.MF2D if len(dims) == 2 else umf.MF3D)(A, *dims, width=1, maxSideLength=maxs)
    O = Factorization(A, dims, 1, maxs, "orc")
    return co, A, G, O


def _blocks(G, O):
    out = {}
    for ell in range(O.L):
        for b in range(O.nboxes(ell)):
            for blk in (["A22"] if ell == 0 else ["A22", "XR", "S"]):
                out[(ell, b, blk)] = G.front_block(ell, b, blk).copy()
    return out


def _relerr(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _inv_norm(O, n, iters=30):
    x = np.random.default_rng(0).standard_normal(n)
    x /= np.linalg.norm(x)
    q = 0.0
    for _ in range(iters):
        y = O.solve(x)
        q = float(np.dot(x, y))
        x = y / np.linalg.norm(y)
    return q


CASES = [((255, 255), 8), ((31, 31, 31), 2)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])) + f"-m{c[1]}")
def test_mixed_levels_and_refined_solve(case):
    dims, maxs = case
    _, A, G, O = _mk(dims, maxs)
    O.factor()
    assert O.L > TOP  # the case has FP32 levels
    G.SetPrecision("fp32", tol=IR_TOL, max_iter=100)
    G.Factor()
    b32 = _blocks(G, O)
    B = np.random.default_rng(3).standard_normal((A.n, 4))
    G.Solve(B)
    it32 = G.stats()["ir_iters"]
    G.SetPrecision("fp32t64", tol=IR_TOL, max_iter=100)
    G.Factor()
    bm = _blocks(G, O)
    for key, v in bm.items():
        ell = key[0]
        if ell >= TOP:
            assert np.array_equal(v, b32[key]), key
        else:
            ref = O.box_block(*key)
            assert _relerr(v, ref) <= 1e-3, key
    X = G.Solve(B)
    st = G.stats()
    print(dims, "iterations fp32", it32, "fp32t64", st["ir_iters"], "berr", st["ir_berr"])
    assert st["ir_berr"] <= IR_TOL
    assert 1 <= st["ir_iters"] <= it32
    assert berr(A, X, B) <= 4 * IR_TOL
    ainv = _inv_norm(O, A.n)
    for j in range(4):
        xs = O.solve(B[:, j])
        bound = ainv * np.linalg.norm(B[:, j] - A.matvec(X[:, j])) / np.linalg.norm(xs)
        assert _relerr(X[:, j], xs) <= 1.5 * bound + 1e-14
        assert bound <= 1e-6
    # Apply: the mixed factors against the FP64 form's
    x = np.random.default_rng(5).standard_normal(A.n)
    ym = G.Apply(x)
    G.SetPrecision("fp64")
    G.Factor()
    y64 = G.Apply(x)
    assert _relerr(ym, y64) <= 1e-5


def test_three_levels_are_fp64():
    """2D 63^2, maxSideLength 8: three levels, all FP64: the blocks equal the FP64 form's."""
    _, A, G, O = _mk((63, 63), 8)
    assert O.L <= TOP
    G.Factor()
    b64 = _blocks(G, O)
    G.SetPrecision("fp32t64", tol=IR_TOL, max_iter=100)
    G.Factor()
    bm = _blocks(G, O)
    assert all(np.array_equal(bm[k], b64[k]) for k in b64)
    B = np.random.default_rng(4).standard_normal((A.n, 2))
    X = G.Solve(B)
    assert G.stats()["ir_berr"] <= IR_TOL
    assert berr(A, X, B) <= 4 * IR_TOL


def test_mixed_update_equals_factor_and_graph():
    """Factor(A1) -> Update(A2) -> Solve equals Factor(A2) -> Solve bit for bit; the captured
    factorization replays the direct one."""
    dims, maxs = (255, 255), 8
    co, A, G, O = _mk(dims, maxs)
    G.SetPrecision("fp32t64", tol=IR_TOL, max_iter=100)
    G.Factor()
    co2 = co.copy()
    co2[100:110, 40:52] *= 30.0
    A2 = laplacian(dims, co2)
    mod = moddofs(A, A2)
    G.Update(A2, mod)
    B = np.random.default_rng(8).standard_normal((A.n, 3))
    Xu = G.Solve(B)
    bu = _blocks(G, O)
    G.Factor()  # the plan holds A2
    Xf = G.Solve(B)
    assert np.array_equal(Xu, Xf)
    assert all(np.array_equal(bu[k], v) for k, v in _blocks(G, O).items())
    assert berr(A2, Xf, B) <= 4 * IR_TOL
    G.use_graph(True)
    for _ in range(2):  # capture, then replay
        G.Factor()
        assert np.array_equal(G.Solve(B), Xf)
    G.use_graph(False)
