"""GPU: LU fronts for pattern-symmetric, unsymmetric operators against the CPU oracle.

The reference factors every front by LU with partial pivoting (include/LinAlg/dense.hpp:180-227,
src/DE/mf2d.cpp:541-576) and so accepts unsymmetric, pattern-symmetric operators
(include/DE/mf.hpp:9-12).  buildLaplacian with stencil width 2 or 3 and a variable coefficient is
one: it averages different node pairs for (p, p-w) and (p-w, p)
(test_include/DE/buildLaplacian.hpp:64-74 vs :115-124).  Such an operator gets LU fronts here.
Parity is judged on the quantities both factorizations define uniquely (the LU factors and
pivots themselves depend on the elimination order): A22, XL = A12 A22^-1, XR = A22^-1 A21,
S = -A12 XR, A11post and the separator blocks within 1e-11 relative, Solve / Apply within 1e-10,
the dirty sets of Update bit-exact, and Update equal to a fresh factorization bit for bit.
"""
import numpy as np
import pytest

from oracle_py import Factorization, coeff_field, laplacian, moddofs
from test_gpu_parity import FACTOR_TOL, SOLVE_TOL, _compare_blocks, relerr

pytestmark = pytest.mark.gpu

# (dims, maxSideLength, coefficient field, width): variable fields with width 2-3 are unsymmetric
CASES = [
    ((40, 33), 5, "rand", 2),
    ((61, 58), 6, "lognormal", 3),
    ((23, 57), 3, "lognormal", 2),
    ((17, 13, 15), 2, "rand", 2),
    ((14, 13, 12), 3, "lognormal", 3),
]


def ids(c):
    return "x".join(map(str, c[0])) + f"-m{c[1]}-{c[2]}-w{c[3]}"


def _mk(dims, maxs, kind, w, seed=11):
    import umf
    co = coeff_field(dims, kind, seed=seed, width=w)
    A = laplacian(dims, co, width=w)
    D = A.to_scipy()
    assert abs(D - D.T).max() > 0  # the operator is unsymmetric
    G = (umf.MF2D if len(dims) == 2 else umf.MF3D)(A, *dims, width=w, maxSideLength=maxs)
    assert G.lu
    O = Factorization(A, dims, w, maxs, "orc")
    return co, A, G, O


@pytest.mark.parametrize("case", CASES, ids=ids)
def test_lu_blocks_solve_apply_match_oracle(case):
    dims, maxs, kind, w = case
    _, A, G, O = _mk(dims, maxs, kind, w)
    G.Factor()
    O.factor()
    worst = _compare_blocks(G, O)
    print(ids(case), {k: f"{v:.1e}" for k, v in worst.items()})
    for k, v in worst.items():
        assert v <= FACTOR_TOL, (k, v)
    rng = np.random.default_rng(5)
    X = rng.uniform(-1, 1, (A.n, 3))
    B = np.stack([A.matvec(X[:, j]) for j in range(3)], axis=1)
    Xg = G.Solve(B)
    for j in range(3):
        assert relerr(Xg[:, j], O.solve(B[:, j])) <= SOLVE_TOL
        assert relerr(Xg[:, j], X[:, j]) <= 1e3 * SOLVE_TOL  # Solve(Ax) = x
    Yg = G.Apply(X)
    for j in range(3):
        assert relerr(Yg[:, j], O.apply(X[:, j])) <= SOLVE_TOL
        assert relerr(Yg[:, j], B[:, j]) <= 1e-12  # Apply(x) = Ax


@pytest.mark.parametrize("case", CASES[:2] + CASES[3:4], ids=ids)
def test_lu_update_matches_oracle_and_fresh_factor(case):
    dims, maxs, kind, w = case
    co, A, G, O = _mk(dims, maxs, kind, w)
    G.Factor()
    O.factor()
    rng = np.random.default_rng(7)
    co2 = co.copy()
    # the box stays in the lower half: an uneven last leaf box that holds >= 2 extra points would
    # send the reference's leaf map out of range (refused here, UMF_EUPDATEMAP)
    lo = [int(rng.integers(0, d // 2 - 2)) for d in dims]
    co2[tuple(slice(l, l + 4) for l in lo)] *= 9.0
    A2 = laplacian(dims, co2, width=w)
    mod = moddofs(A, A2)
    G.Update(A2, mod)
    O.update(A2, mod)
    for ell in range(O.L):
        for which in (0, 1, 2):
            assert np.array_equal(G.dirty(ell, which), O.trace(ell, which)), (ell, which)
    worst = _compare_blocks(G, O)
    for k, v in worst.items():
        assert v <= FACTOR_TOL, (k, v)
    B = rng.standard_normal((A.n, 2))
    Xu = G.Solve(B)
    for j in range(2):
        assert relerr(Xu[:, j], O.solve(B[:, j])) <= SOLVE_TOL
    G.Factor()  # a fresh factorization of A2: the same, bit for bit
    assert np.array_equal(G.Solve(B), Xu)


def test_lu_on_symmetric_operator_equals_cholesky():
    """LU fronts forced on a symmetric positive definite operator give the Cholesky fronts'
    blocks and solution to rounding (both are exact factorizations of the same matrices)."""
    import umf
    dims = (47, 39)
    A = laplacian(dims, coeff_field(dims, "lognormal", seed=3))
    C = umf.MF2D(A, *dims, width=1, maxSideLength=4)
    G = umf.MF2D(A, *dims, width=1, maxSideLength=4)
    assert not G.lu
    G.set_lu(True)
    assert G.lu
    C.Factor()
    G.Factor()
    for ell in range(G.levels):
        for b in range(G.nboxes(ell)):
            for blk in (["A22"] if ell == 0 else ["A22", "XR", "XL", "S"]):
                assert relerr(G.front_block(ell, b, blk), C.front_block(ell, b, blk)) <= 1e-12, (ell, b, blk)
    B = np.random.default_rng(1).standard_normal((A.n, 4))
    assert relerr(G.Solve(B), C.Solve(B)) <= 1e-12
    G.set_lu(False)  # back to Cholesky: identical to the other plan
    G.Factor()
    assert np.array_equal(G.Solve(B), C.Solve(B))


def test_lu_zero_pivot_reported():
    """A singular pivot block is reported with its front and column (the reference asserts on
    LAPACK info != 0, include/LinAlg/dense.hpp:187)."""
    import umf
    dims = (30, 27)
    A = laplacian(dims, coeff_field(dims, "rand", seed=2, width=2), width=2)
    val = A.val.copy()
    val[:] = 0.0
    val[A.rp[0] + 1] = 1.0  # A(0, 1) = 1, A(1, 0) = 0: asymmetric (LU fronts), the pivots zero
    Z = umf.CSR(A.rp, A.ci, val, A.n, A.n)
    G = umf.MF2D(Z, *dims, width=2, maxSideLength=5)
    assert G.lu
    with pytest.raises(ArithmeticError, match="zero pivot"):
        G.Factor()


def test_lu_refuses_other_forms():
    import umf
    _, A, G, O = _mk((40, 33), 5, "rand", 2)
    with pytest.raises(ValueError, match="FP64"):
        G.SetPrecision("fp32")
    with pytest.raises(ValueError, match="not symmetric"):
        G.set_lu(False)
