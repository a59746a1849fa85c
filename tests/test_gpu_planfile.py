"""GPU: plan files (umf_plan_save / umf_plan_load).  A restored plan holds the same factors, so
Solve / Apply / factor blocks are bit-identical to the saved plan's, after Factor, after an
Update (the file carries the current operator values), in the FP32 form, and for a plan saved
before factoring (restored unfactored, then factored).  The restored plan's solves are also
checked against the oracle's factorization of the same operator (relative 2-norm <= 1e-10, the
FP64 form's parity bound), so the round trip is pinned to the reference's arithmetic, not only
to the plan it came from."""
import numpy as np
import pytest

from oracle_py import Factorization, coeff_field, laplacian, moddofs


def _oracle_solve(A, dims, maxs, B):
    O = Factorization(A, dims, 1, maxs, "orc")
    O.factor()
    return np.stack([O.solve(B[:, j]) for j in range(B.shape[1])], axis=1)


def _rel(x, y):
    return np.linalg.norm(x - y) / np.linalg.norm(y)

pytestmark = pytest.mark.gpu


def _mk(dims, maxs, seed=5):
    import umf
    co = coeff_field(dims, "lognormal", seed=seed)
    A = laplacian(dims, co)
    G = (umf.MF2D if len(dims) == 2 else umf.MF3D)(A, *dims, width=1, maxSideLength=maxs)
    return co, A, G


@pytest.mark.parametrize("dims,maxs", [((31, 31), 3), ((15, 15, 15), 3)])
def test_saved_plan_solves_identically(tmp_path, dims, maxs):
    import umf
    co, A, G = _mk(dims, maxs)
    G.Factor()
    B = np.random.default_rng(1).standard_normal((A.n, 3))
    p = str(tmp_path / "plan.umf")
    G.save(p)
    H = umf.load_plan(p)
    assert H.dims == G.dims and H.n == G.n
    assert np.array_equal(H.Solve(B), G.Solve(B))
    assert _rel(H.Solve(B), _oracle_solve(A, dims, maxs, B)) <= 1e-10
    assert np.array_equal(H.Apply(B), G.Apply(B))
    for blk in ("A22", "XR", "S"):
        assert np.array_equal(H.front_block(1, 0, blk), G.front_block(1, 0, blk)), blk
    # after an Update the file carries the new operator values and factors
    co2 = co.copy()
    co2[tuple(slice(3, 9) for _ in dims)] *= 30.0
    A2 = laplacian(dims, co2)
    mod = moddofs(A, A2)
    G.Update(A2, mod)
    G.save(p)
    H2 = umf.load_plan(p)
    assert np.array_equal(H2.Solve(B), G.Solve(B))
    assert _rel(H2.Solve(B), _oracle_solve(A2, dims, maxs, B)) <= 1e-10
    # and the restored plan updates like the original
    A3 = laplacian(dims, co)
    G.Update(A3, mod)
    H2.Update(A3, mod)
    assert np.array_equal(H2.Solve(B), G.Solve(B))


def test_saved_fp32_form_and_unfactored_plan(tmp_path):
    import umf
    co, A, G = _mk((40, 33), 5)
    b = np.random.default_rng(2).standard_normal(A.n)
    p = str(tmp_path / "u.umf")
    G.save(p)  # before any Factor
    H = umf.load_plan(p)
    with pytest.raises(RuntimeError):
        H.Solve(b)
    H.Factor()
    G.Factor()
    assert np.array_equal(H.Solve(b), G.Solve(b))
    G.SetPrecision("fp32", tol=1e-12)
    G.Factor()
    x = G.Solve(b)
    G.save(p)
    H32 = umf.load_plan(p)
    assert np.array_equal(H32.Solve(b), x)
    assert H32.stats()["ir_iters"] == G.stats()["ir_iters"]
    # the BF16-split form stores FP32 fronts too; its plan file restores the form
    G.SetPrecision("bf16x3", tol=1e-12, max_iter=100)
    G.Factor()
    x = G.Solve(b)
    G.save(p)
    Hbf = umf.load_plan(p)
    assert np.array_equal(Hbf.Solve(b), x)
    Hbf.Factor()  # refactors in the restored form
    assert np.array_equal(Hbf.Solve(b), x)


def test_batch_plans_are_not_saved(tmp_path):
    import umf
    d = (31, 31)
    A = laplacian(d, coeff_field(d, "uniform"))
    B = umf.Batch([(A, d, 1, 3), (A, d, 1, 3)])
    with pytest.raises(ValueError, match="batch plans"):
        B.save(str(tmp_path / "b.umf"))
