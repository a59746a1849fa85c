"""GPU: the BF16-split precision form ('bf16x3') against the CPU oracle (FP64).

The form keeps the fronts in FP32 like the 'fp32' form and runs the factorization's Schur and
trailing-update GEMMs (k_gemm) on v_mfma_f32_16x16x32_bf16: every FP32 operand x is split into
two bfloat16 pieces x0 + x1 and the products x0 y0 + x0 y1 + x1 y0 are accumulated in FP32.
The contract is the FP64 one on the refined solution (normwise backward error <= the requested
tolerance, solution within the problem's conditioning of the oracle's).  The forward check is
the rigorous one: x - x* = A^-1 (b - A x), so each refined column must lie within
||A^-1||_2 ||b - A x|| of the oracle's solution (||A^-1||_2 from inverse iteration on the
oracle's FP64 solve).  The factor blocks show that the split product ran: two bfloat16 pieces
carry about 16 significant bits of each operand, so the blocks are less accurate than the FP32
form's (2^-24 per product; worse over the blocks, measured 1.6-10x: the panel GEMMs stay on
FP32 MFMA) yet within the FP32 form's block bound times ten.

The split runs on the levels whose GEMM work lies mostly in fronts of 1024+ pivots; levels of
smaller fronts keep FP32 MFMA, so on a problem without such fronts the form equals 'fp32'.
"""
import numpy as np
import pytest

from metrics import berr
from oracle_py import Factorization, coeff_field, laplacian, moddofs

pytestmark = pytest.mark.gpu

IR_TOL = 1e-12
FWD_TOL = 1e-6    # bound on ||A^-1|| ||b - A x|| / ||x|| at backward error 1e-12
FORMS = ["bf16x3"]

CASES = [
    ((64, 47), 5, "lognormal"),
    ((127, 127), 8, "lognormal"),
    ((15, 15, 15), 3, "lognormal"),
    ((23, 23, 23), 5, "lognormal"),   # fronts of 1500+ rows: the 128 x 128-tile instances
]


def ids(c):
    return "x".join(map(str, c[0])) + f"-m{c[1]}-{c[2]}"


def _mk(dims, maxs, kind, seed=11):
    import umf
    co = coeff_field(dims, kind, seed=seed)
    A = laplacian(dims, co)
    G = (umf.MF2D if len(dims) == 2 else umf.MF3D)(A, *dims, width=1, maxSideLength=maxs)
    O = Factorization(A, dims, 1, maxs, "orc")
    return co, A, G, O


def relerr(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def inv_norm(A, O, iters=30):
    """||A^-1||_2 of the SPD operator: Rayleigh quotient of A^-1 after inverse iterations."""
    x = np.random.default_rng(0).standard_normal(A.n)
    x /= np.linalg.norm(x)
    q = 0.0
    for _ in range(iters):
        y = O.solve(x)
        q = float(np.dot(x, y))
        x = y / np.linalg.norm(y)
    return q


def check_forward(A, O, X, B):
    """Each column within ||A^-1|| ||b - A x|| (relative) of the oracle's solve, that bound
    itself small."""
    ainv = inv_norm(A, O)
    X = X.reshape(A.n, -1)
    B = B.reshape(A.n, -1)
    for j in range(X.shape[1]):
        xs = O.solve(B[:, j])
        bound = ainv * np.linalg.norm(B[:, j] - A.matvec(X[:, j])) / np.linalg.norm(xs)
        e = relerr(X[:, j], xs)
        print("column", j, "forward error", e, "bound", bound)
        assert e <= 1.5 * bound + 1e-14
        assert bound <= FWD_TOL


def worst_block(G, O):
    w = 0.0
    for ell in range(O.L):
        for b in range(O.nboxes(ell)):
            for blk in (["A22"] if ell == 0 else ["A22", "XR", "S"]):
                w = max(w, relerr(G.front_block(ell, b, blk), O.box_block(ell, b, blk)))
    return w


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("case", CASES, ids=ids)
def test_split_solve_reaches_fp64_accuracy(case, form):
    dims, maxs, kind = case
    _, A, G, O = _mk(dims, maxs, kind)
    G.SetPrecision(form, tol=IR_TOL, max_iter=100)
    G.Factor()
    O.factor()
    B = np.random.default_rng(3).standard_normal((A.n, 4))
    X = G.Solve(B)
    st = G.stats()
    print(form, dims, "iterations", st["ir_iters"], "berr", st["ir_berr"])
    assert 1 <= st["ir_iters"] <= 100
    assert st["ir_berr"] <= IR_TOL
    assert berr(A, X, B) <= 4 * IR_TOL
    check_forward(A, O, X, B)
    assert st["bytes_front_pool"] > 0


@pytest.mark.parametrize("case", [((27, 27, 27), 5, "lognormal"), CASES[3]], ids=ids)
def test_split_blocks_against_fp32(case):
    """Factor blocks of the two FP32-storage forms against the FP64 oracle: the split product is
    less accurate than FP32 MFMA (the split path really ran) and stays within ten times the FP32
    form's block bound."""
    dims, maxs, kind = case
    _, A, G, O = _mk(dims, maxs, kind)
    O.factor()
    err = {}
    blocks = {}
    for form in ["fp32"] + FORMS:
        G.SetPrecision(form)
        G.Factor()
        err[form] = worst_block(G, O)
        # the root front (1500+ pivots: its level splits) carries the split's error
        err[form + "_root"] = relerr(G.front_block(0, 0, "A22"), O.box_block(0, 0, "A22"))
        blocks[form] = G.front_block(0, 0, "A22").copy()
    print(dims, err)
    assert err["fp32"] <= 1e-3
    assert err["bf16x3"] <= 10 * max(err["fp32"], 1e-3)
    assert err["bf16x3_root"] > err["fp32_root"]  # the split ran there
    assert not np.array_equal(blocks["bf16x3"], blocks["fp32"])


def test_small_front_levels_keep_fp32_mfma():
    """2D 127^2, maxSideLength 8: no front has 1024 pivots, so no level splits and the form's
    factors equal the FP32 form's bit for bit."""
    dims, maxs, kind = CASES[1]
    _, A, G, O = _mk(dims, maxs, kind)
    blocks = {}
    for form in ["fp32", "bf16x3"]:
        G.SetPrecision(form)
        G.Factor()
        blocks[form] = [G.front_block(ell, b, "A22").copy() for ell in range(O.L) for b in range(O.nboxes(ell))]
    assert all(np.array_equal(a, b) for a, b in zip(blocks["fp32"], blocks["bf16x3"]))


@pytest.mark.parametrize("form", FORMS)
def test_split_update_then_solve(form):
    dims, maxs = (64, 47), 5
    co, A, G, O = _mk(dims, maxs, "lognormal")
    G.SetPrecision(form, tol=IR_TOL, max_iter=100)
    G.Factor()
    O.factor()
    co2 = co.copy()
    co2[20:26, 10:16] *= 50.0
    A2 = laplacian(dims, co2)
    mod = moddofs(A, A2)
    G.Update(A2, mod)
    O.update(A2, mod)
    for ell in range(O.L):
        for which in (0, 1, 2):
            assert np.array_equal(G.dirty(ell, which), O.trace(ell, which)), (ell, which)
    b = np.random.default_rng(9).standard_normal(A.n)
    x = G.Solve(b)
    assert berr(A2, x, b) <= 4 * IR_TOL
    check_forward(A2, O, x, b)


def test_split_matches_fp64_at_scale():
    """3D 63^3 (root front ~12k rows, the 128 x 128-tile Schur and trailing updates at full
    size).  No oracle at this size: FP64 Solve(Ax) = x, each split form reaches the requested
    backward error and agrees with the FP64 solve."""
    import umf
    dims, maxs = (63, 63, 63), 8
    A = laplacian(dims, coeff_field(dims, "lognormal", seed=6))
    G = umf.MF3D(A, *dims, width=1, maxSideLength=maxs)
    G.Factor()
    rng = np.random.default_rng(2)
    x = rng.standard_normal((A.n, 2))
    B = np.stack([A.matvec(x[:, j]) for j in range(2)], axis=1)
    X64 = G.Solve(B)
    assert relerr(X64, x) <= 1e-7
    for form in FORMS:
        G.SetPrecision(form, tol=IR_TOL, max_iter=100)
        G.Factor()
        X = G.Solve(B)
        st = G.stats()
        print(form, "iterations", st["ir_iters"], "berr", st["ir_berr"])
        assert berr(A, X, B) <= 4 * IR_TOL
        assert relerr(X, X64) <= 1e-7


def test_split_batch_refined():
    """The forest (many independent problems) plan in the bf16x3 form, refined to 1e-12."""
    import umf
    from test_gpu_batch import MIX, _mk as mk_batch
    cos, As, B, Os = mk_batch(MIX[:5])
    B.SetPrecision("bf16x3", tol=1e-12, max_iter=100)
    B.Factor()
    RHS = np.random.default_rng(3).standard_normal((B.n, 2))
    X = B.Solve(RHS)
    st = B.stats()
    print("iterations", st["ir_iters"], "berr", st["ir_berr"])
    assert st["ir_berr"] <= 1e-12
    for O in Os:
        O.factor()
    for p, (O, A, Xp, Rp) in enumerate(zip(Os, As, B.split(X), B.split(RHS))):
        print("problem", p)
        check_forward(A, O, Xp[:, :1], Rp[:, :1])
    assert umf.PRECISIONS["bf16x3"] == 2
