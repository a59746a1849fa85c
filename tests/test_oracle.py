"""CPU: the oracle (oracle/liboracle.so) pinned against the reference and the mathematics.

1. Golden fixtures written by the reference itself (tests/golden/*.npz): tree structure, DOF
   lists and their order, dirty sets of Factor and Update must be bit-exact; factor blocks,
   Solve and Apply within 1e-12 (both sides are LU with partial pivoting, MKL vs the oracle's
   loops).
2. The live reference harness (oracle/_ref, marker `ref`) on further sizes and shapes.
3. Closed forms / invariants: the 5- and 7-point stencils for a = 1, Solve(Ax) = x,
   Apply(x) = Ax, the root pivot block equals the dense Schur complement of A onto the root
   separator, Update(M2) equals a fresh factorization of M2, S = -A12 A22^{-1} A21.
4. Sensitivity: each comparison rejects a planted mistake (a swapped DOF, a transposed block,
   a dropped Schur term).
"""
import glob
import os

import numpy as np
import pytest

from oracle_py import Factorization, coeff_field, laplacian, moddofs, Csr

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


def load(path):
    z = np.load(path, allow_pickle=False)
    return {k: z[k] for k in z.files}


def operator(g, second=False):
    dims = tuple(int(v) for v in g["dims"])
    w = int(g["width"])
    co = g["coeff"].copy()
    if second:
        lo, size = g["update_lo"], int(g["update_size"])
        co[tuple(slice(int(l0), int(l0) + size) for l0 in lo)] *= float(g["update_factor"])
    return dims, w, laplacian(dims, co, width=w)


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def dof_lists(F):
    out = {}
    for ell in range(F.L):
        for b in range(F.nboxes(ell)):
            for w in (0, 1):
                out[(0, ell, b, w)] = F.box_dofs(ell, b, w)
        if ell > 0:
            for e in range(F.nseps(ell)):
                out[(1, ell, e, 0)] = F.sep_dofs(ell, e)
    return out


def golden_lists(g, prefix="dofs"):
    data, off, key = g[prefix], g[prefix + "_off"], g[prefix + "_key"]
    return {tuple(int(v) for v in key[i]): data[off[i]:off[i + 1]] for i in range(len(key))}


def check_lists(mine, ref):
    assert set(mine) == set(ref)
    for k in ref:
        assert np.array_equal(mine[k], ref[k]), k


def check_trace(F, g, prefix):
    ref = golden_lists(g, prefix)
    for (ell, w), v in ref.items():
        assert np.array_equal(F.trace(ell, w), v), (prefix, ell, w)


def check_blocks(F, g, prefix, tol=1e-12):
    n = 0
    for k in g:
        if not k.startswith(prefix + "blk_") and not k.startswith(prefix + "sep_"):
            continue
        parts = k[len(prefix):].split("_")
        if parts[0] == "blk":
            ell, b, name = int(parts[1]), int(parts[2]), parts[3]
            M = F.box_block(ell, b, name)
        else:
            M = F.sep_block(int(parts[1]), int(parts[2]))
        assert M.shape == g[k].shape, k
        assert rel(M, g[k]) <= tol, (k, rel(M, g[k]))
        n += 1
    assert n > 0


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_oracle_matches_reference_fixture(path):
    g = load(path)
    dims, w, A = operator(g)
    if "rp" in g:  # the generator itself, bit-exact against the reference's buildLaplacian
        assert np.array_equal(A.rp, g["rp"]) and np.array_equal(A.ci, g["ci"]) and np.array_equal(A.val, g["val"])
    F = Factorization(A, dims, w, int(g["max_side"]), "orc")
    assert F.L == int(g["levels"])
    assert [F.leaf_width(a) for a in range(len(dims))] == list(g["leaf_width"])
    F.factor()
    check_lists(dof_lists(F), golden_lists(g))
    check_trace(F, g, "trace_f")
    check_blocks(F, g, "f_")
    assert rel(F.solve(g["b"]), g["solve_b"]) <= 1e-12
    assert rel(F.apply(g["x"]), g["apply_x"]) <= 1e-12
    _, _, A2 = operator(g, second=True)
    mod = moddofs(A, A2)
    assert np.array_equal(mod, g["mod"])
    F.update(A2, mod)
    check_trace(F, g, "trace_u")
    check_blocks(F, g, "u_")
    b2 = A2.matvec(g["x"])
    assert rel(F.solve(b2), g["solve_b2"]) <= 1e-12
    assert rel(F.apply(g["x"]), g["apply2_x"]) <= 1e-12


LIVE = [((31, 31), 3, "uniform", 1), ((64, 47), 5, "lognormal", 1), ((57, 23), 3, "rand", 2),
        ((33, 40), 5, "lognormal", 3), ((11, 11, 11), 3, "lognormal", 1), ((13, 9, 10), 2, "rand", 1),
        ((12, 12, 12), 2, "rand", 2), ((20, 17, 19), 3, "lognormal", 1)]


@pytest.mark.ref
@pytest.mark.parametrize("case", LIVE, ids=[f"{c[0]}-m{c[1]}-w{c[3]}" for c in LIVE])
def test_oracle_matches_live_reference(case):
    dims, maxs, kind, w = case
    co = coeff_field(dims, kind, seed=9, width=w)
    A = laplacian(dims, co, width=w)
    Ar = laplacian(dims, co, width=w, impl="ref")
    assert np.array_equal(A.val, Ar.val) and np.array_equal(A.ci, Ar.ci)
    O = Factorization(A, dims, w, maxs, "orc")
    R = Factorization(A, dims, w, maxs, "ref")
    O.factor()
    R.factor()
    check_lists(dof_lists(O), dof_lists(R))
    for ell in range(O.L):
        for which in (0, 1, 2):
            assert np.array_equal(O.trace(ell, which), R.trace(ell, which))
    for ell in range(1, O.L):
        for b in range(O.nboxes(ell)):
            for blk in ("A22", "XR", "XL", "S", "A11post"):
                assert rel(O.box_block(ell, b, blk), R.box_block(ell, b, blk)) <= 1e-12
    x = np.random.default_rng(0).uniform(-1, 1, A.n)
    assert rel(O.solve(A.matvec(x)), R.solve(A.matvec(x))) <= 1e-12
    co2 = co.copy()
    co2[tuple(slice(1, 4) for _ in dims)] *= 6.0
    A2 = laplacian(dims, co2, width=w)
    mod = moddofs(A, A2)
    O.update(A2, mod)
    R.update(A2, mod)
    for ell in range(O.L):
        for which in (0, 1, 2):
            assert np.array_equal(O.trace(ell, which), R.trace(ell, which))
    assert rel(O.solve(A2.matvec(x)), R.solve(A2.matvec(x))) <= 1e-12


# ---------------------------------------------------------------------------- closed forms
def test_stencil_closed_forms():
    h = 0.25
    A = laplacian((5, 4), coeff_field((5, 4), "uniform"), h=h)
    d = A.to_scipy().toarray()
    assert np.all(np.diag(d) == 4.0 / h ** 2)
    off = d - np.diag(np.diag(d))
    assert set(np.unique(off)) <= {0.0, -1.0 / h ** 2}
    assert np.allclose(d, d.T) and (off != 0).sum(axis=1).max() == 4
    A3 = laplacian((4, 3, 5), coeff_field((4, 3, 5), "uniform"), h=1.0)
    d3 = A3.to_scipy().toarray()
    assert np.all(np.diag(d3) == 6.0) and (d3 - np.diag(np.diag(d3)) != 0).sum(axis=1).max() == 6
    # DOF id i + j*Nx: the x-neighbour of DOF 0 is DOF 1, the y-neighbour is DOF Nx
    assert d[0, 1] == -1.0 / h ** 2 and d[0, 5] == -1.0 / h ** 2 and d[0, 6] == 0.0


@pytest.mark.parametrize("dims,maxs", [((15, 15), 3), ((9, 9, 9), 3), ((19, 12), 3)])
def test_solve_apply_invariants(dims, maxs):
    A = laplacian(dims, coeff_field(dims, "lognormal", seed=3))
    F = Factorization(A, dims, 1, maxs, "orc")
    F.factor()
    x = np.random.default_rng(1).standard_normal(A.n)
    b = A.matvec(x)
    assert rel(F.solve(b), x) <= 1e-12
    assert rel(F.apply(x), b) <= 1e-13


@pytest.mark.parametrize("dims,maxs", [((15, 15), 3), ((9, 9, 9), 3)])
def test_root_is_dense_schur_complement(dims, maxs):
    A = laplacian(dims, coeff_field(dims, "lognormal", seed=5))
    F = Factorization(A, dims, 1, maxs, "orc")
    F.factor()
    D = A.to_scipy().toarray()
    root = F.box_dofs(0, 0, 0)
    rest = np.setdiff1d(np.arange(A.n), root)
    S = D[np.ix_(root, root)] - D[np.ix_(root, rest)] @ np.linalg.solve(D[np.ix_(rest, rest)], D[np.ix_(rest, root)])
    assert rel(F.box_block(0, 0, "A22"), S) <= 1e-12


def test_front_blocks_are_exact_schur_pieces():
    dims = (31, 31)
    A = laplacian(dims, coeff_field(dims, "lognormal", seed=6))
    F = Factorization(A, dims, 1, 3, "orc")
    F.factor()
    for ell, b in [(F.L - 1, 0), (F.L - 1, 5), (1, 2)]:
        A22, XR, XL, S = (F.box_block(ell, b, k) for k in ("A22", "XR", "XL", "S"))
        A21 = A22 @ XR
        assert rel(S, -(XL @ A22) @ XR) <= 1e-12
        assert rel(XL, XR.T) <= 1e-12  # symmetric operator
        assert rel(XR, np.linalg.solve(A22, A21)) <= 1e-12


@pytest.mark.parametrize("dims,maxs", [((31, 31), 3), ((11, 11, 11), 3)])
def test_update_equals_fresh_factor(dims, maxs):
    co = coeff_field(dims, "lognormal", seed=2)
    A = laplacian(dims, co)
    co2 = co.copy()
    co2[tuple(slice(4, 8) for _ in dims)] *= 20
    A2 = laplacian(dims, co2)
    U = Factorization(A, dims, 1, maxs, "orc")
    U.factor()
    U.update(A2, moddofs(A, A2))
    N = Factorization(A2, dims, 1, maxs, "orc")
    N.factor()
    for ell in range(1, U.L):
        owner = {}
        for e in range(U.nseps(ell)):
            for d in U.sep_dofs(ell, e):
                owner[int(d)] = e
        for b in range(U.nboxes(ell)):
            for blk in ("A22", "XR", "S"):
                assert np.array_equal(U.box_block(ell, b, blk), N.box_block(ell, b, blk)), (ell, b, blk)
            # A11post on its diagonal separator blocks is never read (the parent replaces those
            # blocks from the separators, src/DE/mf2d.cpp:427-443); an unmarked neighbour keeps
            # stale values there.  Everything else must agree bit for bit.
            sep = np.array([owner[int(d)] for d in U.box_dofs(ell, b, 1)])
            live = sep[:, None] != sep[None, :]
            assert np.array_equal(U.box_block(ell, b, "A11post")[live], N.box_block(ell, b, "A11post")[live])
    assert np.array_equal(U.box_block(0, 0, "A22"), N.box_block(0, 0, "A22"))


def test_stale_factor_differs_before_update():
    """The reference's update test: Apply before Update is visibly wrong, after it is exact."""
    dims = (31, 31)
    co = coeff_field(dims, "uniform")
    A = laplacian(dims, co)
    co2 = co.copy()
    co2[:8, :8] *= 2
    A2 = laplacian(dims, co2)
    F = Factorization(A, dims, 1, 3, "orc")
    F.factor()
    x = np.random.default_rng(0).uniform(0, 1, A.n)
    bt = A2.matvec(x)
    before = rel(F.apply(x), bt)
    F.update(A2, moddofs(A, A2))
    after = rel(F.apply(x), bt)
    assert before > 1e-3 and after < 1e-14


def test_update_rejects_dof_outside_leaf_grid():
    # 2D 30 x 30, maxSide 3: leaf width 6, last box holds 2 extra points, i/(bw+w) reaches 4
    dims = (30, 30)
    A = laplacian(dims, coeff_field(dims, "uniform"))
    F = Factorization(A, dims, 1, 3, "orc")
    F.factor()
    assert F.leaf_width(0) == 6
    with pytest.raises(ValueError):
        F.update(A, [29])


# ---------------------------------------------------------------------------- sensitivity
def test_checks_reject_planted_mistakes():
    g = load(os.path.join(HERE, "golden", "2d_31_m3.npz"))
    dims, w, A = operator(g)
    F = Factorization(A, dims, w, int(g["max_side"]), "orc")
    F.factor()
    mine = dof_lists(F)
    ref = golden_lists(g)
    # swapped DOF order inside one list
    k = (0, F.L - 1, 3, 1)
    bad = dict(mine)
    bad[k] = mine[k].copy()
    bad[k][[0, 1]] = bad[k][[1, 0]]
    with pytest.raises(AssertionError):
        check_lists(bad, ref)
    # transposed non-symmetric block (XR vs XL^T of a 1-box slice) and a dropped Schur term
    XR = F.box_block(F.L - 1, 0, "XR")
    assert rel(XR[:, ::-1], g[f"f_blk_{F.L - 1}_0_XR"]) > 1e-3
    S = F.box_block(F.L - 1, 0, "S")
    A11 = F.box_block(F.L - 1, 0, "A11post") - S
    assert rel(A11, g[f"f_blk_{F.L - 1}_0_A11post"]) > 1e-3
    # a solve with one level's contribution dropped misses the fixture
    assert rel(g["b"], g["solve_b"]) > 1e-3


def test_update_before_factor_rejected():
    """The oracle refuses an update of an unfactored object (there are no fronts to update)."""
    dims = (31, 31)
    co = coeff_field(dims, "uniform", seed=1)
    A = laplacian(dims, co)
    F = Factorization(A, dims, 1, 3, "orc")
    with pytest.raises(ValueError):
        F.update(A, np.array([5], dtype=np.int32))
