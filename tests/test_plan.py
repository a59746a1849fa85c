"""CPU: the library's host side -- planner, dirty sets, operator generator, C ABI exports.

The planner runs in libumf_hip.so without touching the GPU (umf_plan_analyze); its tree,
DOF lists (with order), separator numbering and dirty sets must equal the oracle's and the
reference fixtures bit for bit.
"""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

from oracle_py import Factorization, coeff_field, laplacian, moddofs

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLDEN = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))

CASES = [((31, 31), 3, 1), ((23, 57), 3, 1), ((127, 127), 8, 1), ((40, 33), 5, 2), ((64, 47), 5, 1),
         ((9, 9, 9), 3, 1), ((15, 15, 15), 3, 1), ((11, 14, 9), 2, 1), ((17, 13, 15), 2, 2), ((31, 31, 31), 3, 1),
         ((30, 31), 3, 1), ((19, 14, 17), 2, 3)]


def analyzed(A, dims, w, maxs):
    import umf
    cls = umf.MF2D if len(dims) == 2 else umf.MF3D
    return cls(A, *dims, width=w, maxSideLength=maxs, analyze_only=True)


def test_library_exports_every_declared_symbol():
    import umf
    hdr = open(os.path.join(ROOT, "include", "umf.h")).read()
    declared = set(re.findall(r"^\s*(?:int|void\*|const char\*)\s+(umf_\w+)\s*\(", hdr, re.M))
    assert declared, "no declarations parsed"
    lib = ctypes.CDLL(umf.LIB_PATH)
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    assert declared <= set(umf.EXPORTED) | {"umf_stream"}, declared - set(umf.EXPORTED)


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-m{c[1]}-w{c[2]}" for c in CASES])
def test_planner_matches_oracle(case):
    dims, maxs, w = case
    co = coeff_field(dims, "lognormal", seed=2, width=w)
    A = laplacian(dims, co, width=w)
    O = Factorization(A, dims, w, maxs, "orc")
    O.factor()
    P = analyzed(A, dims, w, maxs)
    assert P.levels == O.L
    assert [P.leaf_width(a) for a in range(len(dims))] == [O.leaf_width(a) for a in range(len(dims))]
    for ell in range(O.L):
        assert P.nboxes(ell) == O.nboxes(ell)
        for b in range(O.nboxes(ell)):
            for which in (0, 1):
                assert np.array_equal(P.box_dofs(ell, b, which), O.box_dofs(ell, b, which)), (ell, b, which)
        if ell > 0:
            assert P.nseps(ell) == O.nseps(ell)
            for e in range(O.nseps(ell)):
                assert np.array_equal(P.sep_dofs(ell, e), O.sep_dofs(ell, e)), (ell, e)
        for which in (0, 1, 2):
            assert np.array_equal(P.dirty(ell, which), O.trace(ell, which)), (ell, which)
    # an update's dirty sets
    rng = np.random.default_rng(len(dims) * 100 + maxs)
    co2 = co.copy()
    lo = [int(rng.integers(0, d - 2)) for d in dims]
    co2[tuple(slice(l0, l0 + 3) for l0 in lo)] *= 5.0
    A2 = laplacian(dims, co2, width=w)
    mod = moddofs(A, A2)
    try:
        O.update(A2, mod)
    except ValueError:  # the DOF-to-leaf map leaves the grid (uneven last box): both reject
        with pytest.raises(ValueError):
            P.analyze_update(mod)
        return
    P.analyze_update(mod)
    for ell in range(O.L):
        for which in (0, 1, 2):
            assert np.array_equal(P.dirty(ell, which), O.trace(ell, which)), ("update", ell, which)


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_planner_matches_reference_fixture(path):
    z = np.load(path, allow_pickle=False)
    dims = tuple(int(v) for v in z["dims"])
    w = int(z["width"])
    A = laplacian(dims, z["coeff"], width=w)
    P = analyzed(A, dims, w, int(z["max_side"]))
    data, off, key = z["dofs"], z["dofs_off"], z["dofs_key"]
    for i in range(len(key)):
        kind, ell, idx, which = (int(v) for v in key[i])
        ref = data[off[i]:off[i + 1]]
        mine = P.box_dofs(ell, idx, which) if kind == 0 else P.sep_dofs(ell, idx)
        assert np.array_equal(mine, ref), (kind, ell, idx, which)
    P.analyze_update(z["mod"])
    data, off, key = z["trace_u"], z["trace_u_off"], z["trace_u_key"]
    for i in range(len(key)):
        ell, which = (int(v) for v in key[i])
        assert np.array_equal(P.dirty(ell, which), data[off[i]:off[i + 1]]), (ell, which)


@pytest.mark.parametrize("dims,w", [((13, 11), 1), ((13, 11), 2), ((13, 11), 3), ((6, 5, 4), 1), ((9, 9, 9), 3)])
def test_generator_matches_oracle(dims, w):
    import umf
    co = coeff_field(dims, "lognormal", seed=4, width=w)
    A = laplacian(dims, co, width=w, h=0.37)
    rp, ci, v = umf.build_laplacian(dims, co, w, 0.37)
    assert np.array_equal(rp, A.rp) and np.array_equal(ci, A.ci) and np.array_equal(v, A.val)


def test_geometry_errors():
    import umf
    dims = (9, 9)
    A = laplacian(dims, coeff_field(dims, "uniform"))
    with pytest.raises(RuntimeError, match="recursion depth is 1 level"):
        umf.MF2D(A, *dims, width=1, maxSideLength=8, analyze_only=True)
    with pytest.raises(ValueError):
        umf.MF2D(A, 9, 8, width=1, maxSideLength=2, analyze_only=True)  # n != Nx*Ny


def test_operator_checks():
    import umf
    dims = (15, 15)
    A = laplacian(dims, coeff_field(dims, "uniform"))
    bad = type(A)(A.rp, A.ci.copy(), A.val)
    bad.ci[1] = bad.ci[0]  # duplicate entry in row 0
    with pytest.raises(ValueError, match="duplicate"):
        umf.MF2D(bad, *dims, width=1, maxSideLength=3, analyze_only=True)


def test_device_calls_on_analyzed_plan_are_refused():
    import umf
    dims = (15, 15)
    A = laplacian(dims, coeff_field(dims, "uniform"))
    P = umf.MF2D(A, *dims, width=1, maxSideLength=3, analyze_only=True)
    with pytest.raises(RuntimeError):
        P.Factor()


def test_flop_counts():
    """F_factor = sum over fronts of n^3/3 + n^2 m + n m^2 (the algorithmic count)."""
    dims = (63, 63)
    A = laplacian(dims, coeff_field(dims, "uniform"))
    P = analyzed(A, dims, 1, 3)
    want = 0.0
    for ell in range(P.levels):
        for b in range(P.nboxes(ell)):
            n = len(P.box_dofs(ell, b, 0))
            m = len(P.box_dofs(ell, b, 1))
            want += n ** 3 / 3 + n * n * m + n * m * m
    assert abs(P.stats()["flops_factor"] - want) <= 1e-9 * want


BATCH = [((31, 31), 3, 1), ((40, 33), 5, 1), ((9, 9, 9), 3, 1), ((63, 61), 4, 1), ((11, 14, 9), 2, 1)]


def _batch(cases, seed=1):
    import umf
    As = [laplacian(d, coeff_field(d, "lognormal", seed=seed + i)) for i, (d, m, w) in enumerate(cases)]
    B = umf.Batch([(A, d, w, m) for A, (d, m, w) in zip(As, cases)], analyze_only=True)
    singles = [analyzed(A, d, w, m) for A, (d, m, w) in zip(As, cases)]
    return As, B, singles


def test_batch_layout_matches_single_plans():
    As, B, singles = _batch(BATCH)
    assert B.nprob == len(BATCH)
    assert B.levels == max(S.levels for S in singles)
    off = 0
    for q, S in enumerate(singles):
        o, n, lv, sh = B.problem(q)
        assert (o, n, lv, sh) == (off, As[q].n, S.levels, B.levels - S.levels)
        off += n
    assert B.n == off
    st, st1 = B.stats(), [S.stats() for S in singles]
    assert st["n_fronts"] == sum(s["n_fronts"] for s in st1)
    assert st["flops_factor"] == pytest.approx(sum(s["flops_factor"] for s in st1), rel=1e-12)
    assert st["bytes_front_pool"] == sum(s["bytes_front_pool"] for s in st1)


def test_batch_dirty_sets_are_the_problems_own():
    """An update touching problems 1 and 3 marks exactly their own dirty sets, shifted to the
    merged (root-aligned) levels and offset by the boxes of the problems before them."""
    As, B, singles = _batch(BATCH)
    rng = np.random.default_rng(3)
    mods = {1: rng.choice(As[1].n, 5, replace=False), 3: rng.choice(As[3].n, 7, replace=False)}
    for q, m in mods.items():
        singles[q].analyze_update(m)
    B.analyze_update(np.concatenate([m + B.offsets[q] for q, m in mods.items()]))
    L = B.levels
    for lam in range(L):
        for which in (0, 1):
            want = []
            for q, S in enumerate(singles):
                ell = lam - (L - S.levels)
                if ell < 0:
                    continue
                base = sum((S2.nboxes(lam - (L - S2.levels)) if which == 0 else
                            (S2.nseps(lam - (L - S2.levels)) if lam - (L - S2.levels) >= 1 else 0))
                           for S2 in singles[:q] if lam - (L - S2.levels) >= 0)
                if q in mods and (which == 0 or ell >= 1):
                    want += [base + b for b in S.dirty(ell, which)]
            assert np.array_equal(np.sort(B.dirty(lam, which)), np.sort(want)), (lam, which)


def test_batch_errors_name_the_problem():
    import umf
    d = (20, 20)
    A = laplacian(d, coeff_field(d, "uniform", seed=1))
    with pytest.raises(RuntimeError, match="problem 1"):
        umf.Batch([(A, d, 1, 3), (A, d, 1, 40)], analyze_only=True)  # one level only
    with pytest.raises(ValueError, match="tree queries"):
        umf.Batch([(A, d, 1, 3)], analyze_only=True).box_dofs(0, 0, 0)


@pytest.mark.parametrize("dims,w", [((21, 18, 19), 1), ((40, 37), 1), ((30, 27), 2), ((14, 13, 12), 3)])
def test_serve_box_mods_match_rebuild(dims, w):
    """The serving loop's host-side modified DOFs of a box change (used for admission before a
    slot runs it) equal the reference's modDOF rule applied to a full rebuild of the changed
    field, for every stencil width, boxes at the corners, the far corner and inside, and leave
    the base field and operator untouched."""
    import umf
    from serve import box_mods
    rng = np.random.default_rng(7 + w)
    h = 1.0 / (dims[0] + 1)
    co = np.exp(rng.standard_normal(tuple(n + 2 * w for n in dims)))
    rp, ci, v = umf.build_laplacian(dims, co, w, h)
    co0, v0 = co.copy(), v.copy()
    boxes = [(4, [0] * len(dims)), (6, [d + 2 * w - 6 for d in dims]), (5, [3] * len(dims)), (1, [w] * len(dims))]
    boxes += [(int(s), [int(rng.integers(0, d + 2 * w - s + 1)) for d in dims]) for s in rng.integers(1, 9, size=6)]
    for side, lo in boxes:
        f = float(np.exp(rng.standard_normal()))
        c2 = co.copy()
        c2[tuple(slice(l, l + side) for l in lo)] *= f
        _, _, full = umf.build_laplacian(dims, c2, w, h)
        mod = box_mods(dims, co, v, lo, side, f, rp, ci, h, w)
        assert np.array_equal(mod, np.unique(ci[v != full])), (side, lo)
    assert np.array_equal(co, co0) and np.array_equal(v, v0)
