"""GPU: coefficient-change requests done on the device (umf_box_scale_device) against the host.

The same sequence of box scalings is applied to the padded nodal coefficient field on both
sides.  After each one the host rebuilds the whole operator with the harness generator
(umf.build_laplacian, test_include/DE/buildLaplacian.hpp) and applies the reference's modDOF
rule to it (columns of the entries whose value changed, test_src/DE/testMF2D.cpp:143-150).
The device values must be bit-identical to the rebuild and the modified-DOF sets equal.  The
boxes cover the interior, both padding layers (boxes that touch only padding nodes change only
boundary diagonals), a factor of 1 (no change) and stencil widths 1-3.  For width 1 an Update
with the collected DOFs then solves like a fresh factorization of the changed operator.
"""
import numpy as np
import pytest

from devmem import DevArray
from oracle_py import coeff_field

pytestmark = pytest.mark.gpu

# width-1 grids have even leaf boxes (the reference's modDOF -> leaf box rule maps every DOF,
# which the Update check needs); the wider stencils only check values and DOF sets
CASES = [
    ((31, 31), 1, "lognormal"),
    ((23, 19), 2, "uniform"),
    ((15, 15, 15), 1, "lognormal"),
    ((9, 8, 10), 3, "uniform"),
]


def boxes(dims, w):
    d = len(dims)
    P = [n + 2 * w for n in dims]
    out = [
        ([n // 3 for n in dims], 4, 40.0),              # interior
        ([0] * d, 3, 0.01),                              # touches the low padding
        ([p - 5 for p in P], 5, 7.0),                   # touches the high padding
        ([0] * d, w, 3.0),                               # padding nodes only (corner)
        ([n // 2 for n in dims], 3, 1.0),               # no change
        ([max(n // 2 - 2, 0) for n in dims], 6, 0.5),   # overlaps the first box
    ]
    return out


def ids(c):
    return "x".join(map(str, c[0])) + f"-w{c[1]}-{c[2]}"


@pytest.mark.parametrize("case", CASES, ids=ids)
def test_box_scale_matches_host_rebuild(case):
    import umf
    dims, w, kind = case
    h = 1.0 / (dims[0] + 1)
    co = coeff_field(dims, kind, seed=7, width=w)
    rp, ci, val = umf.build_laplacian(dims, co, w, h)
    A = umf.CSR(rp, ci, val, len(rp) - 1, len(rp) - 1)
    G = (umf.MF2D if len(dims) == 2 else umf.MF3D)(A, *dims, width=w, maxSideLength=3)
    d_co = DevArray(co)
    d_val = DevArray(val)
    cur = val.copy()
    mods = []
    for lo, side, f in boxes(dims, w):
        print("box", lo, side, f, flush=True)
        mod = G.box_scale_device(d_co.data_ptr(), h, lo, side, f, d_val.data_ptr())
        co[tuple(slice(l, l + side) for l in lo)] *= f
        _, _, v2 = umf.build_laplacian(dims, co, w, h)
        got = d_val.numpy()
        assert np.array_equal(got, v2), np.max(np.abs(got - v2))
        assert np.array_equal(d_co.numpy(), co)
        want = np.unique(ci[cur != v2]).astype(np.int32)
        assert np.array_equal(mod, want), (len(mod), len(want))
        if f == 1.0:
            assert len(mod) == 0
        cur = v2
        mods.append(mod)
    if w != 1:
        return
    # width 1: the operator stays symmetric; an Update with the collected DOFs solves like a
    # fresh factorization of the changed operator
    G.Factor()
    mod = np.unique(np.concatenate(mods))
    G.update_device(d_val.data_ptr(), mod)
    F = (umf.MF2D if len(dims) == 2 else umf.MF3D)(umf.CSR(rp, ci, cur, len(rp) - 1, len(rp) - 1), *dims,
                                                    width=w, maxSideLength=3)
    F.Factor()
    b = np.random.default_rng(3).standard_normal(len(rp) - 1)
    x1, x2 = G.Solve(b), F.Solve(b)
    assert np.linalg.norm(x1 - x2) <= 1e-10 * np.linalg.norm(x2)


def test_box_scale_errors():
    import umf
    dims, w = (20, 17), 1
    co = coeff_field(dims, "uniform", width=w)
    rp, ci, val = umf.build_laplacian(dims, co, w, 1.0)
    G = umf.MF2D(umf.CSR(rp, ci, val, len(rp) - 1, len(rp) - 1), *dims, width=w, maxSideLength=3)
    d_co = DevArray(co)
    d_val = DevArray(val)
    with pytest.raises(ValueError, match="outside the padded field"):
        G.box_scale_device(d_co.data_ptr(), 1.0, [19, 0], 4, 2.0, d_val.data_ptr())
    with pytest.raises(ValueError, match="bad h or box side"):
        G.box_scale_device(d_co.data_ptr(), 1.0, [0, 0], 0, 2.0, d_val.data_ptr())
    assert np.array_equal(d_co.numpy(), co)  # refused requests change nothing
