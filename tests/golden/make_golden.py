"""Writes the golden fixtures in tests/golden/*.npz from the reference itself.

Runs only where the reference harness oracle/_ref/libumfref.so has been built from
/root/reference (make -C oracle).  Each fixture holds data only: the inputs (grid, coefficient
field, right-hand side, update) and the reference's outputs (tree structure, DOF lists, dirty
sets, selected factor blocks, Solve and Apply results).  The reference publishes no golden
vectors of its own (its tests print, test_src/DE/testMF2D.cpp:183-189), so these pin the
oracle and the library at small sizes.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from oracle_py import Factorization, coeff_field, have_ref, laplacian, moddofs  # noqa: E402

# name: (dims, maxSideLength, width, field, seed, update block (lo, size, factor) in nodal coords)
CASES = {
    "c1_2d_127_m8": ((127, 127), 8, 1, "uniform", 0, ((60, 60), 8, 10.0)),   # the 2D 127^2 config
    "2d_31_m3": ((31, 31), 3, 1, "uniform", 0, ((2, 2), 5, 3.0)),
    "2d_23x57_m3": ((23, 57), 3, 1, "lognormal", 1, ((10, 30), 6, 0.1)),
    "2d_40x33_m5_w2": ((40, 33), 5, 2, "rand", 2, ((3, 20), 5, 4.0)),
    "3d_9_m3": ((9, 9, 9), 3, 1, "rand", 3, ((1, 1, 1), 4, 2.0)),
    "3d_15_m3": ((15, 15, 15), 3, 1, "lognormal", 4, ((6, 6, 6), 4, 5.0)),
    "3d_11x14x9_m2": ((11, 14, 9), 2, 1, "lognormal", 5, ((3, 8, 2), 3, 7.0)),
}

CSR_KEPT = {"2d_31_m3", "2d_40x33_m5_w2", "3d_9_m3"}


def lists(F):
    """All DOF lists of the tree, concatenated, with offsets."""
    data, off, key = [], [0], []
    for ell in range(F.L):
        for b in range(F.nboxes(ell)):
            for w in (0, 1):
                v = F.box_dofs(ell, b, w)
                data.append(v); off.append(off[-1] + len(v)); key.append((0, ell, b, w))
        if ell > 0:
            for e in range(F.nseps(ell)):
                v = F.sep_dofs(ell, e)
                data.append(v); off.append(off[-1] + len(v)); key.append((1, ell, e, 0))
    return (np.concatenate(data).astype(np.int32), np.array(off, np.int64), np.array(key, np.int32))


def traces(F):
    data, off, key = [], [0], []
    for ell in range(F.L):
        for w in (0, 1, 2):
            v = F.trace(ell, w)
            data.append(v); off.append(off[-1] + len(v)); key.append((ell, w))
    return (np.concatenate(data).astype(np.int32) if data else np.zeros(0, np.int32), np.array(off, np.int64),
            np.array(key, np.int32))


def pick_blocks(F):
    """A few blocks of modest size: first and last leaf box, one box per upper level, root."""
    out = {}
    L = F.L
    sel = [(L - 1, 0), (L - 1, F.nboxes(L - 1) - 1)] + [(ell, F.nboxes(ell) // 2) for ell in range(1, L - 1)]
    for ell, b in sel:
        for blk in ("A22", "XR", "XL", "S", "A11post"):
            M = F.box_block(ell, b, blk)
            if M.size <= 12000:
                out[f"blk_{ell}_{b}_{blk}"] = M
    R = F.box_block(0, 0, "A22")
    if R.size <= 12000:
        out["blk_0_0_A22"] = R
    e = F.nseps(L - 1) // 2
    out[f"sep_{L - 1}_{e}"] = F.sep_block(L - 1, e)
    return out


def make(name, spec):
    dims, maxs, w, kind, seed, (lo, size, fac) = spec
    co = coeff_field(dims, kind, seed=seed, width=w)
    A = laplacian(dims, co, width=w, impl="ref")
    F = Factorization(A, dims, w, maxs, "ref")
    F.factor()
    rng = np.random.default_rng(100 + seed)
    x = rng.uniform(-1.0, 1.0, A.n)
    b = A.matvec(x)
    d = {
        "dims": np.array(dims, np.int32), "max_side": np.int32(maxs), "width": np.int32(w),
        "coeff": co,
        "levels": np.int32(F.L), "leaf_width": np.array([F.leaf_width(a) for a in range(len(dims))], np.int32),
        "x": x, "b": b, "solve_b": F.solve(b), "apply_x": F.apply(x),
    }
    d["dofs"], d["dofs_off"], d["dofs_key"] = lists(F)
    d["trace_f"], d["trace_f_off"], d["trace_f_key"] = traces(F)
    if name in CSR_KEPT:  # pins the operator generator; elsewhere it is rebuilt from `coeff`
        d.update({"rp": A.rp, "ci": A.ci, "val": A.val})
    for k, v in pick_blocks(F).items():
        d["f_" + k] = v
    co2 = co.copy()
    co2[tuple(slice(l0, l0 + size) for l0 in lo)] *= fac
    A2 = laplacian(dims, co2, width=w, impl="ref")
    mod = moddofs(A, A2)
    F.update(A2, mod)
    b2 = A2.matvec(x)
    d.update({"update_lo": np.array(lo, np.int32), "update_size": np.int32(size), "update_factor": fac,
              "mod": mod, "solve_b2": F.solve(b2), "apply2_x": F.apply(x)})
    d["trace_u"], d["trace_u_off"], d["trace_u_key"] = traces(F)
    for k, v in pick_blocks(F).items():
        d["u_" + k] = v
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **d)
    print(f"{name}: n={A.n} levels={F.L} -> {os.path.getsize(path) / 1024:.0f} KiB")


if __name__ == "__main__":
    if not have_ref():
        sys.exit("reference harness oracle/_ref/libumfref.so is not built (make -C oracle)")
    for name, spec in CASES.items():
        make(name, spec)
