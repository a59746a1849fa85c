"""GPU: the benchmark configurations at their full sizes, through the C ABI.

At these sizes the CPU oracle would take hours, so the checks are the properties the reference
itself relies on, which fix the result without an oracle run:

- Update invariant: Factor(A1) -> Update(A2) -> Solve equals Factor(A2) -> Solve bit for bit
  (the reference's Update re-runs the same per-front code on the fronts whose inputs changed,
  src/DE/mf3d.cpp:208-248; here every front's arithmetic is fixed by the plan, not by the pass);
- Solve(A x) = x for the plain FP64 form (the reference's own check, test_src/DE/testMF3D.cpp),
  and an infinity-norm backward error at the level of FP64 rounding;
- the refined lower-precision forms reach backward error <= 1e-10, and each column's distance
  to the FP64 form's solution is within what the two residuals allow, ||A^-1|| (||r|| + ||r64||)
  (||A^-1||_2 by inverse iteration on the FP64 factorization); at these condition numbers a
  1e-10 backward error allows forward errors of order 1e-5, which the test prints;
- the 256-problem batch (the batch benchmark's own generator): problems up to 127^2 against the
  CPU oracle, every problem by the properties above.

Sizes: 2D 1023^2 (11 GB FP64 front pool), 3D 127^3 (170 GB FP64 + 85 GB FP32, one MI355X's HBM).
"""
import os
import sys

import numpy as np
import pytest

from metrics import berr, forward_within_bound, inv_norm, relerr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

pytestmark = pytest.mark.gpu

ETA_FP64 = 1e-14   # infinity-norm backward error of the FP64 solve
FWD_FP64 = 1e-8    # Solve(A x) vs x, FP64 (lognormal coefficients: condition numbers up to ~1e7)
ETA_IR = 1e-10     # refined forms: the requested backward error


class Op:
    """Host CSR (rp, ci, val) with the oracle's matvec interface."""

    def __init__(self, rp, ci, val):
        from oracle_py import Csr
        self.c = Csr(rp, ci, val)
        self.rp, self.ci, self.val, self.n = self.c.rp, self.c.ci, self.c.val, self.c.n

    def matvec(self, x):
        return self.c.matvec(x)


def _full_size(name, forms=("fp32", "bf16x3")):
    import bench
    import umf
    from devmem import DevArray
    spec = bench.WORKLOADS[name]
    dims = spec["dims"]
    print(f"[{name}] generating {dims}", flush=True)
    rp, ci, v1, v2, mod, B, lo = bench.make_problem(**dict(spec, nrhs=16))
    A1, A2 = Op(rp, ci, v1), Op(rp, ci, v2)
    n = A1.n
    G = (umf.MF2D if len(dims) == 2 else umf.MF3D)(A1.c, *dims, 1, spec["max_side"])
    print(f"[{name}] planned: {G.levels} levels, pool {G.stats()['bytes_front_pool'] / 1e9:.1f} GB", flush=True)
    d_v2 = DevArray(v2)
    rng = np.random.default_rng(17)
    x = rng.standard_normal((n, 4))
    b = np.stack([A2.matvec(x[:, j]) for j in range(4)], axis=1)
    Bc = np.asfortranarray(B.T)  # n x 16
    # FP64: Factor(A1) -> Update(A2) -> Solve against Factor(A2) -> Solve
    G.Factor()
    G.update_device(d_v2.data_ptr(), mod)
    Xu = G.Solve(Bc)
    xu = G.Solve(b)
    print(f"[{name}] FP64 update + solve done", flush=True)
    G.Factor()  # the plan now holds A2: a fresh factorization of it
    Xf = G.Solve(Bc)
    assert np.array_equal(Xu, Xf), "Update differs from a fresh Factor of the new operator"
    eta = berr(A2, Xf, Bc)
    fwd = max(relerr(xu[:, j], x[:, j]) for j in range(4))
    print(f"[{name}] FP64: eta_inf {eta:.2e}, Solve(Ax) vs x {fwd:.2e}", flush=True)
    assert eta <= ETA_FP64
    assert fwd <= FWD_FP64
    invn = inv_norm(G.Solve, n, iters=20)
    print(f"[{name}] ||A^-1||_2 ~ {invn:.3e}", flush=True)
    for form in forms:
        G.SetPrecision(form, tol=ETA_IR, max_iter=100)
        G.Factor()
        X = G.Solve(Bc)
        st = G.stats()
        e = berr(A2, X, Bc)
        fe = max(relerr(X[:, j], Xf[:, j]) for j in range(16))
        print(f"[{name}] {form}: iterations {st['ir_iters']}, eta_inf {e:.2e} (device {st['ir_berr']:.2e}), "
              f"vs FP64 {fe:.2e}", flush=True)
        ok, ratio = forward_within_bound(A2, X, Xf, Bc, invn)
        print(f"[{name}] {form}: forward error / residual bound {ratio:.3f}", flush=True)
        assert st["ir_berr"] <= ETA_IR
        assert e <= 2 * ETA_IR
        assert ok, ratio
    del G, d_v2


def test_c2_2d1023():
    _full_size("2d1023")


@pytest.mark.timeout(900)
def test_c3_3d127():
    _full_size("3d127")


@pytest.mark.timeout(900)
def test_c4_batch256_bf16x3():
    """The batch benchmark's generator (256 2D problems, N ~ U{63..511}, maxSideLength 8) in the
    BF16-split form refined in FP64: each problem up to 127^2 against the CPU oracle's FP64
    solve (the FP64 form within 1e-10, the split form within its residual's bound), every
    problem by its own backward error."""
    import bench
    from oracle_py import Factorization
    plan, rp, ci, v1, v2, mod, B, desc = bench.make_batch(nrhs=4)
    print(f"[batch] {desc}", flush=True)
    Bc = np.asfortranarray(B.T)
    plan.Factor()
    X64 = plan.Solve(Bc)
    plan.SetPrecision("bf16x3", tol=ETA_IR, max_iter=100)
    plan.Factor()
    X = plan.Solve(Bc)
    st = plan.stats()
    print(f"[batch] bf16x3: iterations {st['ir_iters']}, eta_inf {st['ir_berr']:.2e}", flush=True)
    assert st["ir_berr"] <= ETA_IR
    checked = 0
    for q in range(plan.nprob):
        o, nq, _, _ = plan.problem(q)
        sl = slice(o, o + nq)
        lo, hi = plan._nnz0[q], plan._nnz0[q + 1]
        Aq = Op(rp[o:o + nq + 1] - rp[o], ci[lo:hi] - o, v1[lo:hi])
        assert berr(Aq, X[sl], Bc[sl]) <= 2 * ETA_IR, q
        assert berr(Aq, X64[sl], Bc[sl]) <= ETA_FP64, q
        N = int(round(np.sqrt(nq)))
        if N <= 127:
            O = Factorization(Aq.c, (N, N), 1, 8, "orc")
            O.factor()
            Xo = np.stack([O.solve(Bc[sl, j]) for j in range(4)], axis=1)
            for j in range(4):
                assert relerr(X64[sl, j], Xo[:, j]) <= 1e-10, (q, j)
            ok, ratio = forward_within_bound(Aq, X[sl], Xo, Bc[sl], inv_norm(O.solve, nq, iters=30))
            assert ok, (q, ratio)
            checked += 1
    print(f"[batch] {checked} problems of side <= 127 checked against the oracle", flush=True)
    assert checked > 0
    del plan
