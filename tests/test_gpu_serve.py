"""GPU: serving requests against a resident factorization (umf_server_*, serve.SlotServer).

A request is a coefficient box change plus right-hand sides, answered against the unchanged
base factorization.  The checks:

- every request's solution is the same, bit for bit, when it is served alone (one slot, one
  request at a time) and in the mix (several slots, all requests admitted as slots and pool
  memory free up, their work overlapping on the device);
- and the same, bit for bit, as the reference's way to get it: Update of the factored base
  with the request's operator, then Solve (src/DE/mf3d.cpp:208-248; an Update here equals a
  fresh factorization of the new operator bit for bit, which the full-size tests check);
- admission by footprint: with a pool that holds one request at a time the mix still completes
  (requests wait for memory), a request larger than the pool is refused, and the pool is empty
  again at the end;
- the base plan is immutable while a server lives (Factor / Update / SetPrecision refused).

Sizes: 3D 95^3 (the serving configuration: base 42 GB FP64, boxes of side 4..16, 1..32
right-hand sides) and a 2D problem in the FP32 form with FP64 refinement.
"""
import numpy as np
import pytest

from devmem import DevArray

pytestmark = pytest.mark.gpu


def _setup(dims, seed, prec="fp64"):
    import umf
    rng = np.random.default_rng(seed)
    w = 1
    coeff = np.exp(rng.standard_normal(tuple(n + 2 * w for n in dims)))
    h = 1.0 / (dims[0] + 1)
    rp, ci, v = umf.build_laplacian(dims, coeff, w, h)
    A = umf.CSR(rp, ci, v, len(rp) - 1, len(rp) - 1)
    G = (umf.MF2D if len(dims) == 2 else umf.MF3D)(A, *dims, w, 8)
    if prec != "fp64":
        G.SetPrecision(prec, tol=1e-10, max_iter=100)
    G.Factor()
    return rng, coeff, h, rp, ci, v, G


def _requests(rng, dims, n, count, smin, smax, kmax):
    from serve import Request
    reqs = []
    for i in range(count):
        side = int(rng.integers(smin, smax + 1))
        lo = [int(x) for x in rng.integers(0, [d + 2 - side for d in dims])]
        fac = float(np.exp(2.0 * rng.standard_normal()))
        k = int(rng.integers(1, kmax + 1))
        B = rng.standard_normal((k, n))
        reqs.append(Request(B, arrival=0.0, box=(lo, side, fac)))
    return reqs


def _fresh_buffers(reqs):
    for r in reqs:
        r.d_b = DevArray(r.rhs)
        r.d_x = DevArray(np.full_like(r.rhs, np.nan))


def _serve_alone_and_mixed(G, coeff, h, rp, ci, v, reqs, slots, pool):
    from serve import SlotServer
    d_coeff = DevArray(coeff)
    S = SlotServer(G, None, coeff, d_coeff, v, rp, ci, h, slots=1, pool_bytes=pool)
    alone = []
    for r in reqs:  # one request at a time
        _fresh_buffers([r])
        S.run([r], realtime=False)
        alone.append(r.d_x.numpy())
    S.close()
    S = SlotServer(G, None, coeff, d_coeff, v, rp, ci, h, slots=slots, pool_bytes=pool)
    _fresh_buffers(reqs)
    S.run(reqs, realtime=False)  # every request queued at once: admitted as slots and memory free up
    mixed = [r.d_x.numpy() for r in reqs]
    total, used = S.srv.pool()
    assert used == 0.0
    inflight = max(c for _, c in S.inflight_trace)
    S.close()
    return alone, mixed, inflight


def _check_against_update(G, coeff, h, rp, ci, v, reqs, alone, dims):
    """Update of the base with each request's operator, then Solve: the reference's path."""
    import umf
    prev = np.zeros(0, np.int32)
    for r, xa in zip(reqs, alone):
        lo, side, f = r.box
        c2 = coeff.copy()
        c2[tuple(slice(l, l + side) for l in lo)] *= f
        _, _, v2 = umf.build_laplacian(dims, c2, 1, h)
        assert np.array_equal(np.unique(ci[v != v2]), r.mod)
        d_v2 = DevArray(v2)
        G.update_device(d_v2.data_ptr(), np.union1d(prev, r.mod))  # from the previous request's operator
        prev = r.mod
        dX = DevArray(np.zeros_like(r.rhs))
        G.solve_device(r.d_b.data_ptr(), dX.data_ptr(), len(r.rhs))
        assert np.array_equal(dX.numpy(), xa), r.box


@pytest.mark.timeout(600)
def test_serve_95_alone_equals_mix_equals_update():
    dims = (95, 95, 95)
    n = int(np.prod(dims))
    rng, coeff, h, rp, ci, v, G = _setup(dims, 5)
    reqs = _requests(rng, dims, n, 6, 4, 16, 32)
    print("requests:", [(r.box[1], len(r.rhs)) for r in reqs], flush=True)
    alone, mixed, inflight = _serve_alone_and_mixed(G, coeff, h, rp, ci, v, reqs, slots=3, pool=40e9)
    print(f"max requests in flight {inflight}; footprints GB {[round(r.footprint / 1e9, 2) for r in reqs]}", flush=True)
    assert inflight >= 2  # they did share the device
    for r, xa, xm in zip(reqs, alone, mixed):
        assert np.isfinite(xa).all()
        assert np.array_equal(xa, xm), r.box
    _check_against_update(G, coeff, h, rp, ci, v, reqs, alone, dims)


def test_serve_fp32_refined_2d():
    dims = (255, 255)
    n = int(np.prod(dims))
    rng, coeff, h, rp, ci, v, G = _setup(dims, 9, prec="fp32")
    reqs = _requests(rng, dims, n, 5, 2, 12, 8)
    alone, mixed, inflight = _serve_alone_and_mixed(G, coeff, h, rp, ci, v, reqs, slots=4, pool=2e9)
    for r, xa, xm in zip(reqs, alone, mixed):
        assert np.array_equal(xa, xm), r.box
        assert r.ir_iters >= 1
    _check_against_update(G, coeff, h, rp, ci, v, reqs, alone, dims)


def test_serve_admission_by_footprint():
    import umf
    from serve import SlotServer
    dims = (127, 127)
    n = int(np.prod(dims))
    rng, coeff, h, rp, ci, v, G = _setup(dims, 3)
    reqs = _requests(rng, dims, n, 6, 3, 10, 4)
    d_coeff = DevArray(coeff)
    S = SlotServer(G, None, coeff, d_coeff, v, rp, ci, h, slots=4, pool_bytes=1e9)
    for r in reqs:
        S.mods(r)
    big = max(S.srv.footprint(r.mod) for r in reqs)
    S.close()
    # a pool that holds the largest request but not two of it: a second one waits for memory
    S = SlotServer(G, None, coeff, d_coeff, v, rp, ci, h, slots=4, pool_bytes=big * 1.5)
    rbig = max(reqs, key=lambda r: S.srv.footprint(r.mod))
    t = S.srv.admit(rbig.mod)
    assert t is not None and S.srv.pool()[1] == S.srv.footprint(rbig.mod)
    assert S.srv.admit(rbig.mod) is None  # busy: no room now
    S.srv.release(t)
    t = S.srv.admit(rbig.mod)  # room again after the release
    assert t is not None
    S.srv.release(t)
    _fresh_buffers(reqs)
    S.run(reqs, realtime=False)  # completes: requests that do not fit wait for memory
    assert all(np.isfinite(r.d_x.numpy()).all() for r in reqs)
    assert S.srv.pool()[1] == 0.0
    with pytest.raises(ValueError, match="pool holds"):
        S.srv.admit(np.arange(n, dtype=np.int32))  # every leaf dirty: more than the pool
    # the base is immutable while the server lives
    with pytest.raises(RuntimeError, match="server"):
        G.Factor()
    with pytest.raises(RuntimeError, match="server"):
        G.SetPrecision("fp32")
    S.close()
    G.Factor()  # allowed again
