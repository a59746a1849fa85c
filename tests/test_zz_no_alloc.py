"""GPU: umf_solve_device_noalloc, the solve that allocates nothing, against umf_solve_device.

Both run the same sweep kernels in the same order; the no-alloc entry point only takes its work
block, per-front scratch and id list from the library's file-scope device array instead of plan
buffers, so the results must agree bit for bit.  Two calls in a row, the second on a smaller
plan with fewer right-hand sides (the array keeps the first call's contents), then the errors:
FP32 factors, and a plan whose solve does not fit the array.
"""
import numpy as np
import pytest

from devmem import DevArray
from oracle_py import Factorization, coeff_field, laplacian

pytestmark = pytest.mark.gpu


def _plan(dims, maxs, seed):
    import umf
    A = laplacian(dims, coeff_field(dims, "lognormal", seed=seed))
    G = umf.MF2D(A, *dims, width=1, maxSideLength=maxs) if len(dims) == 2 else \
        umf.MF3D(A, *dims, width=1, maxSideLength=maxs)
    G.Factor()
    return A, G


def _both(G, n, nrhs, seed):
    B = np.random.default_rng(seed).standard_normal((nrhs, n))  # row r = column r of n x nrhs
    dB = DevArray(B)
    dX1 = DevArray(np.zeros_like(B))
    dX2 = DevArray(np.full_like(B, np.nan))
    G.solve_device(dB.data_ptr(), dX1.data_ptr(), nrhs)
    G.solve_device_noalloc(dB.data_ptr(), dX2.data_ptr(), nrhs)
    return B, dX1.numpy(), dX2.numpy()


def test_noalloc_matches_solve_device_twice():
    import umf
    cases = [((127, 127), 8, 21, 1), ((47, 39), 4, 3, 2), ((15, 14, 13), 3, 5, 3)]  # 2nd smaller than 1st
    for dims, maxs, nrhs, seed in cases:
        A, G = _plan(dims, maxs, seed)
        print(f"no-alloc solve: {dims} maxSide {maxs} nrhs {nrhs}", flush=True)
        B, X1, X2 = _both(G, A.n, nrhs, seed)
        assert np.array_equal(X1, X2), dims
        O = Factorization(A, dims, 1, maxs, "orc")
        O.factor()
        xo = O.solve(B[0])
        assert np.linalg.norm(X2[0] - xo) <= 1e-10 * np.linalg.norm(xo)
        del G


def test_noalloc_rejects_fp32_and_oversized():
    import umf
    A, G = _plan((63, 63), 4, 4)
    G.SetPrecision("fp32")
    G.Factor()
    dB = DevArray(np.ones((2, A.n)))
    dX = DevArray(np.zeros((2, A.n)))
    with pytest.raises(RuntimeError, match="FP64 factors"):
        G.solve_device_noalloc(dB.data_ptr(), dX.data_ptr(), 2)
    del G
    dims = (511, 511)  # its solve needs more than the 64 MiB array
    A, G = _plan(dims, 16, 5)
    dB = DevArray(np.ones((1, A.n)))
    dX = DevArray(np.zeros((1, A.n)))
    with pytest.raises(ValueError, match="MiB of scratch"):
        G.solve_device_noalloc(dB.data_ptr(), dX.data_ptr(), 1)
    # the regular entry point still serves it
    G.solve_device(dB.data_ptr(), dX.data_ptr(), 1)
    assert np.isfinite(dX.numpy()).all()
