"""CPU: the GEMM-stage tuning knobs are checked before any device work (umf_set_tune on an
analysis-only plan)."""
import pytest

from oracle_py import coeff_field, laplacian


def test_set_tune_checks_arguments():
    import umf
    dims = (31, 29)
    A = laplacian(dims, coeff_field(dims, "uniform", seed=2))
    G = umf.MF2D(A, *dims, width=1, maxSideLength=5, analyze_only=True)
    assert G.tune() == {"gw2": 1024, "tile": 0}
    for bad in ((300, 0), (0, 0), (16384, 0), (1024, 32), (1024, -1)):
        with pytest.raises(ValueError):
            G.set_tune(*bad)
    G.set_tune(2048, 128)
    assert G.tune() == {"gw2": 2048, "tile": 128}
