import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "updating-multifrontal_amd"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on cuda:0)")
    config.addinivalue_line("markers", "ref: needs the reference harness oracle/_ref (built from /root/reference)")


def pytest_collection_modifyitems(config, items):
    from oracle_py import have_ref
    if not have_ref():
        skip = pytest.mark.skip(reason="reference harness oracle/_ref not built here")
        for it in items:
            if "ref" in it.keywords:
                it.add_marker(skip)
