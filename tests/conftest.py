import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "postgis-cuda_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs through libgsj's C ABI")


def sorted_rows(pairs, *cols):
    """rows ordered by (a, b) so that two result sets can be compared element-wise"""
    import numpy as np

    pairs = np.asarray(pairs).reshape(-1, 2)
    order = np.lexsort((pairs[:, 1], pairs[:, 0]))
    return (pairs[order],) + tuple(None if c is None else np.asarray(c)[order] for c in cols)


@pytest.fixture(scope="session")
def ctx():
    import gsj

    c = gsj.Context(0)
    # the built-in launch shapes, whatever table ships beside the library: small test joins then use every work class
    # (rows are the same under every table; tests/test_gpu_tuning.py checks the shipped one)
    c.set_tuning("bucket np>=0\n")
    yield c
    c.close()
