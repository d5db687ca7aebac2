import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "radialbasisfinitedifferences.jl_amd")
for p in (ROOT, PKG_DIR):
    if p not in sys.path:
        sys.path.insert(0, p)

REF_DATA = os.path.join(ROOT, "tests", "golden", "reference_data")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs librbffd kernels)")


@pytest.fixture(scope="session")
def O():
    from oracle import oracle

    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def fitted_nodes():
    x = np.loadtxt(os.path.join(REF_DATA, "x_nodes_fitted.csv"), delimiter=",")
    y = np.loadtxt(os.path.join(REF_DATA, "y_nodes_fitted.csv"), delimiter=",")
    return x, y


@pytest.fixture(scope="session")
def ctx():
    import rbffd_amd

    c = rbffd_amd.Context(0)
    yield c
    c.close()
