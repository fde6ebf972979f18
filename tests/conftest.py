"""Shared fixtures.  The oracle (oracle/) is the parity checker; the product
is driven through pureibm_ps2d_fd1d_amd over the C-ABI."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.build()
    return oracle


def make_setup(tmp_path, name, kind="simple", precision=0, **kw):
    """Write a namelist deck with configs.py and run the Fortran host set-up."""
    from pureibm_ps2d_fd1d_amd import configs, solver
    d = str(tmp_path)
    if kind == "simple":
        configs.simple_cubic(d, name, **kw)
    else:
        configs.random_bed(d, name, **kw)
    return solver.RunSetup(os.path.join(d, name), precision=precision)


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)
