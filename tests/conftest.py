import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oceanwave3d-fortran90_amd"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def oracle_params(nx, ny, nz, **kw):
    """Oracle parameter dict with the deck defaults used throughout the tests."""
    p = dict(nx=nx, ny=ny, nz=nz, alpha=3, beta=3, gamma=3, gridz=1, lx=1.0, ly=1.0, lz=1.0, g=9.81,
             nonlinear=1, solver=1, precond=3, gmres_m=23, reltol=1e-8, abstol=0.0, maxit=1, nu1=1, nu2=1,
             mg_maxgrids=20, extrapolation=1, mg_mode=1, cd_quirk=0, gmres_side=0)
    p.update(kw)
    return p


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.build()
    return oracle
