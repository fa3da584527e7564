import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "butterflyfactorizations.jl_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


from bfgpu.geometry import ellipsoid_surface, fibonacci_sphere, plane_waves, two_squares  # noqa: E402,F401


def crandn(rng, *shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2)


@pytest.fixture(scope="session")
def bfgpu():
    import bfgpu as m
    return m


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    return o
