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


def two_squares(n, seed=0, side=16.0, gap=32.0):
    """Observers uniform on [0,side]^2 x {0}, sources on [0,side]^2 x {gap}
    (the single-block geometry of the two-squares workloads)."""
    rng = np.random.default_rng(seed)
    obs = np.c_[rng.uniform(0, side, n), rng.uniform(0, side, n), np.zeros(n)]
    src = np.c_[rng.uniform(0, side, n), rng.uniform(0, side, n), np.full(n, gap)]
    return obs, src


def fibonacci_sphere(n, radius):
    i = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * i / n)
    theta = np.pi * (1 + 5 ** 0.5) * i
    return radius * np.c_[np.cos(theta) * np.sin(phi), np.sin(theta) * np.sin(phi), np.cos(phi)]


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
