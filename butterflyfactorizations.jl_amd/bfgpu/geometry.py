"""Point sets and excitations of the benchmark geometries (host-side
generators; every geometry is seeded and reproducible)."""
import numpy as np


def two_squares(n, seed=0, side=16.0, gap=32.0):
    """Observers uniform on [0,side]^2 x {0}, sources uniform on
    [0,side]^2 x {gap}: the single admissible block of the two-squares
    workloads."""
    rng = np.random.default_rng(seed)
    obs = np.c_[rng.uniform(0, side, n), rng.uniform(0, side, n), np.zeros(n)]
    src = np.c_[rng.uniform(0, side, n), rng.uniform(0, side, n), np.full(n, gap)]
    return obs, src


def fibonacci_sphere(n, radius):
    """n points of the Fibonacci lattice on a sphere of the given radius."""
    i = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * i / n)
    theta = np.pi * (1 + 5 ** 0.5) * i
    return radius * np.c_[np.cos(theta) * np.sin(phi), np.sin(theta) * np.sin(phi), np.cos(phi)]


def ellipsoid_surface(n, axes, seed=0):
    """n points uniform in surface area on the ellipsoid with semi-axes
    (a, b, c): points uniform on the unit sphere, mapped by diag(a, b, c),
    kept with probability proportional to the area stretch of the map
    (|diag(a,b,c)^-1 u| a b c, maximal at the shortest axis)."""
    a = np.asarray(axes, float)
    rng = np.random.default_rng(seed)
    out = []
    have = 0
    gmax = np.prod(a) / a.min()
    while have < n:
        m = max(1024, int(1.5 * (n - have)))
        u = rng.standard_normal((m, 3))
        u /= np.linalg.norm(u, axis=1)[:, None]
        g = np.linalg.norm(u / a, axis=1) * np.prod(a)
        keep = rng.uniform(0, gmax, m) < g
        p = u[keep] * a
        out.append(p)
        have += p.shape[0]
    return np.concatenate(out)[:n]


def plane_waves(points, k, n_dirs, seed=0):
    """Columns exp(-i k d . x) for n_dirs directions d uniform on the
    sphere (Fortran-ordered, points x n_dirs)."""
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((n_dirs, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    return np.asfortranarray(np.exp(-1j * k * (points @ d.T)))
