"""EFIE / RWG plug-in: the oracle and the host-side mesh, RWG and sample
construction (CPU).  The oracle restates the far-interaction entries of the
BEAST plug-in (ext/BFBEAST/kernelmatrix.jl:22-82); BEAST is absent, so the
entries are pinned by what the mathematics fixes: the triangle rules are
exact on monomials up to their degree, RWG functions have zero total
divergence and unit normal flux across their edge, the Galerkin matrix with
equal test and trial rules is symmetric, and the degree-(2, 3) entries of
separated functions converge to the high-order ones."""
import math

import numpy as np
import pytest

from oracle import efie

K = 2 * np.pi


@pytest.fixture(scope="module")
def mesh():
    from bfgpu import rwg
    return rwg.icosphere(1.0, 2)


@pytest.mark.parametrize("degree", [1, 2, 3, 4, 5])
def test_triangle_rules_exact_on_monomials(degree):
    from bfgpu import rwg
    for bary, w in (efie.RULES[degree], rwg.triangle_rule(degree)):
        assert abs(w.sum() - 1.0) < 1e-14
        x, y = bary[:, 1], bary[:, 2]  # reference triangle (0,0), (1,0), (0,1): area 1/2
        for a in range(degree + 1):
            for b in range(degree + 1 - a):
                exact = math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
                assert abs(0.5 * np.sum(w * x ** a * y ** b) - exact) < 1e-14, (degree, a, b)
    # one degree more is not exact (the rule is not over-stated)
    bary, w = efie.RULES[degree]
    x = bary[:, 1]
    n = degree + 1
    assert abs(0.5 * np.sum(w * x ** n) - math.factorial(n) / math.factorial(n + 2)) > 1e-6


def test_mesh_and_rwg_functions(mesh):
    from bfgpu import rwg
    V, T = mesh
    r = rwg.RWG(V, T)
    assert len(V) - len(r) + len(T) == 2                       # closed genus-0 surface
    assert len(r) == 3 * len(T) // 2 and np.allclose(np.linalg.norm(V, axis=1), 1.0)
    E, tp, tm, op, om = efie.rwg_functions(V, T)
    np.testing.assert_array_equal(E, r.edges)
    np.testing.assert_array_equal(tp, r.plus)
    np.testing.assert_array_equal(tm, r.minus)
    for n in range(0, len(r), 37):
        # unit flux across the edge from both sides: f . nu = 1 on the edge,
        # nu the in-plane normal of the edge pointing from plus to minus
        a, b = V[E[n, 0]], V[E[n, 1]]
        mid = 0.5 * (a + b)
        l = np.linalg.norm(b - a)
        for tri, free, s in ((tp[n], op[n], 1.0), (tm[n], om[n], -1.0)):
            P = V[T[tri]]
            area = 0.5 * np.linalg.norm(np.cross(P[1] - P[0], P[2] - P[0]))
            f = s * l / (2 * area) * (mid - V[free])
            e = (b - a) / l
            nu = (mid - V[free]) - np.dot(mid - V[free], e) * e
            nu /= np.linalg.norm(nu)
            assert abs(np.dot(f, s * nu) - 1.0) < 1e-12
    # zero total divergence, and the host samples carry the same divergence
    S = r.samples(3)
    assert np.abs(S[:, :, 6].sum(axis=1)).max() < 1e-12
    xa, fa, da = efie._samples(V, T, (E, tp, tm, op, om), np.arange(len(r)), 3)
    np.testing.assert_allclose(S[:, :, 0:3], xa, atol=1e-15)
    np.testing.assert_allclose(S[:, :, 3:6], fa, atol=1e-15)
    np.testing.assert_allclose(S[:, :, 6], da, atol=1e-15)


def test_galerkin_symmetry_and_convergence(mesh):
    V, T = mesh
    E = efie.rwg_functions(V, T)[0]
    mid = 0.5 * (V[E[:, 0]] + V[E[:, 1]])
    rows = np.nonzero(mid[:, 2] > 0.7)[0][:12]
    cols = np.nonzero(mid[:, 2] < -0.7)[0][:15]
    # equal rules on both sides: Z[rows, cols] = Z[cols, rows]^T
    Z = efie.efie_block(V, T, K, rows, cols, rules=(3, 3))
    Zt = efie.efie_block(V, T, K, cols, rows, rules=(3, 3))
    np.testing.assert_allclose(Z, Zt.T, rtol=1e-13, atol=0)
    # the (2, 3) entries of separated functions are close to degree-5 ones
    lo = efie.efie_block(V, T, K, rows, cols, rules=(2, 3))
    hi = efie.efie_block(V, T, K, rows, cols, rules=(5, 5))
    assert np.linalg.norm(lo - hi) < 2e-2 * np.linalg.norm(hi)
    assert np.linalg.norm(efie.efie_block(V, T, K, rows, cols, rules=(4, 4)) - hi) < np.linalg.norm(lo - hi)
    # the two terms separately: alpha only / beta only add up
    a = efie.efie_block(V, T, K, rows, cols, beta=0.0)
    b = efie.efie_block(V, T, K, rows, cols, alpha=0.0)
    np.testing.assert_allclose(a + b, lo, rtol=1e-13)


def test_host_samples_reproduce_the_oracle_entries(mesh):
    """The plug-in's host samples, summed on the CPU as the device fill sums
    them, give the oracle's entries (the device fill is checked against the
    oracle on the GPU)."""
    from bfgpu import rwg
    V, T = mesh
    km = rwg.EFIEKernelMatrix(rwg.RWG(V, T), K)
    ts, ss = km.rwg.samples(2), km.trial_rwg.samples(3)
    rows, cols = np.arange(0, 480, 41), np.arange(5, 480, 29)
    Z = np.zeros((len(rows), len(cols)), complex)
    for i, m in enumerate(rows):
        for j, n in enumerate(cols):
            d = ts[m, :, None, 0:3] - ss[n, None, :, 0:3]
            r = np.linalg.norm(d, axis=2)
            with np.errstate(divide="ignore", invalid="ignore"):
                G = np.where(r > 0, np.exp(-1j * K * r) / (4 * np.pi * r), 0)
            w = km.alpha * (ts[m, :, None, 3:6] * ss[n, None, :, 3:6]).sum(2) + km.beta * ts[m, :, None, 6] * ss[n, None, :, 6]
            Z[i, j] = (G * w).sum()
    want = efie.efie_block(V, T, K, rows, cols)
    assert np.linalg.norm(Z - want) <= 1e-13 * np.linalg.norm(want)
