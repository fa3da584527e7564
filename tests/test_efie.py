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


def _subdivided_integrals(P, X, levels=5):
    """im1, iv of triangle P at X by the degree-5 rule on 4^levels
    sub-triangles (accurate away from the triangle)."""
    tris = [np.asarray(P, float)]
    for _ in range(levels):
        nt = []
        for a, b, c in tris:
            ab, bc, ca = (a + b) / 2, (b + c) / 2, (c + a) / 2
            nt += [np.array([a, ab, ca]), np.array([ab, b, bc]), np.array([ca, bc, c]), np.array([ab, bc, ca])]
        tris = nt
    bq, w = efie.RULES[5]
    n = np.cross(P[1] - P[0], P[2] - P[0])
    n /= np.linalg.norm(n)
    rho = X - ((X - P[0]) @ n)[:, None] * n
    im = np.zeros(len(X))
    iv = np.zeros((len(X), 3))
    for t in tris:
        A = 0.5 * np.linalg.norm(np.cross(t[1] - t[0], t[2] - t[0]))
        y = bq @ t
        r = np.linalg.norm(X[:, None, :] - y[None], axis=2)
        im += (w * A / r).sum(1)
        iv += np.einsum("nq,nqd->nd", w * A / r, y[None] - rho[:, None])
    return im, iv


def test_potential_integrals_closed_forms():
    """The closed-form potential integrals (Wilton et al. 1984) against
    independent evaluations: a subdivided high-order rule off the triangle;
    for points in the plane inside it, the polar form sum_i h_i
    [ln(sec t + tan t)] over the sub-triangles seen from the point; and, for
    the vector integral everywhere, the boundary form sum_i u_i int_edge R dl
    (the surface gradient of R is (rho_y - rho_x)/R)."""
    P = np.array([[0.0, 0.0, 0.0], [1.0, 0.1, 0.0], [0.2, 0.9, 0.05]])
    off = np.array([[0.3, 0.3, 0.5], [2.0, 1.0, 0.3], [1.5, 1.5, 0.0], [-0.5, 0.2, 0.2], [0.4, 0.2, -0.8]])
    a, b = efie.potential_integrals(P, off)
    c, d = _subdivided_integrals(P, off)
    assert np.max(np.abs(a - c) / np.abs(c)) < 1e-10
    assert np.max(np.linalg.norm(b - d, axis=1) / np.linalg.norm(d, axis=1)) < 1e-10
    # in the plane, inside: polar closed form
    n = np.cross(P[1] - P[0], P[2] - P[0])
    n /= np.linalg.norm(n)
    for bary in ([1 / 3, 1 / 3, 1 / 3], [0.7, 0.2, 0.1], [0.05, 0.05, 0.9]):
        x = np.asarray(bary) @ P
        im1, _ = efie.potential_integrals(P, x[None])
        ref = 0.0
        for i in range(3):
            a_, b_ = P[i], P[(i + 1) % 3]
            l = (b_ - a_) / np.linalg.norm(b_ - a_)
            foot = a_ + np.dot(x - a_, l) * l
            h = np.linalg.norm(foot - x)
            t1, t2 = np.dot(a_ - foot, l) / h, np.dot(b_ - foot, l) / h  # tan of the angles from the normal
            ref += h * (np.arcsinh(t2) - np.arcsinh(t1))                   # h [ln(sec t + tan t)]
        assert abs(im1[0] - ref) < 1e-13 * abs(ref)
    # vector integral: boundary form by Gauss-Legendre on each edge
    g, gw = np.polynomial.legendre.leggauss(60)
    pts = np.vstack([off, [np.asarray([0.7, 0.2, 0.1]) @ P]])
    _, iv = efie.potential_integrals(P, pts)
    for x, v in zip(pts, iv):
        ref = np.zeros(3)
        for i in range(3):
            a_, b_ = P[i], P[(i + 1) % 3]
            L = np.linalg.norm(b_ - a_)
            l = (b_ - a_) / L
            u = np.cross(l, n)
            s = 0.5 * (g + 1) * L
            R = np.linalg.norm(a_[None] + s[:, None] * l[None] - x[None], axis=1)
            ref += u * np.sum(0.5 * L * gw * R)
        assert np.linalg.norm(v - ref) < 1e-12 * np.linalg.norm(ref)


def test_near_quadrature_entries(mesh):
    """Near-field entries by singularity extraction: they converge with the
    rule order (degree 3 -> 5 changes them by more than degree 4 -> 5), the
    Galerkin matrix is symmetric to the quadrature accuracy, and pairs of
    functions farther apart than the extraction distance reproduce the far
    entries of the same double rule."""
    V, T = mesh
    from bfgpu import rwg
    r = rwg.RWG(V, T)
    c = r.centers
    rows = np.argsort(np.linalg.norm(c - c[10], axis=1))[:8]
    Z5 = efie.efie_near_block(V, T, K, rows, rows, near_rules=(5, 5), rules=(5, 5))
    Z4 = efie.efie_near_block(V, T, K, rows, rows, near_rules=(4, 4), rules=(4, 4))
    Z3 = efie.efie_near_block(V, T, K, rows, rows, near_rules=(3, 3), rules=(3, 3))
    d35 = np.linalg.norm(Z3 - Z5) / np.linalg.norm(Z5)
    d45 = np.linalg.norm(Z4 - Z5) / np.linalg.norm(Z5)
    assert d45 < 0.25 * d35 and d45 < 0.02
    assert np.linalg.norm(Z5 - Z5.T) / np.linalg.norm(Z5) < 1e-3
    far = np.argsort(np.linalg.norm(c - c[10], axis=1))[-5:]
    Zn = efie.efie_near_block(V, T, K, rows[:4], far, near_rules=(5, 5), rules=(5, 5), near_factor=3.0)
    Zf = efie.efie_block(V, T, K, rows[:4], far, rules=(5, 5))
    assert np.linalg.norm(Zn - Zf) < 1e-13 * np.linalg.norm(Zf)
    # the static part is what makes the self and neighbour entries finite:
    # the double rule alone misses them by a wide margin
    Zd = efie.efie_block(V, T, K, rows, rows, rules=(5, 5))
    assert np.linalg.norm(Zd - Z5) / np.linalg.norm(Z5) > 0.1
