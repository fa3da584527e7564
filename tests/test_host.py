"""Host-side checks of the C-ABI library that need no GPU: symbol exports,
cluster trees (the tree contract of src/treeinterface.jl), interaction lists
and argument errors."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT, fibonacci_sphere, two_squares


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "bfgpu.h")).read()
    return sorted(set(re.findall(r"\b(bfgpu_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(bfgpu):
    from bfgpu import _lib
    h = _lib.lib()
    declared = header_symbols()
    assert len(declared) > 30
    for name in declared:
        assert hasattr(h, name), name
    assert set(declared) == set(_lib.EXPORTED)
    assert b"gfx950" in h.bfgpu_version()


def _tree_invariants(T, leaf, mode):
    n = T.n_values
    assert sorted(T.perm.tolist()) == list(range(n))
    assert T.parent[0] == -1 and T.val_begin[0] == 0 and T.val_end[0] == n
    for v in range(T.n_nodes):
        ch = T.cluster_children(v)
        if ch:
            assert all(T.parent[c] == v for c in ch)
            assert T.val_begin[ch[0]] == T.val_begin[v] and T.val_end[ch[-1]] == T.val_end[v]
            for a, b in zip(ch[:-1], ch[1:]):
                assert T.val_end[a] == T.val_begin[b]
            assert all(T.val_end[c] > T.val_begin[c] for c in ch)
        pts = T.points[T.cluster_values(v)]
        assert np.all(np.abs(pts[:, :2] - T.center[v, :2]) <= T.halfsize[v] * (1 + 1e-9))
    leaves = T.leaves()
    if mode == "maxpoints":
        assert all(T.val_end[v] - T.val_begin[v] <= leaf for v in leaves)
    else:
        assert len(set(T.level[leaves].tolist())) == 1


@pytest.mark.parametrize("mode", ["balanced", "maxpoints"])
def test_quadtree_and_octree_contract(bfgpu, mode):
    obs, _ = two_squares(4096)
    T = bfgpu.ClusterTree(obs, dims=2, leaf_size=64, mode=mode)
    _tree_invariants(T, 64, mode)
    if mode == "balanced":
        assert len(T.cluster_levels()) == 4 and len(T.leaves()) == 64
    S = bfgpu.ClusterTree(fibonacci_sphere(5000, 3.0), dims=3, leaf_size=64, mode=mode)
    _tree_invariants(S, 64, mode)
    assert S.cluster_radius(0) == pytest.approx(S.halfsize[0] * np.sqrt(3))


def test_tree_from_arrays_round_trip(bfgpu):
    obs, _ = two_squares(1000, seed=3)
    T = bfgpu.ClusterTree(obs, dims=2, leaf_size=50, mode="maxpoints")
    U = bfgpu.ClusterTree.from_arrays(T.points, T.parent, T.first_child, T.n_children,
                                      T.val_begin, T.val_end, T.center, T.halfsize, T.perm)
    for name in ("parent", "first_child", "n_children", "val_begin", "val_end", "perm", "level"):
        np.testing.assert_array_equal(getattr(T, name), getattr(U, name))
    bad = T.parent.copy()
    bad[1] = 5
    with pytest.raises(bfgpu.ArgumentError):
        bfgpu.ClusterTree.from_arrays(T.points, bad, T.first_child, T.n_children, T.val_begin,
                                      T.val_end, T.center, T.halfsize, T.perm)


def test_nearandfar_matches_oracle(bfgpu, oracle):
    T = bfgpu.ClusterTree(fibonacci_sphere(6000, 4.0), dims=3, leaf_size=64, mode="maxpoints")
    bt = bfgpu.BlockTree(T, T)
    for adm, kind, p in [(bfgpu.isFarFunctor(1.0), 0, 1.0),
                         (bfgpu.CenterDistanceAdmissibility(2.0), 1, 2.0)]:
        for unb in (False, True):
            far, near = bfgpu.nearandfar(bt, adm, unbalancedints=unb)
            ofar, onear = oracle.nearandfar(T, T, kind, p, unbalancedints=unb)
            np.testing.assert_array_equal(far, ofar)
            np.testing.assert_array_equal(near, onear)
    far, near = bfgpu.nearandfar(bt, bfgpu.isFarFunctor(1.0), leafcomp=False)
    ofar, onear = oracle.nearandfar(T, T, 0, 1.0, leafcomp=False)
    np.testing.assert_array_equal(far, ofar)


def test_argument_errors(bfgpu):
    with pytest.raises(bfgpu.ArgumentError):
        bfgpu.ClusterTree(np.zeros((10, 2)))
    from bfgpu import _lib
    h = C.c_void_p()
    pts = np.zeros((4, 3))
    st = _lib.lib().bfgpu_tree_build(_lib.ptr(pts, C.c_double), 4, 5, 64, 0, C.byref(h))
    assert st == 1 and b"dims" in _lib.lib().bfgpu_last_error()
    st = _lib.lib().bfgpu_matvec(None, 0, 1, None, 1, None, 1, None, None, 0)
    assert st == 1
    assert _lib.lib().bfgpu_factorize(None, None, None, 1, None, None, C.byref(h)) == 1


def test_tree_parameters_and_estimator_defaults(bfgpu):
    assert bfgpu.tree_parameters().Ctau == 4.0
    assert bfgpu.tree_parameters(None, bfgpu.isFarFunctor(1.0)).Ctau == pytest.approx(2.9592)
    assert bfgpu.tree_parameters(None, bfgpu.CenterDistanceAdmissibility(2.0)).Ctau == pytest.approx(2.9352)
    assert bfgpu.ButterflyRankEstimator().Rmin == 10


def test_ellipsoid_points_are_uniform_in_area():
    """Area-uniform sampling on a prolate spheroid: the fraction of points in
    the band |z| < c/2 matches the band's share of the surface area
    (surface of revolution, numerical quadrature); points lie on the
    surface.  A sampler without the area correction fails the band test."""
    from scipy.integrate import quad
    from conftest import ellipsoid_surface
    a, c = 1.0, 3.0
    p = ellipsoid_surface(200000, (a, a, c), seed=5)
    assert np.allclose((p[:, 0] / a) ** 2 + (p[:, 1] / a) ** 2 + (p[:, 2] / c) ** 2, 1.0, atol=1e-9)
    # surface of revolution rho(z) = a sqrt(1 - z^2/c^2); dA = 2 pi rho sqrt(1 + rho'^2) dz
    dA = lambda z: 2 * np.pi * np.sqrt(a * a * (1 - z * z / (c * c)) + (a * a * z / (c * c)) ** 2)
    band = quad(dA, -c / 2, c / 2)[0] / quad(dA, -c, c)[0]
    frac = np.mean(np.abs(p[:, 2]) < c / 2)
    assert abs(frac - band) < 0.005
    # the same map without the rejection step puts too few points in the band
    u = np.random.default_rng(5).standard_normal((200000, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    assert abs(np.mean(np.abs(u[:, 2] * c) < c / 2) - band) > 0.05


def test_vector_layouts(bfgpu):
    """Column-major blocks of right-hand sides, with the strides numpy gives
    length-1 dimensions (never used); row-major or overlapping blocks are
    rejected before any call."""
    lay = bfgpu.ButterflyFactorization._layout
    es = 16
    for shape in [(1, 3), (4, 1), (1, 1), (5, 4)]:
        a = np.asfortranarray(np.zeros(shape, np.complex128))
        b = np.ascontiguousarray(np.zeros(shape, np.complex128))
        for arr in (a, b):
            if arr is b and shape[0] > 1 and shape[1] > 1:
                continue
            nr, ld = lay(2, arr.shape, tuple(s // es for s in arr.strides), 1, shape[0], "x")
            assert nr == shape[1] and ld >= shape[0]
            if shape[1] > 1:  # column j starts at j * ld
                assert ld == arr.strides[1] // es
    with pytest.raises(bfgpu.ArgumentError):
        lay(2, (5, 4), (4, 1), 1, 5, "x")  # row-major
    with pytest.raises(bfgpu.ArgumentError):
        lay(2, (5, 4), (1, 2), 1, 5, "x")  # overlapping columns
    with pytest.raises(bfgpu.ArgumentError):
        lay(2, (4, 4), (1, 4), 1, 5, "x")  # wrong row count
    assert lay(1, (1,), (7,), 1, 1, "x") == (1, 1)
    with pytest.raises(bfgpu.ArgumentError):
        lay(1, (3,), (2,), 1, 3, "x")


def test_sparse_level_matrices_apply_the_factorization(bfgpu, oracle):
    """bf_to_mat (ButterflyFactorization_Mat, bfstructs.jl:177-199) of an
    exported factorization: P R ... R Q on x[permQ], scattered by permP,
    equals the factorization's matvec; the permutations are permutations of
    the points; balanced (two squares) and padded (max-points) trees."""
    from conftest import crandn, fibonacci_sphere, two_squares
    rng = np.random.default_rng(0)
    obs, src = two_squares(1024, seed=2)
    To = bfgpu.ClusterTree(obs, dims=2, leaf_size=64)
    Ts = bfgpu.ClusterTree(src, dims=2, leaf_size=64)
    cases = [(To, Ts, 0, 0)]
    pts = fibonacci_sphere(3000, 3.0)
    T = bfgpu.ClusterTree(pts, dims=3, leaf_size=64, mode="maxpoints")
    far, _ = oracle.nearandfar(T, T, 1, 2.0)
    big = max(far, key=lambda p: len(T.cluster_values(p[0])) * len(T.cluster_values(p[1])))
    cases.append((T, T, int(big[0]), int(big[1])))
    for te, tr, no, ns in cases:
        obf = oracle.assemble(te, tr, no, ns, 2 * np.pi, 1e-4)
        M = bfgpu.bf_to_mat(obf.export(), te, tr)
        assert sorted(M.permQ) == sorted(tr.cluster_values(ns))
        assert sorted(M.permP) == sorted(te.cluster_values(no))
        assert M.Q.shape[1] == len(M.permQ) and M.P.shape[0] == len(M.permP)
        for a, b in zip([M.Q] + M.R, M.R + [M.P]):
            assert a.shape[0] == b.shape[1]
        x = crandn(rng, tr.n_values)
        want = obf.matvec(x)
        got = M @ x
        assert np.linalg.norm(got[:len(want)] - want) <= 1e-13 * np.linalg.norm(want)
        X = crandn(rng, tr.n_values, 3)
        assert np.allclose((M @ X)[:, 1], M @ X[:, 1], rtol=0, atol=1e-14 * np.abs(want).max())


def test_public_header_is_plain_c(tmp_path):
    """include/bfgpu.h is a C header (the boundary a Julia ccall, a cgo or
    a ctypes binding reads): it compiles as C99 with warnings as errors."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    src = tmp_path / "h.c"
    src.write_text('#include "bfgpu.h"\nint main(void) { bfgpu_opts o; bfgpu_bf_info i; (void)o; (void)i; return 0; }\n')
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    r = subprocess.run([gcc, "-std=c99", "-Wall", "-Wextra", "-Werror", "-I", inc, "-c", str(src), "-o",
                        str(tmp_path / "h.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
