"""Triangle meshes, RWG functions and the EFIE kernel-matrix plug-in.

The plug-in stands in for the BEAST kernel matrix of
ext/BFBEAST/kernelmatrix.jl:22-82 (Maxwell single layer between RWG /
Raviart-Thomas test and trial functions, far interactions by a double
numerical quadrature, DoubleNumQStrat(2, 3)).  BEAST is not available here,
so the quadrature orders are read as polynomial degrees: a degree-2 rule
(3 points) on test triangles and a degree-3 rule (4 points) on trial
triangles, from the published tables of Strang & Fix and Dunavant (1985).

Mesh format on disk: an .npz with ``vertices`` (V x 3 float64) and
``triangles`` (T x 3 int, 0-based vertex ids, consistently oriented).
"""
from dataclasses import dataclass, field

import numpy as np

from . import _destroy

# symmetric triangle rules: (barycentric points, weights summing to 1)
_RULES = {
    1: (np.array([[1 / 3, 1 / 3, 1 / 3]]), np.array([1.0])),
    2: (np.array([[2 / 3, 1 / 6, 1 / 6], [1 / 6, 2 / 3, 1 / 6], [1 / 6, 1 / 6, 2 / 3]]), np.full(3, 1 / 3)),
    3: (np.array([[1 / 3, 1 / 3, 1 / 3], [0.6, 0.2, 0.2], [0.2, 0.6, 0.2], [0.2, 0.2, 0.6]]),
        np.array([-27 / 48, 25 / 48, 25 / 48, 25 / 48])),
}


def _orbit3(a, w):
    b = 1.0 - 2.0 * a
    return [[b, a, a], [a, b, a], [a, a, b]], [w] * 3


def _dunavant(parts):
    pts, ws = [], []
    for a, w in parts:
        if a is None:
            pts.append([1 / 3, 1 / 3, 1 / 3])
            ws.append(w)
        else:
            p, q = _orbit3(a, w)
            pts += p
            ws += q
    return np.array(pts), np.array(ws)


_RULES[4] = _dunavant([(0.445948490915965, 0.223381589678011), (0.091576213509771, 0.109951743655322)])
_RULES[5] = _dunavant([(None, 0.225), (0.470142064105115, 0.132394152788506),
                       (0.101286507323456, 0.125939180544827)])


def triangle_rule(degree):
    """(barycentric points (q x 3), weights (q,)) exact for polynomials of
    the given degree (1-5) on a triangle; weights sum to 1."""
    if degree not in _RULES:
        raise ValueError("triangle rules of degree 1 to 5")
    b, w = _RULES[degree]
    return b.copy(), w.copy()


def icosphere(radius=1.0, subdivisions=3):
    """Sphere mesh: an icosahedron subdivided `subdivisions` times, vertices
    projected to the sphere; triangles oriented outwards.  20 4^s
    triangles, 30 4^s edges."""
    t = (1.0 + 5 ** 0.5) / 2
    V = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    F = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
         [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
         [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    V = [np.array(v, float) / np.linalg.norm(v) for v in V]
    for _ in range(subdivisions):
        mid = {}
        nF = []

        def midpoint(i, j):
            key = (min(i, j), max(i, j))
            if key not in mid:
                m = V[i] + V[j]
                V.append(m / np.linalg.norm(m))
                mid[key] = len(V) - 1
            return mid[key]

        for a, b, c in F:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nF += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        F = nF
    return radius * np.array(V), np.array(F, np.int64)


def save_mesh(path, vertices, triangles):
    np.savez(path, vertices=np.asarray(vertices, np.float64), triangles=np.asarray(triangles, np.int64))


def load_mesh(path):
    with np.load(path, allow_pickle=False) as z:
        return np.asarray(z["vertices"], np.float64), np.asarray(z["triangles"], np.int64)


@dataclass
class RWG:
    """RWG functions on the interior edges of a triangle mesh (edges shared
    by exactly two triangles), in order of their sorted vertex pairs.
    Function n lives on triangles plus[n] and minus[n]:
    f = l/(2 A+) (r - v+) on the plus triangle, l/(2 A-) (v- - r) on the
    minus triangle, with v+-, the vertices opposite the edge and l its
    length; div f = +l/A+, -l/A-."""
    vertices: np.ndarray
    triangles: np.ndarray
    edges: np.ndarray = field(init=False)
    plus: np.ndarray = field(init=False)
    minus: np.ndarray = field(init=False)
    free_plus: np.ndarray = field(init=False)
    free_minus: np.ndarray = field(init=False)
    length: np.ndarray = field(init=False)

    def __post_init__(self):
        V = np.asarray(self.vertices, np.float64)
        T = np.asarray(self.triangles, np.int64)
        seen = {}
        for t, (a, b, c) in enumerate(T):
            for i, j, opp in ((a, b, c), (b, c, a), (c, a, b)):
                seen.setdefault((min(i, j), max(i, j)), []).append((t, opp))
        keys = sorted(k for k, v in seen.items() if len(v) == 2)
        self.edges = np.array(keys, np.int64).reshape(-1, 2)
        sides = [seen[k] for k in keys]
        self.plus = np.array([s[0][0] for s in sides], np.int64)
        self.minus = np.array([s[1][0] for s in sides], np.int64)
        self.free_plus = np.array([s[0][1] for s in sides], np.int64)
        self.free_minus = np.array([s[1][1] for s in sides], np.int64)
        self.length = np.linalg.norm(V[self.edges[:, 0]] - V[self.edges[:, 1]], axis=1)

    def __len__(self):
        return self.edges.shape[0]

    @property
    def centers(self):
        """Edge midpoints: the positions the cluster trees sort the
        functions by."""
        V = np.asarray(self.vertices, np.float64)
        return 0.5 * (V[self.edges[:, 0]] + V[self.edges[:, 1]])

    def geometry(self):
        """Per function, plus then minus triangle: 13 doubles each -- the
        three vertices, the free vertex v and c with f(x) = c (x - v)
        (c = +l/(2 A+) on the plus, -l/(2 A-) on the minus triangle, so
        div f = 2 c): shape (n, 26)."""
        V = np.asarray(self.vertices, np.float64)
        T = np.asarray(self.triangles, np.int64)
        out = np.zeros((len(self), 26))
        for side, (tri, free, sgn) in enumerate(((self.plus, self.free_plus, 1.0),
                                                 (self.minus, self.free_minus, -1.0))):
            P = V[T[tri]]
            area = 0.5 * np.linalg.norm(np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]), axis=1)
            o = 13 * side
            out[:, o:o + 9] = P.reshape(-1, 9)
            out[:, o + 9:o + 12] = V[free]
            out[:, o + 12] = sgn * self.length / (2 * area)
        return out

    def samples(self, degree):
        """Quadrature samples of every function: (n, 2 q, 7) with rows
        (x, y, z, w A f(x) (3), w A div f) over the plus then the minus
        triangle (A f = l/2 (x - v+) or l/2 (v- - x), A div f = +-l)."""
        b, w = triangle_rule(degree)
        V = np.asarray(self.vertices, np.float64)
        T = np.asarray(self.triangles, np.int64)
        out = np.zeros((len(self), 2 * len(w), 7))
        half = 0.5 * self.length[:, None, None]
        for side, (tri, free, sgn) in enumerate(((self.plus, self.free_plus, 1.0),
                                                 (self.minus, self.free_minus, -1.0))):
            P = V[T[tri]]                                  # (n, 3 vertices, 3)
            x = np.einsum("qk,nkd->nqd", b, P)             # (n, q, 3)
            f = sgn * half * (x - V[free][:, None, :])     # A f
            s = slice(side * len(w), (side + 1) * len(w))
            out[:, s, 0:3] = x
            out[:, s, 3:6] = w[None, :, None] * f
            out[:, s, 6] = sgn * w[None, :] * self.length[:, None]
        return out


@dataclass
class EFIEKernelMatrix:
    """Maxwell single layer (EFIE) between the RWG functions `rwg` (test)
    and `trial_rwg` (trial; default the same), far interactions by double
    quadrature with rules of degree `rules` = (outer, inner) on test / trial
    triangles.  Z = alpha <f, G g> + beta <div f, G div g> with
    G = exp(-ik|x-y|)/(4 pi |x-y|), alpha = -ik, beta = -1/(ik) by default
    (gamma = ik).  Near-field blocks (``near_block``, and the near field of
    ``PetrovGalerkinBF``) use singularity extraction for triangle pairs
    closer than ``near_factor`` longest edges (outer / inner rule degrees
    ``near_rules``; closed-form static part, Wilton et al. 1984) and the
    double rule ``near_far_rules`` for the other pairs."""
    rwg: RWG
    k: float
    trial_rwg: RWG = None
    alpha: complex = None
    beta: complex = None
    rules: tuple = (2, 3)
    near_rules: tuple = (5, 5)
    near_far_rules: tuple = (5, 5)
    near_factor: float = 3.0
    _handles: dict = field(default_factory=dict, repr=False)

    def __post_init__(self):
        if self.trial_rwg is None:
            self.trial_rwg = self.rwg
        if self.alpha is None:
            self.alpha = -1j * self.k
        if self.beta is None:
            self.beta = -1.0 / (1j * self.k)

    def handle(self, ctx):
        """The library's plug-in handle on context `ctx` (built once)."""
        import ctypes as C
        from . import _lib as L
        from ._lib import check, lib
        key = id(ctx)
        if key not in self._handles:
            ts = np.ascontiguousarray(self.rwg.samples(self.rules[0]))
            ss = np.ascontiguousarray(self.trial_rwg.samples(self.rules[1]))
            a = (C.c_double * 2)(float(np.real(self.alpha)), float(np.imag(self.alpha)))
            b = (C.c_double * 2)(float(np.real(self.beta)), float(np.imag(self.beta)))
            h = C.c_void_p()
            dp = L.P(C.c_double)
            check(lib().bfgpu_kernel_rwg_create(ctx.handle, a, b, ts.shape[0], ts.shape[1],
                                                ts.ctypes.data_as(dp), ss.shape[0], ss.shape[1],
                                                ss.ctypes.data_as(dp), C.byref(h)))
            self._handles[key] = (h, ctx)
            tg = np.ascontiguousarray(self.rwg.geometry())
            sg = np.ascontiguousarray(self.trial_rwg.geometry())
            check(lib().bfgpu_kernel_rwg_set_near(h, tg.ctypes.data_as(dp), sg.ctypes.data_as(dp),
                                                  int(self.near_rules[0]), int(self.near_rules[1]),
                                                  int(self.near_far_rules[0]), int(self.near_far_rules[1]),
                                                  float(self.near_factor)))
        return self._handles[key][0]

    def block(self, rows, cols, acctype=np.complex128, ctx=None):
        """Z[rows, cols] computed by the plug-in's device fill."""
        import ctypes as C
        from . import _lib as L
        from ._lib import check, lib
        from . import default_context
        ctx = ctx or default_context()
        rows = np.ascontiguousarray(rows, np.int32)
        cols = np.ascontiguousarray(cols, np.int32)
        dt = np.dtype(acctype)
        Z = np.zeros((len(rows), len(cols)), dt, order="F")
        prec = L.BFGPU_C128 if dt == np.complex128 else L.BFGPU_C64
        check(lib().bfgpu_kernel_fill(self.handle(ctx), float(self.k), prec, len(rows),
                                      rows.ctypes.data_as(L.P(C.c_int32)), len(cols),
                                      cols.ctypes.data_as(L.P(C.c_int32)), C.c_void_p(Z.ctypes.data)))
        return Z

    def near_block(self, rows, cols, acctype=np.complex128, ctx=None):
        """Z[rows, cols] of the near-field quadrature, computed on the device."""
        import ctypes as C
        from . import _lib as L
        from ._lib import check, lib
        from . import default_context
        ctx = ctx or default_context()
        rows = np.ascontiguousarray(rows, np.int32)
        cols = np.ascontiguousarray(cols, np.int32)
        dt = np.dtype(acctype)
        Z = np.zeros((len(rows), len(cols)), dt, order="F")
        prec = L.BFGPU_C128 if dt == np.complex128 else L.BFGPU_C64
        check(lib().bfgpu_kernel_fill_near(self.handle(ctx), float(self.k), prec, len(rows),
                                           rows.ctypes.data_as(L.P(C.c_int32)), len(cols),
                                           cols.ctypes.data_as(L.P(C.c_int32)), C.c_void_p(Z.ctypes.data)))
        return Z

    def __del__(self, _destroy=_destroy):
        for h, _ctx in getattr(self, "_handles", {}).values():
            _destroy("bfgpu_kernel_destroy", h)
        self._handles = {}
