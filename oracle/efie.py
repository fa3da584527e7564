"""CPU FP64 restatement of the EFIE / RWG kernel-matrix entries that the
BEAST plug-in computes for far interactions (ext/BFBEAST/kernelmatrix.jl:
22-82: Maxwell single layer, Raviart-Thomas / RWG functions, double
numerical quadrature DoubleNumQStrat(2, 3), read here as triangle rules of
polynomial degree 2 on test and 3 on trial triangles).

TEST INFRASTRUCTURE ONLY (the checker of the GPU plug-in); the product never
imports it.  BEAST itself is absent, so its exact rule tables cannot be
compared: the rules below are the published Strang-Fix / Dunavant (1985)
ones, pinned by exactness on monomials; the entries are pinned by symmetry,
quadrature convergence and the RWG identities (tests/test_efie.py).

    Z[m, n] = alpha int int f_m(x) . f_n(y) G(x, y) + beta int int div f_m div f_n G,
    G = exp(-ik|x-y|) / (4 pi |x-y|),  alpha = -ik, beta = -1/(ik).
"""
import numpy as np

_A4, _W4a, _B4, _W4b = 0.445948490915965, 0.223381589678011, 0.091576213509771, 0.109951743655322
_A5, _W5a, _B5, _W5b = 0.470142064105115, 0.132394152788506, 0.101286507323456, 0.125939180544827


def _orbit(a):
    return [[1 - 2 * a, a, a], [a, 1 - 2 * a, a], [a, a, 1 - 2 * a]]


RULES = {
    1: (np.array([[1, 1, 1]]) / 3.0, np.array([1.0])),
    2: (np.array([[4, 1, 1], [1, 4, 1], [1, 1, 4]]) / 6.0, np.ones(3) / 3.0),
    3: (np.array([[1 / 3, 1 / 3, 1 / 3], [0.6, 0.2, 0.2], [0.2, 0.6, 0.2], [0.2, 0.2, 0.6]]),
        np.array([-27.0, 25.0, 25.0, 25.0]) / 48.0),
    4: (np.array(_orbit(_A4) + _orbit(_B4)), np.array([_W4a] * 3 + [_W4b] * 3)),
    5: (np.array([[1 / 3, 1 / 3, 1 / 3]] + _orbit(_A5) + _orbit(_B5)), np.array([0.225] + [_W5a] * 3 + [_W5b] * 3)),
}


def rwg_functions(V, T):
    """Interior edges (shared by two triangles), sorted by vertex pair:
    (edge vertex pairs, plus triangle, minus triangle, vertex opposite the
    edge in each) with plus the lower-numbered triangle."""
    T = np.asarray(T, np.int64)
    E = np.sort(np.concatenate([T[:, [0, 1]], T[:, [1, 2]], T[:, [2, 0]]]), axis=1)
    opp = np.concatenate([T[:, 2], T[:, 0], T[:, 1]])
    tri = np.tile(np.arange(T.shape[0]), 3)
    order = np.lexsort((tri, E[:, 1], E[:, 0]))
    E, opp, tri = E[order], opp[order], tri[order]
    same = np.all(E[1:] == E[:-1], axis=1)
    i = np.nonzero(same)[0]
    return E[i], tri[i], tri[i + 1], opp[i], opp[i + 1]


def _samples(V, T, fns, idx, degree):
    """Points, w A f and w A div f of the functions idx over their plus and
    minus triangles: arrays (len(idx), 2q, 3), (.., 2q, 3), (.., 2q)."""
    E, tp, tm, op, om = fns
    bary, w = RULES[degree]
    V = np.asarray(V, np.float64)
    pts, fv, dv = [], [], []
    for tri, opv, s in ((tp, op, 1.0), (tm, om, -1.0)):
        P = V[np.asarray(T)[tri[idx]]]                         # (n, 3, 3)
        area = 0.5 * np.linalg.norm(np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]), axis=1)
        ln = np.linalg.norm(V[E[idx, 0]] - V[E[idx, 1]], axis=1)
        x = np.einsum("qk,nkd->nqd", bary, P)
        f = s * (ln / (2 * area))[:, None, None] * (x - V[opv[idx]][:, None, :])
        pts.append(x)
        fv.append(w[None, :, None] * area[:, None, None] * f)
        dv.append(w[None, :] * area[:, None] * (s * ln / area)[:, None])
    return np.concatenate(pts, 1), np.concatenate(fv, 1), np.concatenate(dv, 1)


def efie_block(V, T, k, rows, cols, rules=(2, 3), alpha=None, beta=None, trial_mesh=None):
    """Z[rows, cols] (complex128): test functions `rows` of mesh (V, T),
    trial functions `cols` of trial_mesh (default the same mesh)."""
    alpha = -1j * k if alpha is None else alpha
    beta = -1.0 / (1j * k) if beta is None else beta
    Vs, Ts = trial_mesh if trial_mesh is not None else (V, T)
    xa, fa, da = _samples(V, T, rwg_functions(V, T), np.asarray(rows), rules[0])
    yb, fb, db = _samples(Vs, Ts, rwg_functions(Vs, Ts), np.asarray(cols), rules[1])
    Z = np.zeros((len(rows), len(cols)), np.complex128)
    for i in range(len(rows)):
        d = xa[i][:, None, None, :] - yb[None, :, :, :]          # (qa, n, qb, 3)
        r = np.linalg.norm(d, axis=3)
        with np.errstate(divide="ignore", invalid="ignore"):
            G = np.where(r > 0, np.exp(-1j * k * r) / (4 * np.pi * r), 0.0)
        fg = np.einsum("ad,nbd->anb", fa[i], fb)
        de = da[i][:, None, None] * db[None, :, :]
        Z[i] = np.sum(G * (alpha * fg + beta * de), axis=(0, 2))
    return Z
