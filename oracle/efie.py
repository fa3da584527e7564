"""CPU FP64 restatement of the EFIE / RWG kernel-matrix entries that the
BEAST plug-in computes for far interactions (ext/BFBEAST/kernelmatrix.jl:
22-82: Maxwell single layer, Raviart-Thomas / RWG functions, double
numerical quadrature DoubleNumQStrat(2, 3), read here as triangle rules of
polynomial degree 2 on test and 3 on trial triangles).

TEST INFRASTRUCTURE ONLY (the reference results for the GPU plug-in); the product never
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


def potential_integrals(P, X):
    """Closed-form potential integrals of a flat triangle P (3 x 3 vertices)
    at observation points X (N x 3), by the published method of Wilton,
    Rao, Glisson, Schaubert, Al-Bundak and Butler (IEEE Trans. AP 32(3),
    1984):
        Im1 = int_P 1 / |x - y| dS_y,
        Iv  = int_P (rho_y - rho_x) / |x - y| dS_y   (in-plane vector),
    rho the projection onto the plane of P.  Per edge i (endpoints a -> b,
    unit tangent l, outward in-plane normal u = l x n, n the triangle's unit
    normal for its vertex order), with d the height of x above the plane:
        l+- = (b|a - rho) . l,  P0 = (a - rho) . u,  R0^2 = P0^2 + d^2,
        R+- = sqrt(R0^2 + l+-^2),  f = ln((R+ + l+) / (R- + l-)),
        Im1 = sum_i P0 f - |d| [atan(P0 l+ / (R0^2 + |d| R+)) - atan(P0 l- / (R0^2 + |d| R-))],
        Iv  = 1/2 sum_i u (R0^2 f + l+ R+ - l- R-).
    f is taken in the form without cancellation (R0^2 = R^2 - l^2); an edge
    whose line passes through the projected point (R0 = 0) adds only its
    l R terms."""
    P = np.asarray(P, np.float64)
    X = np.atleast_2d(np.asarray(X, np.float64))
    nrm = np.cross(P[1] - P[0], P[2] - P[0])
    n = nrm / np.linalg.norm(nrm)
    d = (X - P[0]) @ n
    rho = X - d[:, None] * n
    ad = np.abs(d)
    im1 = np.zeros(len(X))
    iv = np.zeros((len(X), 3))
    scale = max(np.linalg.norm(P[1] - P[0]), np.linalg.norm(P[2] - P[1]), np.linalg.norm(P[0] - P[2]))
    for i in range(3):
        a, b = P[i], P[(i + 1) % 3]
        l = (b - a) / np.linalg.norm(b - a)
        u = np.cross(l, n)
        lp = (b - rho) @ l
        lm = (a - rho) @ l
        p0 = (a - rho) @ u
        r02 = p0 * p0 + d * d
        rp = np.sqrt(r02 + lp * lp)
        rm = np.sqrt(r02 + lm * lm)
        ok = r02 > (1e-14 * scale) ** 2
        with np.errstate(divide="ignore", invalid="ignore"):
            f = np.where(lm >= 0, np.log((rp + lp) / (rm + lm)),
                         np.where(lp <= 0, np.log((rm - lm) / (rp - lp)),
                                  np.log((rp + lp) * (rm - lm) / r02)))
            f = np.where(ok, f, 0.0)
            at = np.where(ok, np.arctan2(p0 * lp, r02 + ad * rp) - np.arctan2(p0 * lm, r02 + ad * rm), 0.0)
        im1 += p0 * f - ad * at
        iv += 0.5 * (r02 * f + lp * rp - lm * rm)[:, None] * u[None, :]
    return im1, iv


def rwg_sides(V, T, fns, idx):
    """Per function in idx and side (plus, minus): triangle vertices
    (n, 2, 3, 3), free vertex (n, 2, 3) and c with f = c (x - free),
    div f = 2 c (n, 2)."""
    E, tp, tm, op, om = fns
    V = np.asarray(V, np.float64)
    T = np.asarray(T)
    idx = np.asarray(idx)
    ln = np.linalg.norm(V[E[idx, 0]] - V[E[idx, 1]], axis=1)
    tri, free, c = [], [], []
    for t, o, sg in ((tp, op, 1.0), (tm, om, -1.0)):
        P = V[T[t[idx]]]
        area = 0.5 * np.linalg.norm(np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]), axis=1)
        tri.append(P)
        free.append(V[o[idx]])
        c.append(sg * ln / (2 * area))
    return np.stack(tri, 1), np.stack(free, 1), np.stack(c, 1)


def _g_smooth(k, r):
    """(exp(-ikr) - 1) / (4 pi r), with its limit -ik / (4 pi) at r = 0."""
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(r > 0, np.expm1(-1j * k * r) / (4 * np.pi * r), -1j * k / (4 * np.pi))


def _pair_extracted(k, Pm, fm, cm, Pn, fn, cn, alpha, beta, rules):
    """Entry part of test triangle Pm (f = cm (x - fm)) against trial
    triangle Pn (cn, fn) with the static kernel 1/(4 pi R) integrated in
    closed form over Pn (potential_integrals) and the rest
    (exp(-ikR) - 1)/(4 pi R) by the inner rule; outer rule over Pm."""
    bo, wo = RULES[rules[0]]
    bi, wi = RULES[rules[1]]
    am = 0.5 * np.linalg.norm(np.cross(Pm[1] - Pm[0], Pm[2] - Pm[0]))
    an = 0.5 * np.linalg.norm(np.cross(Pn[1] - Pn[0], Pn[2] - Pn[0]))
    x = bo @ Pm                                       # outer points
    y = bi @ Pn                                       # inner points
    im1, iv = potential_integrals(Pn, x)
    nrm = np.cross(Pn[1] - Pn[0], Pn[2] - Pn[0])
    nrm /= np.linalg.norm(nrm)
    rho = x - ((x - Pn[0]) @ nrm)[:, None] * nrm
    # int_Pn (y - fn) G dS_y and int_Pn G dS_y at every outer point
    vec = (iv + (rho - fn) * im1[:, None]) / (4 * np.pi)
    sca = im1 / (4 * np.pi)
    r = np.linalg.norm(x[:, None, :] - y[None, :, :], axis=2)
    gs = _g_smooth(k, r) * (wi * an)[None, :]
    vec = vec + gs @ (y - fn)
    sca = sca + gs.sum(axis=1)
    fx = cm * (x - fm)
    return np.sum(wo * am * (alpha * cn * np.sum(fx * vec, axis=1) + beta * (2 * cm) * (2 * cn) * sca))


def _pair_regular(k, Pm, fm, cm, Pn, fn, cn, alpha, beta, rules):
    bo, wo = RULES[rules[0]]
    bi, wi = RULES[rules[1]]
    am = 0.5 * np.linalg.norm(np.cross(Pm[1] - Pm[0], Pm[2] - Pm[0]))
    an = 0.5 * np.linalg.norm(np.cross(Pn[1] - Pn[0], Pn[2] - Pn[0]))
    x = bo @ Pm
    y = bi @ Pn
    r = np.linalg.norm(x[:, None, :] - y[None, :, :], axis=2)
    G = np.exp(-1j * k * r) / (4 * np.pi * r)
    fx = cm * (x - fm)
    fy = cn * (y - fn)
    w = (wo * am)[:, None] * (wi * an)[None, :]
    return np.sum(w * G * (alpha * (fx @ fy.T) + beta * (2 * cm) * (2 * cn)))


def near_pair(Pm, Pn, near_factor):
    """Triangles whose centroids are closer than near_factor times the
    longer of their longest edges take the singularity extraction."""
    h = max(np.max(np.linalg.norm(Pm - np.roll(Pm, 1, axis=0), axis=1)),
            np.max(np.linalg.norm(Pn - np.roll(Pn, 1, axis=0), axis=1)))
    return np.linalg.norm(Pm.mean(axis=0) - Pn.mean(axis=0)) < near_factor * h


def efie_near_block(V, T, k, rows, cols, near_rules=(5, 5), rules=(5, 5), near_factor=3.0, alpha=None,
                    beta=None, trial_mesh=None):
    """Z[rows, cols] of the near-field quadrature: per pair of triangles,
    singularity extraction (outer rule near_rules[0] on the test triangle,
    closed-form static part and rule near_rules[1] for the rest on the
    trial triangle) when near_pair holds, else the double rule `rules`."""
    alpha = -1j * k if alpha is None else alpha
    beta = -1.0 / (1j * k) if beta is None else beta
    Vs, Ts = trial_mesh if trial_mesh is not None else (V, T)
    tm_, fm_, cm_ = rwg_sides(V, T, rwg_functions(V, T), rows)
    tn_, fn_, cn_ = rwg_sides(Vs, Ts, rwg_functions(Vs, Ts), cols)
    Z = np.zeros((len(rows), len(cols)), np.complex128)
    for i in range(len(rows)):
        for j in range(len(cols)):
            z = 0.0
            for sm in range(2):
                for sn in range(2):
                    Pm, Pn = tm_[i, sm], tn_[j, sn]
                    args = (k, Pm, fm_[i, sm], cm_[i, sm], Pn, fn_[j, sn], cn_[j, sn], alpha, beta)
                    if near_pair(Pm, Pn, near_factor):
                        z += _pair_extracted(*args, near_rules)
                    else:
                        z += _pair_regular(*args, rules)
            Z[i, j] = z
    return Z


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
