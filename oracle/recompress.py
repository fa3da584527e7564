"""CPU FP64 restatement of the reference's algebraic recompression
(src/Butterflyalgebra/algrecomp.jl:19-148), over factorizations in the
exported reference layout.

TEST INFRASTRUCTURE ONLY (the reference results for the GPU recompression); the
product never imports it.

A factorization is a dict as ``export()`` returns it: ``table`` rows
(stage, obs_out, src_out, obs_in, src_in, m, n, identity, out_ptr, in_ptr),
``data_ofs`` and ``data`` (complex128, each block column-major).  Stage 0 is
Q, stages 1..n_stages-2 the R levels, the last stage P.
"""
import numpy as np

from . import oracle


def pqr(A, rtol):
    """pqr(A; rtol) as algrecomp.jl:80 uses it: column-pivoted Householder
    QR truncated at the first step k >= 1 with |R_kk| <= rtol |R_11| (the
    oracle's CPQR, pinned against LAPACK ZGEQP3).  Returns the thin Q
    (m x r, formed from the reflectors as ZUNG2R does), R (r x n, in pivot
    order) and the pivot permutation p (A[:, p] = Q R up to truncation)."""
    A = np.asarray(A, np.complex128)
    m, n = A.shape
    r, piv, F, tau = oracle.cpqr(A, rtol)
    R = np.triu(F[:r, :])
    Q = np.zeros((m, r), np.complex128)
    Q[:r, :r] = np.eye(r)
    for k in range(r - 1, -1, -1):  # Q = H_1 ... H_r [I; 0]
        v = np.zeros(m - k, np.complex128)
        v[0] = 1.0
        v[1:] = F[k + 1:, k]
        w = v.conj() @ Q[k:, :]
        Q[k:, :] -= tau[k] * np.outer(v, w)
    return Q, R, piv.astype(np.int64)


def _levels(ex):
    """(Q, R levels, P) as lists of [obs_out, src_out, obs_in, src_in, data]
    (data None for identity blocks), plus the table columns kept as they are."""
    ns = int(ex["n_stages"])
    st = [[] for _ in range(ns)]
    for row, o in zip(ex["table"], ex["data_ofs"]):
        s, oo, so, oi, si, m, n, ident = (int(v) for v in row[:8])
        d = None
        if not ident:
            d = ex["data"][o:o + m * n].reshape((m, n), order="F").copy()
        st[s].append([oo, so, oi, si, d, int(row[8]), int(row[9]), m])
    return st


def _to_export(ex, st):
    table, ofs, data = [], [], []
    pos = 0
    for s, blocks in enumerate(st):
        for oo, so, oi, si, d, op, ip, mid in blocks:
            if d is None:
                table.append([s, oo, so, oi, si, mid, mid, 1, op, ip])
                ofs.append(-1)
            else:
                m, n = d.shape
                table.append([s, oo, so, oi, si, m, n, 0, op, ip])
                ofs.append(pos)
                data.append(np.asfortranarray(d).reshape(-1, order="F"))
                pos += m * n
    out = dict(ex)
    out["table"] = np.array(table, np.int32).reshape(-1, 10)
    out["data_ofs"] = np.array(ofs, np.int64)
    out["data"] = np.concatenate(data) if data else np.zeros(0, np.complex128)
    return out


def _right(R, tau):
    """recompress_BF_right (algrecomp.jl:46-148) on the R levels (lists of
    blocks [obs_out, src_out, obs_in, src_in, data, ...]), in place.
    Groups holding an identity block are left alone (the reference would
    fail to stack them)."""
    lr = len(R)
    for l in range(1, lr):
        lold = lr - l  # 0-based level lr - l + 1
        cur = R[lold]
        groups = {}
        for b in cur:
            groups.setdefault((b[2], b[3]), []).append(b)  # column key (obs_in, src_in)
        R_u = {}
        for key in sorted(groups):
            blocks = groups[key]
            if any(b[4] is None for b in blocks):
                continue
            A = np.vstack([b[4] for b in blocks])
            Q, Rm, p = pqr(A, tau)
            T = np.zeros_like(Rm)
            T[:, p] = Rm  # R[:, invperm(p)]
            row = 0
            for b in blocks:
                mb = b[4].shape[0]
                b[4] = Q[row:row + mb, :].copy()
                row += mb
            R_u[key] = T
        for b in R[lold - 1]:
            T = R_u.get((b[0], b[1]))  # row key (obs_out, src_out)
            if T is not None:
                b[4] = T.copy() if b[4] is None else T @ b[4]  # T * I = T


def _adjoint(R):
    """The R levels of the adjoint (bfadjtr.jl:13-58): levels reversed,
    keys swapped, data conjugate-transposed."""
    out = []
    for lvl in reversed(R):
        out.append([[b[3], b[2], b[1], b[0], None if b[4] is None else b[4].conj().T.copy()] + b[5:]
                    for b in lvl])
    return out


def recompress(ex, tau):
    """recompress_BF (algrecomp.jl:19-33): right recompression, then left
    recompression as the adjoint of the right one of the adjoint."""
    st = _levels(ex)
    R = st[1:-1]
    _right(R, tau)
    Ra = _adjoint(R)
    _right(Ra, tau)
    R = _adjoint(Ra)
    return _to_export(ex, [st[0]] + R + [st[-1]])
