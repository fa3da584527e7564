"""Writes tests/golden/cpqr_zgeqp3.npz: column-pivoted QR of seeded complex
matrices computed by LAPACK ZGEQP3 (through scipy), the textbook routine the
oracle's pivoted QR restates.  Shapes keep min(m, n) below LAPACK's blocking
crossover so ZGEQP3 runs its unblocked ZLAQP2 path.  Run from the repo root:
    python tests/golden/make_golden.py
"""
import os

import numpy as np
from scipy.linalg import lapack

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpqr_zgeqp3.npz")


def main():
    rng = np.random.default_rng(20240601)
    shapes = [(8, 8), (12, 30), (30, 12), (43, 64), (45, 120), (64, 43), (20, 90), (90, 20)]
    rec = {}
    for i, (m, n) in enumerate(shapes):
        # mix of generic and numerically low-rank matrices
        if i % 2 == 0:
            A = (rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))) / np.sqrt(2)
        else:
            r = min(m, n) // 3 + 1
            U = rng.standard_normal((m, r)) + 1j * rng.standard_normal((m, r))
            V = rng.standard_normal((r, n)) + 1j * rng.standard_normal((r, n))
            s = np.logspace(0, -9, r)
            A = (U * s) @ V + 1e-12 * (rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n)))
        A = np.asfortranarray(A)
        qr, jpvt, tau, work, info = lapack.zgeqp3(A.copy(order="F"))
        assert info == 0
        rec[f"A{i}"] = A
        rec[f"jpvt{i}"] = (jpvt - 1).astype(np.int32)
        rec[f"rdiag{i}"] = np.abs(np.diag(qr))
        rec[f"R{i}"] = np.triu(qr[: min(m, n), :])
    rec["n"] = np.array(len(shapes))
    np.savez_compressed(OUT, **rec)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
