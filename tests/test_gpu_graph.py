"""Per-level launches replayed from a captured HIP graph (BFGPU_OPT_GRAPH).

The reference's matvec is a per-call loop over levels and factorizations
(src/matrixalgebra/matrixvector.jl:41-61); with the graph option the same
launches are captured once per call signature (op, x, y, alpha, beta, RHS
count, stream) and replayed.  Every replay must equal the per-level
launches bit for bit (the same kernels run), with eager work and
synchronisations between replays; a new y pointer or new alpha/beta must
force a new capture; and the FP64 replays must match the oracle's matvec
to 1e-12.
"""
import sys

import numpy as np
import pytest

from conftest import crandn, fibonacci_sphere, two_squares

pytestmark = pytest.mark.gpu

K = 2 * np.pi


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _cols(torch, X, dev):
    """Column-major device copy of a host matrix."""
    return torch.from_numpy(np.ascontiguousarray(X.T)).to(dev).t()


def _check_replays(torch, view, x, y0, alpha, beta, n_replays=3):
    """Graph replays of y = alpha op(A) x + beta y0 against the per-level
    launches; eager torch work between replays."""
    view.set_option("graph", 0)
    ref = y0.clone()
    view.mul(ref, x, alpha, beta)
    torch.cuda.synchronize()
    view.set_option("graph", 1)
    y = y0.clone()
    for i in range(n_replays):
        y.copy_(y0)  # the same pointer each time: one capture, then replays
        view.mul(y, x, alpha, beta)
        junk = torch.randn(1 << 16, device=x.device).sum()  # eager work between replays
        torch.cuda.synchronize()
        assert torch.equal(y, ref), f"replay {i} differs"
        del junk
    view.set_option("graph", 0)
    return ref


def test_graph_replays_equal_per_level_launches(bfgpu, oracle):
    import torch
    dev = torch.device("cuda:0")
    ctx = bfgpu.default_context()

    say("graph: fp64 two squares 4096, N / C / T, alpha-beta, 3 RHS")
    obs, src = two_squares(4096, seed=31)
    bt = bfgpu.BlockTree(bfgpu.ClusterTree(obs, dims=2, leaf_size=64),
                         bfgpu.ClusterTree(src, dims=2, leaf_size=64))
    g = bfgpu.assemble_BF(bfgpu.HelmholtzKernelMatrix(K), bt, 0, 0, K, 1e-4, ctx=ctx)
    imp = oracle.import_bf(bt.testcluster, bt.trialcluster, g.export())
    rng = np.random.default_rng(8)
    xh = crandn(rng, 4096)
    yh = crandn(rng, 4096)
    x = torch.from_numpy(xh).to(dev)
    y0 = torch.from_numpy(yh).to(dev)
    for view, ref_op, name in ((g, imp, "N"), (g.adjoint(), imp.transform(1), "C"),
                               (g.transpose(), imp.transform(0), "T")):
        say("graph fp64", name)
        out = _check_replays(torch, view, x, y0, 1.0, 0.0)
        assert rel(out.cpu().numpy(), ref_op.matvec(xh)) < 1e-12
        out = _check_replays(torch, view, x, y0, 0.5 - 1j, 2.0 + 0.5j)
        assert rel(out.cpu().numpy(), ref_op.matvec(xh, alpha=0.5 - 1j, beta=2.0 + 0.5j, y=yh)) < 1e-12
    X = crandn(rng, 4096, 3)
    Xd, Y0 = _cols(torch, X, dev), _cols(torch, crandn(rng, 4096, 3), dev)
    out = _check_replays(torch, g, Xd, Y0, 1.0, 0.0)
    assert rel(out.cpu().numpy(), imp.matvec(X)) < 1e-12

    say("graph: a new y pointer and new alpha force a new capture")
    g.set_option("graph", 1)
    y1 = torch.zeros(4096, dtype=torch.complex128, device=dev)
    y2 = torch.zeros(4096, dtype=torch.complex128, device=dev)
    g.mul(y1, x)
    g.mul(y2, x)          # another y: recaptured, writes y2 only
    g.mul(y1, x, 2.0, 0)  # another alpha: recaptured
    torch.cuda.synchronize()
    g.set_option("graph", 0)
    ref = imp.matvec(xh)
    assert rel(y2.cpu().numpy(), ref) < 1e-12
    assert rel(y1.cpu().numpy(), 2.0 * ref) < 1e-12

    say("graph: fp32 sphere far field 6000, matrix-free last level, forward / adjoint, 40 RHS")
    pts = fibonacci_sphere(6000, 4.0)
    T = bfgpu.ClusterTree(pts, dims=3, leaf_size=64, mode="maxpoints")
    B = bfgpu.BlockTree(T, T)
    op = bfgpu.PetrovGalerkinBF(bfgpu.HelmholtzKernelMatrix(K), B, K, tol=1e-4,
                                admissibility=bfgpu.CenterDistanceAdmissibility(2.0), acctype=np.complex64,
                                farfieldonly=True, ctx=ctx)
    assert op.get_option("pfree") == 1
    xs = torch.from_numpy(crandn(rng, 6000).astype(np.complex64)).to(dev)
    ys = torch.from_numpy(crandn(rng, 6000).astype(np.complex64)).to(dev)
    for view, name in ((op, "N"), (op.adjoint(), "C"), (op.transpose(), "T")):
        say("graph fp32 sphere", name)
        _check_replays(torch, view, xs, ys, 1.0, 0.0)
        _check_replays(torch, view, xs, ys, 1.0 + 1j, -0.5)
    Xs = _cols(torch, crandn(rng, 6000, 40).astype(np.complex64), dev)
    Ys = torch.zeros((40, 6000), dtype=torch.complex64, device=dev).t()
    _check_replays(torch, op, Xs, Ys, 1.0, 0.0)
