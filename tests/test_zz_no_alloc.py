"""The allocation-free matvec (bfgpu_matvec_noalloc) against bfgpu_matvec.

The entry point keeps the skeleton vectors in the library's static scratch
array; its products must equal the allocating entry point's on the same
device inputs (FP64 to 1e-12 relative, and bit for bit where both run the
same launches), on consecutive calls whose second call is smaller, and it
must reject a call whose skeleton vectors exceed the array.
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


def test_noalloc_matvec_matches_matvec(bfgpu, oracle):
    import torch
    ctx = bfgpu.default_context()
    dev = torch.device("cuda:0")

    say("noalloc: fp64 two squares 4096, 4 RHS")
    obs, src = two_squares(4096, seed=21)
    bt = bfgpu.BlockTree(bfgpu.ClusterTree(obs, dims=2, leaf_size=64),
                         bfgpu.ClusterTree(src, dims=2, leaf_size=64))
    g = bfgpu.assemble_BF(bfgpu.HelmholtzKernelMatrix(K), bt, 0, 0, K, 1e-4, ctx=ctx)
    rng = np.random.default_rng(5)
    X = np.asfortranarray(crandn(rng, 4096, 4))
    Xd = torch.from_numpy(X.T.copy()).to(dev).t()  # column-major on the device
    Y1 = torch.zeros((4, 4096), dtype=torch.complex128, device=dev).t()
    Y2 = torch.zeros((4, 4096), dtype=torch.complex128, device=dev).t()
    g.mul(Y1, Xd)
    g.mul_noalloc(Y2, Xd)
    torch.cuda.synchronize()
    y1, y2 = Y1.cpu().numpy(), Y2.cpu().numpy()
    assert rel(y2, y1) < 1e-12
    ref = oracle.import_bf(bt.testcluster, bt.trialcluster, g.export())
    assert rel(y2[:, 2], ref.matvec(X[:, 2])) < 1e-12

    say("noalloc: fp32 sphere 2000 far field, 1 RHS, forward and adjoint (smaller second call)")
    pts = fibonacci_sphere(2000, 3.0)
    T = bfgpu.ClusterTree(pts, dims=3, leaf_size=64, mode="maxpoints")
    B = bfgpu.BlockTree(T, T)
    op = bfgpu.PetrovGalerkinBF(bfgpu.HelmholtzKernelMatrix(K), B, K, tol=1e-4,
                                admissibility=bfgpu.CenterDistanceAdmissibility(2.0), acctype=np.complex64,
                                farfieldonly=True, ctx=ctx)
    assert op.info().slot_len < g.info().slot_len * 4
    xs = torch.from_numpy(crandn(np.random.default_rng(6), 2000).astype(np.complex64)).to(dev)
    for view, name in ((op, "forward"), (op.adjoint(), "adjoint")):
        say(f"noalloc: fp32 sphere {name}")
        a = torch.zeros(2000, dtype=torch.complex64, device=dev)
        b = torch.zeros(2000, dtype=torch.complex64, device=dev)
        view.mul(a, xs)
        view.mul_noalloc(b, xs)
        torch.cuda.synchronize()
        # one right-hand side: the same launches, so the same bits
        assert torch.equal(a, b), name

    say("noalloc: fp64 two squares again, 1 RHS with alpha and beta")
    y0 = crandn(np.random.default_rng(7), 4096)
    a = torch.from_numpy(y0).to(dev)
    b = torch.from_numpy(y0).to(dev)
    x1 = Xd[:, 1].contiguous()
    g.mul(a, x1, alpha=0.5 - 2j, beta=1.5 + 0.25j)
    g.mul_noalloc(b, x1, alpha=0.5 - 2j, beta=1.5 + 0.25j)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    ya = b.cpu().numpy()
    assert rel(ya, (0.5 - 2j) * ref.matvec(X[:, 1]) + (1.5 + 0.25j) * y0) < 1e-12

    say("noalloc: a call beyond the scratch array is rejected")
    limit = 64 * 1024 * 1024
    slot_len = int(g.info().slot_len)
    n_big = limit // (slot_len * 16) + 1
    Xb = torch.zeros((n_big, 4096), dtype=torch.complex128, device=dev).t()
    Yb = torch.zeros((n_big, 4096), dtype=torch.complex128, device=dev).t()
    with pytest.raises(bfgpu.ArgumentError):
        g.mul_noalloc(Yb, Xb)
