"""Batching configuration: 256 clouds of sizes log-uniform in [2048, 16384],
half 2D and half 3D, l uniform in [0.05, 0.3], leaf 32, eps = 1e-6: every
cloud's KD-tree against the oracle bit for bit, the batched build and the
graph-replayed 8-RHS matvec against 256 sampled dense rows per cloud
(relative error <= eps).  256 clouds of up to 65536 points need about
280 GB of FP64 factors at eps = 1e-6, more than one GPU holds beside the
workspaces, so the sizes stop at 16384 here.

The dense rows are the reference's EXPAND_MATRIX check (src/expandMatrix.cu:
202-234) restricted to sampled rows, computed by the GPU kernel
hmx_denseRowsXVector for every sampled row and by the CPU oracle for a subset
of them (the two agree within 1e-12, which ties the full sample to the
oracle).  One module per configuration: a module's H-matrix is freed before
the next one is built.
"""
import numpy as np
import pytest

import hmx
import oracle as O
from test_gpu_pipeline import to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def ctx(torch_cuda):
    return hmx.Context()


def rel(torch, a, b):
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))


# ---- batching, 256 clouds ------------------------------------------------------------
def cloud_specs(count, seed, nmin=2048, nmax=16384):
    rng = np.random.default_rng(seed)
    ns = np.floor(np.exp(rng.uniform(np.log(nmin), np.log(nmax), count))).astype(int)
    ells = rng.uniform(0.05, 0.3, count)
    return [(int(ns[i]), 2 if i < count // 2 else 3, float(ells[i])) for i in range(count)]


def test_c4_batch_256_clouds(torch_cuda, ctx):
    torch = torch_cuda
    specs = cloud_specs(256, 2024)
    hp = [O.generate_points(n, d, 500 + i) for i, (n, d, ell) in enumerate(specs)]
    pts = [to_dev(torch, p) for p in hp]
    print(f"[configs] batched build of {len(specs)} clouds, {sum(s[0] for s in specs)} points", flush=True)
    bp = hmx.BatchPipeline(specs, b=32, tol=1e-6, ctx=ctx, points=pts)
    print(f"[configs] built {bp.timings}", flush=True)
    for i, (n, d, ell) in enumerate(specs):
        perm, off, _, _ = bp.trees[i].copy(ctx, n, d)
        rperm, roff, _, _, _ = O.kdtree(hp[i], n, d, 32)
        assert np.array_equal(perm, rperm) and np.array_equal(off, roff), i
    tot = bp.batch.info()["totalPoints"]
    g = torch.Generator("cuda").manual_seed(7)
    X = torch.randn(8, tot, dtype=torch.float64, device="cuda", generator=g)
    Y = torch.empty_like(X)
    bp.matvec(X, Y, graph=True)
    eager = bp.matvec(X)
    assert torch.equal(Y, eager)
    worst = 0.0
    for i, (n, d, ell) in enumerate(specs):
        a, e = bp.batch.offsets[i], bp.batch.offsets[i + 1]
        perm = bp.trees[i].copy(ctx, n, d)[0]
        xi = X[:, a:e].contiguous()
        gr = torch.Generator().manual_seed(i)
        rows = torch.randperm(n, generator=gr)[:256].to(torch.int32)
        yd = torch.empty(8, 256, dtype=torch.float64, device="cuda")
        hmx.denseRowsXVector(ctx, n, d, bp.points[i], bp.trees[i].device_perm(), ell, rows.cuda(), 8, xi, yd)
        if i % 16 == 0:  # the GPU dense rows against the CPU oracle
            yo = O.sampled_rows(hp[i], n, d, perm, ell, rows[:32].numpy(), xi.cpu().numpy().T)
            assert np.linalg.norm(yd[:, :32].cpu().numpy().T - yo) <= 1e-12 * np.linalg.norm(yo), i
        err = rel(torch, Y[:, a:e][:, rows.long().cuda()], yd)
        worst = max(worst, err)
        assert err <= 1e-6, (i, n, d, ell, err)
    print(f"256 clouds: worst rel err vs 256 dense rows per cloud {worst:.3e}")
