"""Precision configuration at full size: N = 2^20 points in [0,1]^3,
l = 0.2, leaf 64, eps = 1e-5, direct assembly: FP64 with 1 and 64 right-hand
sides against 2048 sampled dense rows (<= 1e-6); the mixed FP32/FP16 product
with 64 right-hand sides against the FP64 product and the dense rows
(<= 1e-5).

The dense rows are the reference's EXPAND_MATRIX check (src/expandMatrix.cu:
202-234) restricted to sampled rows, computed by the GPU kernel
hmx_denseRowsXVector for every sampled row and by the CPU oracle for a subset
of them (the two agree within 1e-12, which ties the full sample to the
oracle).  One module per configuration: a module's H-matrix is freed before
the next one is built.
"""
import gc

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


def dense_rows_check(torch, ctx, pipe, host_points, perm, X, nrows, ncpu, seed):
    """GPU dense rows for nrows sampled rows, CPU oracle for the first ncpu of
    them; returns (rows, GPU dense rows (k x nrows) on the device)."""
    n, d = pipe.n, pipe.d
    g = torch.Generator().manual_seed(seed)
    rows = torch.randperm(n, generator=g)[:nrows].to(torch.int32)
    k = X.shape[0]
    yd = torch.empty(k, nrows, dtype=torch.float64, device="cuda")
    hmx.denseRowsXVector(ctx, n, d, pipe.points, pipe.tree.device_perm(), pipe.ell, rows.cuda(), k, X, yd)
    Xh = X[:, :].cpu().numpy().T
    yo = O.sampled_rows(host_points, n, d, perm, pipe.ell, rows[:ncpu].numpy(), Xh)
    yg = yd[:, :ncpu].cpu().numpy().T
    assert np.linalg.norm(yg - yo) <= 1e-12 * np.linalg.norm(yo)
    return rows.long().cuda(), yd


def rel(torch, a, b):
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))


# ---- precision, N = 2^20, 3D -------------------------------------------------------
@pytest.fixture(scope="module")
def c3(torch_cuda, ctx):
    n, d, b, tol = 1 << 20, 3, 64, 1e-5
    print(f"[configs] building n={n} d={d} b={b} eps={tol} (direct)", flush=True)
    p = O.generate_points(n, d, 1234)
    pipe = hmx.Pipeline(n, d, b, tol, adm="maxDiam", max_rank_policy=hmx.MAXRANK_FULL, points=to_dev(torch_cuda, p),
                        ctx=ctx, direct=True)
    print(f"[configs] built {pipe.timings}", flush=True)
    perm = pipe.tree.copy(ctx, n, d)[0]
    X = torch_cuda.from_numpy(np.ascontiguousarray(O.generate_rhs(n, 64, 31).T)).cuda()
    rows, yd = dense_rows_check(torch_cuda, ctx, pipe, p, perm, X, 2048, 64, 5)
    yield pipe, X, rows, yd
    del pipe
    gc.collect()
    torch_cuda.cuda.empty_cache()


def test_c3_fp64_vs_dense_rows(c3, torch_cuda):
    torch = torch_cuda
    pipe, X, rows, yd = c3
    y1 = pipe.matvec(X[0].contiguous())
    e1 = rel(torch, y1[rows], yd[0])
    Y = pipe.matvec(X)
    e64 = rel(torch, Y[:, rows], yd)
    print(f"N=2^20 3D FP64: 1 RHS rel err {e1:.3e}, 64 RHS {e64:.3e} (2048 dense rows)")
    assert e1 <= 1e-6 and e64 <= 1e-6
    # the two FP64 paths (streaming, matrix cores) agree
    assert rel(torch, Y[0], y1) < 1e-13


def test_c3_mixed_vs_fp64_and_dense(c3, torch_cuda):
    torch = torch_cuda
    pipe, X, rows, yd = c3
    target = 1e-5
    info = pipe.make_mixed(target)
    x32 = X.float()
    y32 = pipe.matvec_mixed(x32).double()
    y64 = pipe.matvec(x32.double())
    ef = rel(torch, y32, y64)
    ed = rel(torch, y32[:, rows], yd)
    print(f"N=2^20 3D mixed: {info['loCols']} FP16 / {info['hiCols']} FP32 columns, {info['bytes']} bytes; "
          f"rel vs FP64 {ef:.3e}, vs dense rows {ed:.3e}")
    assert ef <= target and ed <= target
