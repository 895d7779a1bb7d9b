"""FP64 configuration at full size: N = 65536 points in [0,1]^2, l = 0.1,
leaf 32, eps = 1e-6, max-diameter admissibility, through the TLR path
(createColumnMajorLRMatrix, Morton conversion, generateHMatrixFromStruct):
KD-tree permutation and offsets, structure tile lists and JDS bit for bit
against the oracle; the H-matvec with 1 and 16 right-hand sides against 1024
sampled rows of the dense FP64 kernel (relative error <= eps).

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


# ---- FP64, N = 65536, 2D --------------------------------------------------------
@pytest.fixture(scope="module")
def c2(torch_cuda, ctx):
    n, d, b, tol = 65536, 2, 32, 1e-6
    print(f"[configs] building n={n} d={d} b={b} eps={tol} (TLR path)", flush=True)
    p = O.generate_points(n, d, 1234)
    pipe = hmx.Pipeline(n, d, b, tol, adm="maxDiam", max_rank_policy=hmx.MAXRANK_FULL, points=to_dev(torch_cuda, p),
                        ctx=ctx)
    print(f"[configs] built {pipe.timings}", flush=True)
    yield pipe, p
    del pipe
    gc.collect()
    torch_cuda.cuda.empty_cache()


def test_c2_tree_structure_jds_bitwise(c2, ctx):
    pipe, p = c2
    n, d, b = pipe.n, pipe.d, pipe.b
    perm, off, bbmin, bbmax = pipe.tree.copy(ctx, n, d)
    rperm, roff, rmin, rmax, _ = O.kdtree(p, n, d, b)
    assert np.array_equal(perm, rperm) and np.array_equal(off, roff)
    ref = O.structure(n, d, b, "maxDiam", 1.0, rmin, rmax)
    L = pipe.structure.num_levels()
    for l in range(1, L):
        tiles = pipe.structure.tiles(l)
        assert np.array_equal(tiles, ref[l - 1]), l
        ri, ci, tp, ip = pipe.structure.jds(l)
        ro, co, to, io = O.jds(l, tiles)
        assert np.array_equal(ri, ro) and np.array_equal(ci, co) and np.array_equal(tp, to) and np.array_equal(ip, io), l


@pytest.mark.parametrize("k", [1, 16])
def test_c2_matvec_vs_dense_rows(c2, ctx, torch_cuda, k):
    torch = torch_cuda
    pipe, p = c2
    perm = pipe.tree.copy(ctx, pipe.n, pipe.d)[0]
    X = torch.from_numpy(np.ascontiguousarray(O.generate_rhs(pipe.n, k, 21 + k).T)).cuda()
    Y = pipe.matvec(X if k > 1 else X[0].contiguous()).reshape(k, pipe.n)
    rows, yd = dense_rows_check(torch, ctx, pipe, p, perm, X, 1024, 128, 2 + k)
    e = rel(torch, Y[:, rows], yd)
    print(f"N=65536 2D, {k} RHS: rel err vs 1024 dense rows {e:.3e}")
    assert e <= 1e-6
