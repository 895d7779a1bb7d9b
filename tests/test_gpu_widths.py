"""Right-hand-side counts and partial leaves on every product the library
offers beside the forward FP64 one.

Bars:
  * the transposed H-matvec at widths that are not multiples of the
    streaming kernels' 4/8/16 (2, 3, 5, 12, 20) equals the transposed dense
    expansion of the factors within 1e-13, on full and on partial leaves;
  * factor gradients at such widths, on partial leaves: the loss
    sum((H X) * W) is linear in every factor block, so a central difference
    with step 1e-3 equals the directional derivative up to rounding (1e-8
    relative), and the padding rows of the gradients are zero;
  * the mixed-precision product on partial leaves and at widths 1, 5, 20, 64:
    within the target relative error of the FP64 product, and each column
    independent of the other columns in its call.
"""
import numpy as np
import pytest

import hmx
import oracle as O
from test_gpu_pipeline import expand_gpu, to_dev

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


@pytest.fixture(scope="module", params=[(4096, 2, 32), (3000, 2, 32), (5001, 3, 64)],
                ids=["full-2d-b32", "partial-2d-b32", "partial-3d-b64"])
def pipe(request, torch_cuda, ctx):
    n, d, b = request.param
    p = O.generate_points(n, d, 91)
    pp = hmx.Pipeline(n, d, b, 1e-6, adm="weak", max_rank_policy=hmx.MAXRANK_FULL, points=to_dev(torch_cuda, p),
                      ctx=ctx, direct=True)
    pp.E = expand_gpu(ctx, pp)
    return pp


@pytest.mark.parametrize("k", [2, 3, 5, 12, 20])
def test_transpose_any_width(torch_cuda, pipe, k):
    X = O.generate_rhs(pipe.n, k, 7 + k)
    y = pipe.matvec(to_dev(torch_cuda, X.T), transpose=True).cpu().numpy().T
    ref = pipe.E.T @ X
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-13


@pytest.mark.parametrize("k", [5, 12])
def test_factor_gradients_any_width(torch_cuda, pipe, k):
    from hmx.autograd import HMatrixOperator
    torch = torch_cuda
    op = HMatrixOperator(pipe)
    g = torch.Generator(device="cuda").manual_seed(k)
    X = torch.randn(k, pipe.n, generator=g, device="cuda", dtype=torch.float64)
    W = torch.randn(k, pipe.n, generator=g, device="cuda", dtype=torch.float64)
    (op(X) * W).sum().backward()
    h = 1e-3
    L = pipe.structure.num_levels()
    S = pipe.tree.info()[0]
    offs = pipe.tree.copy(pipe.ctx, pipe.n, pipe.d)[1]
    counts = np.diff(offs)
    rng = torch.Generator(device="cuda").manual_seed(100 + k)
    for l in op.levels[:: max(1, len(op.levels) // 4)]:
        for P in (op.U[l], op.V[l]):
            E = torch.randn(P.shape, generator=rng, device="cuda", dtype=torch.float64)
            analytic = float((P.grad * E).sum())
            with torch.no_grad():
                P.add_(h * E)
                lp = float((op(X) * W).sum())
                P.sub_(2 * h * E)
                lm = float((op(X) * W).sum())
                P.add_(h * E)
            fd = (lp - lm) / (2 * h)
            assert abs(fd - analytic) <= 1e-8 * max(1.0, abs(analytic)), (l, fd, analytic)
    op.sync()
    # gradients of the padding rows: tile rows map to padded positions of the
    # leaf space, and a position past a leaf's point count carries no point
    if pipe.n != S * pipe.b:
        padpos = np.concatenate([np.arange(c, pipe.b) + q * pipe.b for q, c in enumerate(counts)])
        for l in op.levels:
            m = (1 << (L - 1 - l)) * pipe.b
            ranks = pipe.H.level(pipe.ctx, l)[0]
            tiles = pipe.structure.tiles(l)
            dU = op.U[l].grad.cpu().numpy()
            dV = op.V[l].grad.cpu().numpy()
            off = np.concatenate([[0], np.cumsum(ranks)])
            for t in range(len(tiles)):
                u, v = O._mo_to_xy(int(tiles[t]))
                r = int(ranks[t])
                if r == 0:
                    continue
                rows_pad = padpos[(padpos >= v * m) & (padpos < (v + 1) * m)] - v * m
                cols_pad = padpos[(padpos >= u * m) & (padpos < (u + 1) * m)] - u * m
                gu = dU[off[t] * m:(off[t] + r) * m].reshape(r, m)
                gv = dV[off[t] * m:(off[t] + r) * m].reshape(r, m)
                assert np.all(gu[:, rows_pad] == 0) and np.all(gv[:, cols_pad] == 0), (l, t)


@pytest.fixture(scope="module")
def mixed_pipe(torch_cuda, ctx):
    n, d, b = 5001, 3, 64
    p = O.generate_points(n, d, 92)
    pp = hmx.Pipeline(n, d, b, 1e-6, adm="weak", max_rank_policy=hmx.MAXRANK_FULL, points=to_dev(torch_cuda, p),
                      ctx=ctx, direct=True)
    pp.make_mixed(3e-6)
    return pp


@pytest.mark.parametrize("k", [1, 5, 20, 64])
def test_mixed_partial_leaves_any_width(torch_cuda, mixed_pipe, k):
    torch = torch_cuda
    pp = mixed_pipe
    g = torch.Generator(device="cuda").manual_seed(k)
    X = torch.randn(64, pp.n, generator=g, device="cuda", dtype=torch.float32)
    full = pp.matvec_mixed(X)
    x = X[:k].contiguous() if k > 1 else X[0].contiguous()
    y = pp.matvec_mixed(x).reshape(k, pp.n)
    y64 = pp.matvec(x.double().contiguous()).reshape(k, pp.n)
    rel = float(torch.linalg.norm(y.double() - y64) / torch.linalg.norm(y64))
    print(f"mixed k={k} partial leaves: rel vs FP64 {rel:.3e}")
    assert rel <= 3e-6
    # a column's result does not depend on the other columns of its call
    # (per-column right-hand-side scales, columns past the count masked)
    col = float(torch.linalg.norm(y.double() - full[:k].double()) / torch.linalg.norm(full[:k].double()))
    assert col < 1e-6, col
