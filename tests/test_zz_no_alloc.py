"""hmx_HMatrixVecMultNoAlloc: the H-matvec that allocates nothing (its
workspaces live in a fixed per-GPU array), called through the C ABI and
checked against hmx_HMatrixVecMult on the same H-matrix and inputs.

Bars: one right-hand side runs the same kernels in the same order as
hmx_HMatrixVecMult, so the outputs are equal bit for bit; wider calls run the
streaming kernels in passes where hmx_HMatrixVecMult runs the matrix cores,
so they agree within 1e-13 relative.  Two calls in a row, the second on a
smaller H-matrix (partial leaves: the padded-space gather and scatter), so the
second runs on the array the first left behind.  An H-matrix whose passes
need more than the array holds is rejected with HMX_EINVAL, and one without a
streaming plan with HMX_ESTATE.
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


def noalloc(ctx, pipe, x, y):
    k = 1 if x.dim() == 1 else x.shape[0]
    return hmx.lib().hmx_HMatrixVecMultNoAlloc(ctx.handle, pipe.H.handle, k, x.data_ptr(), y.data_ptr())


def build(torch, ctx, n, d, b, adm="weak", seed=61):
    p = O.generate_points(n, d, seed)
    return hmx.Pipeline(n, d, b, 1e-6, adm=adm, max_rank_policy=hmx.MAXRANK_FULL, points=to_dev(torch, p), ctx=ctx,
                        direct=True)


def test_no_alloc_matches_hmatvec_two_calls(torch_cuda, ctx):
    torch = torch_cuda
    big = build(torch, ctx, 16384, 3, 64, adm="maxDiam")
    small = build(torch, ctx, 3000, 2, 32)
    g = torch.Generator(device="cuda").manual_seed(5)
    for pipe in (big, small):  # the second call is the smaller input
        x = torch.randn(pipe.n, generator=g, device="cuda", dtype=torch.float64)
        ref = pipe.matvec(x)  # builds the streaming plan
        y = torch.full_like(x, float("nan"))
        st = noalloc(ctx, pipe, x, y)
        assert st == hmx.HMX_OK, hmx.lib().hmx_last_error(ctx.handle)
        torch.cuda.synchronize()
        print(f"n={pipe.n}: no-alloc vs hmx_HMatrixVecMult equal: {torch.equal(y, ref)}")
        assert torch.equal(y, ref)
    # several right-hand sides (passes 16 + 4 + 1 on the streaming kernels)
    X = torch.randn(21, small.n, generator=g, device="cuda", dtype=torch.float64)
    ref = small.matvec(X)
    Y = torch.full_like(X, float("nan"))
    assert noalloc(ctx, small, X, Y) == hmx.HMX_OK
    rel = float(torch.linalg.norm(Y - ref) / torch.linalg.norm(ref))
    assert rel < 1e-13, rel
    # and the big one again after the small one left its contents behind
    x = torch.randn(big.n, generator=g, device="cuda", dtype=torch.float64)
    y = torch.empty_like(x)
    assert noalloc(ctx, big, x, y) == hmx.HMX_OK
    assert torch.equal(y, big.matvec(x))


def test_no_alloc_rejections(torch_cuda, ctx):
    torch = torch_cuda
    pipe = build(torch, ctx, 1024, 2, 32)
    x = torch.randn(pipe.n, device="cuda", dtype=torch.float64)
    y = torch.empty_like(x)
    assert noalloc(ctx, pipe, x, y) == hmx.HMX_ESTATE  # no streaming plan yet: nothing is built here
    pipe.matvec(x)
    assert noalloc(ctx, pipe, x, y) == hmx.HMX_OK
    assert torch.equal(y, pipe.matvec(x))
    # partial leaves of 2^17 + 1000 points (8192 leaves of 32 positions): in
    # passes of 16 right-hand sides the padded x and y copies alone
    # (2 * 262144 * 16 doubles) fill the 64 MiB, z and the partial sums come on top
    big = build(torch, ctx, (1 << 17) + 1000, 2, 32)
    X = torch.randn(16, big.n, device="cuda", dtype=torch.float64)
    Y = torch.empty_like(X)
    big.matvec(X[0].contiguous())
    assert noalloc(ctx, big, X, Y) == hmx.HMX_EINVAL
    msg = hmx.lib().hmx_last_error(ctx.handle).decode()
    assert "MiB" in msg, msg
    # one right-hand side of the same H-matrix fits
    y1 = torch.empty(big.n, device="cuda", dtype=torch.float64)
    assert noalloc(ctx, big, X[0].contiguous(), y1) == hmx.HMX_OK
    assert torch.equal(y1, big.matvec(X[0].contiguous()))
