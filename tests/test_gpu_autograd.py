"""torch.autograd binding: input gradient against the dense kernel and the
transposed expansion, factor gradients against central finite differences
(step 1e-3, FP64; the loss is linear in each factor block, so the central
difference is exact up to rounding) and against per-tile outer products."""
import numpy as np
import pytest

import hmx
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def setup(torch_cuda):
    from hmx.autograd import HMatrixOperator
    torch = torch_cuda
    n, d, b, tol = 4096, 2, 32, 1e-6
    p = O.generate_points(n, d, 1234)
    pipe = hmx.Pipeline(n, d, b, tol, max_rank_policy=hmx.MAXRANK_FULL, points=torch.from_numpy(p).cuda())
    op = HMatrixOperator(pipe)
    perm = O.kdtree(p, n, d, b)[0]
    K = O.dense_matrix(p, n, d, perm, 0.1)
    return pipe, op, K


def test_input_gradient(torch_cuda, setup):
    torch = torch_cuda
    pipe, op, K = setup
    n = pipe.n
    g = torch.Generator(device="cuda").manual_seed(0)
    X = torch.randn(32, n, generator=g, device="cuda", dtype=torch.float32, requires_grad=True)
    W = torch.randn(32, n, generator=g, device="cuda", dtype=torch.float32)
    Y = op(X)
    assert Y.dtype == torch.float32 and Y.shape == X.shape
    (Y * W).sum().backward()
    Wn = W.double().cpu().numpy().T
    dX_dense = (K.T @ Wn).T
    rel = np.linalg.norm(X.grad.double().cpu().numpy() - dX_dense) / np.linalg.norm(dX_dense)
    assert rel < 1e-5
    # exactly the transposed H-matvec (up to the FP32 output rounding)
    dX_h = pipe.matvec(W.double().contiguous(), transpose=True)
    assert torch.allclose(X.grad.double(), dX_h, rtol=1e-6, atol=1e-6)


def test_factor_gradients_finite_differences(torch_cuda, setup):
    torch = torch_cuda
    pipe, op, K = setup
    n = pipe.n
    g = torch.Generator(device="cuda").manual_seed(1)
    X = torch.randn(32, n, generator=g, device="cuda", dtype=torch.float64)
    W = torch.randn(32, n, generator=g, device="cuda", dtype=torch.float64)
    for prm in op.parameters():
        prm.grad = None
    (op(X) * W).sum().backward()
    h = 1e-3
    rng = torch.Generator(device="cuda").manual_seed(2)
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
    # the library's H-matrix is back to its original factors
    op.sync()
    y = pipe.matvec(X)
    Kx = torch.from_numpy(K @ X.cpu().numpy().T).cuda().T
    assert float(torch.linalg.norm(y - Kx) / torch.linalg.norm(Kx)) < 1e-6


def test_factor_gradient_tile_outer_product(torch_cuda, setup):
    torch = torch_cuda
    pipe, op, K = setup
    n = pipe.n
    X = torch.randn(16, n, device="cuda", dtype=torch.float64)
    W = torch.randn(16, n, device="cuda", dtype=torch.float64)
    for prm in op.parameters():
        prm.grad = None
    (op(X) * W).sum().backward()
    L = pipe.structure.num_levels()
    l = op.levels[len(op.levels) // 2]
    ranks, Us, Vs = pipe.H.level(pipe.ctx, l)
    m = (1 << (L - 1 - l)) * pipe.b
    tiles = pipe.structure.tiles(l)
    Xn, Wn = X.cpu().numpy().T, W.cpu().numpy().T
    dU = op.U[l].grad.cpu().numpy()
    off = np.concatenate([[0], np.cumsum(ranks)])
    for t in range(min(4, len(tiles))):
        u, v = O._mo_to_xy(int(tiles[t]))
        r = ranks[t]
        ref = Wn[v * m:(v + 1) * m] @ (Vs[t].T @ Xn[u * m:(u + 1) * m]).T
        got = dU[off[t] * m:(off[t] + r) * m].reshape(r, m).T
        assert np.allclose(got, ref, rtol=1e-10, atol=1e-10)


# ---- full size: N = 2^18 points in [0,1]^2, l = 0.1, eps = 1e-6, 32 right-hand sides
@pytest.fixture(scope="module")
def big(torch_cuda):
    from hmx.autograd import HMatrixOperator
    n, d, b, tol = 1 << 18, 2, 32, 1e-6
    print(f"[autograd] building n={n}", flush=True)
    pipe = hmx.Pipeline(n, d, b, tol, max_rank_policy=hmx.MAXRANK_FULL, direct=True)
    return pipe, HMatrixOperator(pipe)


def test_full_size_input_gradient_vs_dense(torch_cuda, big):
    torch = torch_cuda
    pipe, op = big
    n = pipe.n
    g = torch.Generator(device="cuda").manual_seed(5)
    X = torch.randn(32, n, generator=g, device="cuda", dtype=torch.float32, requires_grad=True)
    W = torch.randn(32, n, generator=g, device="cuda", dtype=torch.float32)
    (op(X) * W).sum().backward()
    rows = torch.randperm(n, generator=torch.Generator().manual_seed(6))[:512].to(torch.int32).cuda()
    ref = torch.empty(32, 512, dtype=torch.float64, device="cuda")
    # K is symmetric: (K^T W)[rows] = K[rows, :] W
    hmx.denseRowsXVector(pipe.ctx, n, 2, pipe.points, pipe.tree.device_perm(), pipe.ell, rows, 32,
                         W.double().contiguous(), ref)
    got = X.grad.double()[:, rows.long()]
    rel = float(torch.linalg.norm(got - ref) / torch.linalg.norm(ref))
    print(f"dX vs dense K^T W on 512 rows: {rel:.3e}")
    assert rel < 1e-5


def test_full_size_factor_gradients_fd(torch_cuda, big):
    torch = torch_cuda
    pipe, op = big
    n = pipe.n
    g = torch.Generator(device="cuda").manual_seed(7)
    X = torch.randn(32, n, generator=g, device="cuda", dtype=torch.float64)
    W = torch.randn(32, n, generator=g, device="cuda", dtype=torch.float64)
    for prm in op.parameters():
        prm.grad = None
    (op(X) * W).sum().backward()
    h = 1e-3
    rng = torch.Generator(device="cuda").manual_seed(8)
    for l in (op.levels[0], op.levels[len(op.levels) // 2], op.levels[-1]):
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
            print(f"level {l}: analytic {analytic:.6e} fd {fd:.6e}")
            assert abs(fd - analytic) <= 1e-6 * max(1.0, abs(analytic)), (l, fd, analytic)
    op.sync()
