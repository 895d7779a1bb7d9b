"""Point clouds of uneven, non-power-of-two sizes and the batched H-matvec.

Bars:
  * partial leaves (32- and 64-row leaves; 1, 4 and 16 right-hand sides,
    transposed too): the H-matvec of an H-matrix over n points that do not
    fill the leaves equals the dense expansion of its own factors within 1e-13
    and the dense kernel product within the oracle's compression error;
  * a batch of independent H-matrices (2D and 3D clouds, per-cloud
    correlation length) applied in one launch per phase equals, bit for bit,
    each instance's own H-matvec; hipGraph replays equal the eager launches
    bit for bit, with eager work between the replays; after autotuning
    (another chunk length, another summation order) within 1e-14;
  * every instance against sampled rows of its dense FP64 kernel;
  * the batched datasets and KD-trees (one forest per point dimension) equal
    each cloud's own dataset and tree bit for bit: points, permutation,
    offsets and every bounding box the build writes.
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


@pytest.mark.parametrize("n,d,direct,b", [(3000, 2, False, 32), (5001, 3, True, 32), (2049, 2, True, 32),
                                          (5001, 3, True, 64)])
def test_partial_leaves(torch_cuda, ctx, n, d, direct, b):
    tol = 1e-6
    p = O.generate_points(n, d, 77)
    pipe = hmx.Pipeline(n, d, b, tol, adm="weak", max_rank_policy=hmx.MAXRANK_FULL, points=to_dev(torch_cuda, p),
                        ctx=ctx, direct=direct)
    X = O.generate_rhs(n, 4, 3)
    x = to_dev(torch_cuda, X.T)
    Y = pipe.matvec(x).cpu().numpy().T
    E = expand_gpu(ctx, pipe)
    assert np.linalg.norm(Y - E @ X) / np.linalg.norm(E @ X) < 1e-13
    # one right-hand side: the streaming kernels (with 64-row leaves the
    # per-leaf pass and the 128-row-block pass) over the padded space
    y1 = pipe.matvec(x[:1].contiguous()).cpu().numpy()[0]
    assert np.linalg.norm(y1 - E @ X[:, 0]) / np.linalg.norm(E @ X[:, 0]) < 1e-13
    yt = pipe.matvec(x, transpose=True).cpu().numpy().T
    assert np.linalg.norm(yt - E.T @ X) / np.linalg.norm(E.T @ X) < 1e-13
    K = O.dense_matrix(p, n, d, O.kdtree(p, n, d, b)[0], pipe.ell)
    err = np.linalg.norm(Y - K @ X) / np.linalg.norm(K @ X)
    ref = O.OracleHMatrix(p, n, d, b, tol=tol, max_rank_policy=1, seed=1234)
    err_o = np.linalg.norm(ref.matvec(X) - K @ X) / np.linalg.norm(K @ X)
    print(f"n={n} d={d} direct={direct}: rel err {err:.3e} (oracle {err_o:.3e})")
    assert err <= 2 * err_o + 1e-9
    # 16 right-hand sides: the matrix cores over the padded leaf space
    X16 = O.generate_rhs(n, 16, 5)
    Y16 = pipe.matvec(to_dev(torch_cuda, X16.T)).cpu().numpy().T
    assert np.linalg.norm(Y16 - E @ X16) / np.linalg.norm(E @ X16) < 1e-13


def cloud_specs(count, seed, nmin=2048, nmax=16384):
    """Sizes log-uniform in [nmin, nmax], first half 2D, second half 3D,
    correlation length uniform in [0.05, 0.3]."""
    rng = np.random.default_rng(seed)
    ns = np.floor(np.exp(rng.uniform(np.log(nmin), np.log(nmax), count))).astype(int)
    ells = rng.uniform(0.05, 0.3, count)
    return [(int(ns[i]), 2 if i < count // 2 else 3, float(ells[i])) for i in range(count)]


@pytest.fixture(scope="module")
def batch(torch_cuda, ctx):
    specs = cloud_specs(8, 11)
    pipes, pts = [], []
    for i, (n, d, ell) in enumerate(specs):
        p = O.generate_points(n, d, 100 + i)
        pts.append(p)
        pipes.append(hmx.Pipeline(n, d, 32, 1e-6, adm="weak", ell=ell, max_rank_policy=hmx.MAXRANK_FULL,
                                  points=to_dev(torch_cuda, p), ctx=ctx, direct=True))
    B = hmx.Batch(ctx, [pp.H for pp in pipes])
    return specs, pipes, pts, B


def test_batch_equals_instances(torch_cuda, ctx, batch):
    specs, pipes, pts, B = batch
    tot = B.info()["totalPoints"]
    assert tot == sum(n for n, _, _ in specs)
    k = 8
    X = torch_cuda.randn(k, tot, dtype=torch_cuda.float64, device="cuda")
    Y = B.matvec(X)
    for i, pp in enumerate(pipes):
        a, e = B.offsets[i], B.offsets[i + 1]
        yi = pp.matvec(X[:, a:e].contiguous())
        assert torch_cuda.equal(Y[:, a:e], yi), i
    # one right-hand side: the streaming kernels (per-leaf pass only in the
    # batch; an instance's own plan may add the row-block pass, another order)
    x1 = X[:1].contiguous()
    y1 = B.matvec(x1)
    for i, pp in enumerate(pipes):
        a, e = B.offsets[i], B.offsets[i + 1]
        yi = pp.matvec(x1[:, a:e].contiguous())
        rel = float(torch_cuda.linalg.norm(y1[:, a:e] - yi) / torch_cuda.linalg.norm(yi))
        assert rel < 1e-13, (i, rel)


def test_batch_graph_replay(torch_cuda, ctx, batch):
    specs, pipes, pts, B = batch
    tot = B.info()["totalPoints"]
    k = 8
    X = torch_cuda.randn(k, tot, dtype=torch_cuda.float64, device="cuda")
    ref = B.matvec(X)
    Y = torch_cuda.empty_like(X)
    for rep in range(4):
        Y.fill_(float("nan"))
        B.matvec(X, Y, graph=True)
        torch_cuda.cuda.synchronize()
        assert torch_cuda.equal(Y, ref), rep
        Z = B.matvec(X)  # eager work between replays
        assert torch_cuda.equal(Z, ref)
    # new inputs through the same captured pointers
    X.copy_(torch_cuda.randn(k, tot, dtype=torch_cuda.float64, device="cuda"))
    B.matvec(X, Y, graph=True)
    assert torch_cuda.equal(Y, B.matvec(X))


def test_batch_tune_and_dense_error(torch_cuda, ctx, batch):
    specs, pipes, pts, B = batch
    tot = B.info()["totalPoints"]
    k = 8
    X = torch_cuda.randn(k, tot, dtype=torch_cuda.float64, device="cuda")
    before = B.matvec(X)
    xc, ms = B.tune(k)
    assert xc in (0, 8, 16, 32) and ms > 0
    chunk, ms1 = B.tune(1)
    assert chunk in (512, 1024, 2048, 4096) and ms1 > 0
    assert B.info()["chunk"] == chunk
    after = B.matvec(X)
    assert float(torch_cuda.linalg.norm(after - before) / torch_cuda.linalg.norm(before)) < 1e-14
    g = torch_cuda.Generator().manual_seed(0)
    for i, ((n, d, ell), pp) in enumerate(zip(specs, pipes)):
        a, e = B.offsets[i], B.offsets[i + 1]
        rows = torch_cuda.randperm(n, generator=g)[:256].to(torch_cuda.int32).cuda()
        xi = X[:, a:e].contiguous()
        yd = torch_cuda.empty(k, 256, dtype=torch_cuda.float64, device="cuda")
        hmx.denseRowsXVector(ctx, n, d, pp.points, pp.tree.device_perm(), ell, rows, k, xi, yd)
        rel = float(torch_cuda.linalg.norm(after[:, a:e][:, rows.long()] - yd) / torch_cuda.linalg.norm(yd))
        print(f"cloud {i}: n={n} d={d} l={ell:.3f} rel err {rel:.2e}")
        assert rel < 1e-5


def written_boxes(n, b, L, div):
    """Box slots a tree build writes: every node of levels 0..L-1 for the
    halving methods; the segments that exist at each level for
    POWER_OF_TWO_ON_LEFT (cut at multiples of 2^(L-l) * size)."""
    if div != hmx.POWER_OF_TWO_ON_LEFT:
        return np.arange((1 << L) - 1)
    size = 1
    while size < n:
        size <<= 1
    while size > b:
        size >>= 1
    out = []
    for lev in range(L):
        lsize = size << (L - lev)
        out.extend(range((1 << lev) - 1, (1 << lev) - 1 + (n + lsize - 1) // lsize))
    return np.array(out, dtype=int)


@pytest.mark.parametrize("div", [hmx.FULL_TREE, hmx.POWER_OF_TWO_ON_LEFT])
def test_batched_datasets_and_trees_equal_single(torch_cuda, ctx, div):
    torch = torch_cuda
    b = 32
    specs = [(5000, 2), (1, 3), (31, 2), (100, 1), (70000, 3), (4097, 3), (2048, 2), (33, 1), (9000, 2)]
    ns = [n for n, _ in specs]
    ds = [d for _, d in specs]
    seeds = [500 + 7 * i for i in range(len(specs))]
    pts = [torch.empty(n * d, dtype=torch.float64, device="cuda") for n, d in specs]
    hmx.generateDatasetBatch(ctx, ns, ds, seeds, pts)
    trees = hmx.constructKDTreeBatch(ctx, ns, ds, b, pts, divMethod=div)
    for i, (n, d) in enumerate(specs):
        p1 = torch.empty(n * d, dtype=torch.float64, device="cuda")
        hmx.generateDataset(ctx, n, d, seeds[i], p1)
        assert torch.equal(pts[i], p1), i
        t1 = hmx.constructKDTree(ctx, n, d, b, p1, divMethod=div)
        assert trees[i].info() == t1.info(), i
        pb, ob, bminb, bmaxb = trees[i].copy(ctx, n, d)
        ps, os_, bmins, bmaxs = t1.copy(ctx, n, d)
        assert np.array_equal(pb, ps) and np.array_equal(ob, os_), i
        w = written_boxes(n, b, t1.info()[2], div)
        assert np.array_equal(bminb[w], bmins[w]) and np.array_equal(bmaxb[w], bmaxs[w]), i
    # the trees outlive one another: free some, the rest stay valid
    keep = trees[4]
    del trees
    pb, ob, _, _ = keep.copy(ctx, ns[4], ds[4])
    assert np.array_equal(np.sort(pb), np.arange(ns[4]))


def test_batch_rejects_mixed_bucket_sizes(torch_cuda, ctx):
    a = hmx.Pipeline(2048, 2, 32, 1e-5, ctx=ctx, direct=True)
    b = hmx.Pipeline(4096, 3, 64, 1e-5, ctx=ctx, direct=True)
    with pytest.raises(hmx.HmxError) as ei:
        hmx.Batch(ctx, [a.H, b.H])
    assert ei.value.status == hmx.HMX_EINVAL


def test_batched_build_matches_instances(torch_cuda, ctx):
    """Upper levels of several clouds compressed together equal, tile for
    tile, each cloud's own direct build: same diagonal and leaf level bit for
    bit, the same ranks on every upper-level tile, products within the
    level tolerance; and the batch's matvec against dense kernel rows."""
    specs = cloud_specs(6, 21, nmin=2048, nmax=8192)
    pts = [to_dev(torch_cuda, O.generate_points(n, d, 300 + i)) for i, (n, d, ell) in enumerate(specs)]
    bp = hmx.BatchPipeline(specs, b=32, tol=1e-6, ctx=ctx, points=pts)
    print("batched build timings", bp.timings)
    diffs = []
    for i, ((n, d, ell), H) in enumerate(zip(specs, bp.H)):
        single = hmx.Pipeline(n, d, 32, 1e-6, ell=ell, max_rank_policy=hmx.MAXRANK_FULL, points=pts[i], ctx=ctx,
                              direct=True)
        S = single.tree.info()[0]
        assert np.array_equal(H.diag(ctx, S), single.H.diag(ctx, S))
        L = single.structure.num_levels()
        for l in range(1, L):
            rb, Ub, Vb = H.level(ctx, l)
            rs, Us, Vs = single.H.level(ctx, l)
            assert H.level_info(l)["tol"] == single.H.level_info(l)["tol"]
            if l == L - 1:
                assert np.array_equal(rb, rs)
                continue
            diffs.extend(np.abs(rb.astype(int) - rs.astype(int)).tolist())
            tol = H.level_info(l)["tol"]
            for t in range(0, len(rb), max(1, len(rb) // 4)):
                if rb[t] == rs[t]:
                    assert np.linalg.norm(Ub[t] @ Vb[t].T - Us[t] @ Vs[t].T, 2) <= 2 * tol
    diffs = np.array(diffs)
    print(f"batched vs single upper-level ranks: {len(diffs)} tiles, mismatches {(diffs != 0).sum()}")
    assert diffs.max() == 0
    tot = bp.batch.info()["totalPoints"]
    X = torch_cuda.randn(8, tot, dtype=torch_cuda.float64, device="cuda")
    Y = bp.matvec(X, graph=True)
    g = torch_cuda.Generator().manual_seed(3)
    for i, (n, d, ell) in enumerate(specs):
        a, e = bp.batch.offsets[i], bp.batch.offsets[i + 1]
        rows = torch_cuda.randperm(n, generator=g)[:256].to(torch_cuda.int32).cuda()
        yd = torch_cuda.empty(8, 256, dtype=torch_cuda.float64, device="cuda")
        hmx.denseRowsXVector(ctx, n, d, bp.points[i], bp.trees[i].device_perm(), ell, rows, 8,
                             X[:, a:e].contiguous(), yd)
        rel = float(torch_cuda.linalg.norm(Y[:, a:e][:, rows.long()] - yd) / torch_cuda.linalg.norm(yd))
        assert rel < 1e-5, (i, rel)


def small_batch(torch_cuda, ctx, layout, seed=23):
    specs = cloud_specs(4, seed, 1500, 6000)
    pipes = []
    for i, (n, d, ell) in enumerate(specs):
        p = O.generate_points(n, d, 300 + i)
        pp = hmx.Pipeline(n, d, 32, 1e-6, adm="weak", ell=ell, max_rank_policy=hmx.MAXRANK_FULL,
                          points=to_dev(torch_cuda, p), ctx=ctx, direct=True)
        pp.H.set_fp64_layout(layout)
        pipes.append(pp)
    return specs, pipes, hmx.Batch(ctx, [pp.H for pp in pipes])


@pytest.mark.parametrize("k", [5, 64])
def test_batch_fp64_layouts_agree(torch_cuda, ctx, k):
    """Factors read in place (one item run and one entry per tile) and copied
    into the run layout give the same product up to summation order; within
    one layout the batch equals each instance bit for bit."""
    out = {}
    for layout in (1, 2):
        specs, pipes, B = small_batch(torch_cuda, ctx, layout)
        tot = B.info()["totalPoints"]
        X = torch_cuda.randn(k, tot, dtype=torch_cuda.float64, device="cuda",
                             generator=torch_cuda.Generator("cuda").manual_seed(4))
        Y = B.matvec(X)
        for i, pp in enumerate(pipes):
            a, e = B.offsets[i], B.offsets[i + 1]
            assert torch_cuda.equal(Y[:, a:e], pp.matvec(X[:, a:e].contiguous())), (layout, i)
        out[layout] = Y
    rel = float(torch_cuda.linalg.norm(out[1] - out[2]) / torch_cuda.linalg.norm(out[1]))
    print(f"k={k}: in place vs copy rel diff {rel:.2e}")
    assert rel < 1e-13


def test_batch_plan_follows_factor_updates(torch_cuda, ctx):
    """Overwriting an instance's level (hmx_hmatrix_set_level) after the
    batch's copied FP64 plan was built: the batch, eager and replayed, applies
    the new factors."""
    specs, pipes, B = small_batch(torch_cuda, ctx, 2, seed=29)
    tot = B.info()["totalPoints"]
    k = 8
    X = torch_cuda.randn(k, tot, dtype=torch_cuda.float64, device="cuda")
    Y = torch_cuda.empty_like(X)
    B.matvec(X, Y, graph=True)
    before = Y.clone()
    H = pipes[0].H
    lvl = next(l for l in range(1, H.num_levels) if H.level_info(l)["rankSum"] > 0)
    m = (1 << (H.num_levels - 1 - lvl)) * H.b
    rs = H.level_info(lvl)["rankSum"]
    u = torch_cuda.empty(rs * m, dtype=torch_cuda.float64, device="cuda")
    v = torch_cuda.empty_like(u)
    L = hmx.lib()
    assert L.hmx_hmatrix_level_get(ctx.handle, H.handle, lvl, u.data_ptr(), v.data_ptr()) == hmx.HMX_OK
    u.mul_(2.0)
    assert L.hmx_hmatrix_set_level(ctx.handle, H.handle, lvl, u.data_ptr(), v.data_ptr()) == hmx.HMX_OK
    B.matvec(X, Y, graph=True)
    torch_cuda.cuda.synchronize()
    a, e = B.offsets[0], B.offsets[1]
    y0 = pipes[0].matvec(X[:, a:e].contiguous())
    assert not torch_cuda.equal(Y[:, a:e], before[:, a:e])
    assert torch_cuda.equal(Y[:, a:e], y0)
    assert torch_cuda.equal(B.matvec(X), Y)
    assert torch_cuda.equal(Y[:, e:], before[:, e:])


def test_batch_graph_recaptured_after_eager_changes(torch_cuda, ctx):
    """A replayed graph never runs on plans or workspaces an eager call
    rebuilt or reallocated since the capture: (1) capture at 8 right-hand
    sides, overwrite a level, eager call (rebuilds the batch plan), replay;
    (2) capture at 16, eager call at 64 (grows the workspaces), replay at 16.
    Every replay equals the eager product of the same inputs bit for bit."""
    specs, pipes, B = small_batch(torch_cuda, ctx, 2, seed=31)
    tot = B.info()["totalPoints"]
    g = torch_cuda.Generator("cuda").manual_seed(9)
    X8 = torch_cuda.randn(8, tot, dtype=torch_cuda.float64, device="cuda", generator=g)
    Y8 = torch_cuda.empty_like(X8)
    B.matvec(X8, Y8, graph=True)
    H = pipes[1].H
    lvl = next(l for l in range(1, H.num_levels) if H.level_info(l)["rankSum"] > 0)
    m = (1 << (H.num_levels - 1 - lvl)) * H.b
    rs = H.level_info(lvl)["rankSum"]
    u = torch_cuda.empty(rs * m, dtype=torch_cuda.float64, device="cuda")
    v = torch_cuda.empty_like(u)
    L = hmx.lib()
    assert L.hmx_hmatrix_level_get(ctx.handle, H.handle, lvl, u.data_ptr(), v.data_ptr()) == hmx.HMX_OK
    v.mul_(0.5)
    assert L.hmx_hmatrix_set_level(ctx.handle, H.handle, lvl, u.data_ptr(), v.data_ptr()) == hmx.HMX_OK
    eager = B.matvec(X8)
    for rep in range(3):
        Y8.fill_(float("nan"))
        B.matvec(X8, Y8, graph=True)
        torch_cuda.cuda.synchronize()
        assert torch_cuda.equal(Y8, eager), rep
    X16 = torch_cuda.randn(16, tot, dtype=torch_cuda.float64, device="cuda", generator=g)
    Y16 = torch_cuda.empty_like(X16)
    B.matvec(X16, Y16, graph=True)
    ref16 = B.matvec(X16)
    X64 = torch_cuda.randn(64, tot, dtype=torch_cuda.float64, device="cuda", generator=g)
    B.matvec(X64)
    for rep in range(3):
        Y16.fill_(float("nan"))
        B.matvec(X16, Y16, graph=True)
        torch_cuda.cuda.synchronize()
        assert torch_cuda.equal(Y16, ref16), rep
        B.matvec(X64)  # eager work of another width between replays
    assert torch_cuda.equal(B.matvec(X16), ref16)
