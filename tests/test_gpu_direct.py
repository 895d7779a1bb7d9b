"""Direct H-matrix assembly (no S x S TLR matrix) and 64-point leaves.

Bars:
  * the diagonal blocks and the leaf level of the direct build equal, bit for
    bit, those of the TLR path (same tile kernel, same random stream ids);
  * every upper-level tile of the direct build meets the ARA guarantee on the
    dense kernel tile: spectral error <= the level tolerance, and its rank lies
    between the exact count of singular values above the tolerance and the
    count above tolerance/1000 (numpy SVD);
  * the direct build's expansion error against the dense kernel matrix is
    within 1.5x of the TLR path's;
  * with b = 64 the TLR path's ranks equal the CPU oracle's and the direct
    build's upper-level ranks equal the oracle's ARA on the dense tile with
    the same random stream; both matvec paths (streaming FP64, matrix-core
    FP64, mixed) equal the dense expansion within their bars.
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


CASES = [("weak", 2, 32, 4096, 1e-6), ("maxDiam", 3, 64, 4096, 1e-5), ("weak", 3, 64, 4096, 1e-5)]


@pytest.fixture(scope="module", params=CASES, ids=lambda c: f"{c[0]}-d{c[1]}-b{c[2]}")
def trio(request, torch_cuda, ctx):
    adm, d, b, n, tol = request.param
    print(f"[direct] building {adm} d={d} b={b} n={n} tol={tol}", flush=True)
    p = O.generate_points(n, d, 1234)
    pts = to_dev(torch_cuda, p)
    tlr = hmx.Pipeline(n, d, b, tol, adm=adm, max_rank_policy=hmx.MAXRANK_FULL, points=pts, ctx=ctx)
    direct = hmx.Pipeline(n, d, b, tol, adm=adm, max_rank_policy=hmx.MAXRANK_FULL, points=pts, ctx=ctx, direct=True)
    K = O.dense_matrix(p, n, d, O.kdtree(p, n, d, b)[0], tlr.ell)
    return tlr, direct, p, K


def test_diag_and_leaf_level_bitwise(trio, ctx):
    tlr, direct, p, K = trio
    S = tlr.tree.info()[0]
    assert np.array_equal(tlr.H.diag(ctx, S), direct.H.diag(ctx, S))
    L = tlr.structure.num_levels()
    rt, Ut, Vt = tlr.H.level(ctx, L - 1)
    rd, Ud, Vd = direct.H.level(ctx, L - 1)
    assert np.array_equal(rt, rd)
    for a, c in zip(Ut + Vt, Ud + Vd):
        assert np.array_equal(a, c)
    assert tlr.H.level_info(L - 1)["maxRank"] == direct.H.level_info(L - 1)["maxRank"]


def test_upper_levels_meet_ara_bound(trio, ctx):
    tlr, direct, p, K = trio
    n, b = tlr.n, tlr.b
    L = tlr.structure.num_levels()
    _, off, _, _ = direct.tree.copy(ctx, n, direct.d)
    rng = np.random.default_rng(3)
    checked = 0
    for l in range(1, L - 1):
        info = direct.H.level_info(l)
        if info["numTiles"] == 0:
            continue
        assert info["tol"] == tlr.H.level_info(l)["tol"]
        ranks, Us, Vs = direct.H.level(ctx, l)
        rt = tlr.H.level(ctx, l)[0]
        print(f"level {l}: direct ranks mean {ranks.mean():.1f} (TLR path {rt.mean():.1f}), tol {info['tol']:.3g}")
        bus = 1 << (L - 1 - l)
        tiles = direct.structure.tiles(l)
        for t in rng.choice(len(tiles), min(3, len(tiles)), replace=False):
            u, v = O._mo_to_xy(int(tiles[t]))
            rows = O._cluster_rows(off, v, bus, b)
            cols = O._cluster_rows(off, u, bus, b)
            A = K[np.ix_(rows[1], cols[1])]
            blk = (Us[t] @ Vs[t].T)[np.ix_(rows[0], cols[0])]
            err = np.linalg.norm(A - blk, 2)
            sv = np.linalg.svd(A, compute_uv=False)
            tol = info["tol"]
            r = int(ranks[t])
            assert err <= tol, (l, t, err, tol)
            assert int((sv > tol).sum()) <= r <= int((sv > tol / 1000).sum()), (l, t, r)
            if r:
                assert np.abs(Us[t].T @ Us[t] - np.eye(r)).max() < 1e-11
            checked += 1
    assert checked > 0


def test_direct_expansion_error(trio, ctx):
    tlr, direct, p, K = trio
    et = np.linalg.norm(expand_gpu(ctx, tlr) - K) / np.linalg.norm(K)
    ed = np.linalg.norm(expand_gpu(ctx, direct) - K) / np.linalg.norm(K)
    print(f"expansion error: TLR path {et:.3e}, direct {ed:.3e}")
    assert ed <= 1.5 * et + 1e-9


@pytest.mark.parametrize("k", [1, 16, 64])
def test_matvec_paths_vs_expansion(trio, ctx, torch_cuda, k):
    tlr, direct, p, K = trio
    n = tlr.n
    X = O.generate_rhs(n, k, 11 + k)
    x = to_dev(torch_cuda, X.T)
    for pipe in (tlr, direct):
        E = expand_gpu(ctx, pipe)
        Y = pipe.matvec(x).cpu().numpy().T
        assert np.linalg.norm(Y - E @ X) / np.linalg.norm(E @ X) < 1e-13
        if k == 1:
            yt = pipe.matvec(x, transpose=True).cpu().numpy().T
            assert np.linalg.norm(yt - E.T @ X) / np.linalg.norm(E.T @ X) < 1e-13
            continue
        # every matrix-core plan variant gives the same product
        for opts in ((1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 1)):
            pipe.H.set_plan_options(*opts)
            Y = pipe.matvec(x).cpu().numpy().T
            assert np.linalg.norm(Y - E @ X) / np.linalg.norm(E @ X) < 1e-13, opts
            if opts[0] and tlr.d == 3:  # 3D leaf-level tiles have ranks above b/2
                assert pipe.H.plan_info(0)["denseBlocks"] > 0
        pipe.H.set_plan_options(0, 0, 0)
        # FP64 factors read in place and copied into the run layout
        for layout in (1, 2):
            pipe.H.set_fp64_layout(layout)
            Y = pipe.matvec(x).cpu().numpy().T
            assert np.linalg.norm(Y - E @ X) / np.linalg.norm(E @ X) < 1e-13, layout
        pipe.H.set_fp64_layout(0)


def test_b64_ranks_vs_oracle(torch_cuda, ctx):
    n, d, b, tol = 4096, 3, 64, 1e-5
    p = O.generate_points(n, d, 1234)
    pipe = hmx.Pipeline(n, d, b, tol, adm="maxDiam", max_rank_policy=hmx.MAXRANK_FULL, points=to_dev(torch_cuda, p),
                        ctx=ctx)
    ref = O.OracleHMatrix(p, n, d, b, adm="maxDiam", tol=tol, max_rank_policy=1, seed=1234)
    for l in range(1, ref.num_levels):
        ro = ref.level(l)[0]
        rg = pipe.H.level(ctx, l)[0]
        assert len(rg) == len(ro)
        if len(ro):
            dd = np.abs(rg.astype(int) - ro.astype(int))
            print(f"level {l}: {len(ro)} tiles, rank mismatches {(dd != 0).sum()}, max {dd.max()}")
            assert dd.max() == 0, l


def test_b64_mixed(trio, ctx, torch_cuda):
    tlr, direct, p, K = trio
    n = tlr.n
    target = 1e-5
    direct.H.set_plan_options(1, 1, 1)
    info = direct.make_mixed(target)
    X = O.generate_rhs(n, 64, 8).astype(np.float32)
    x32 = to_dev(torch_cuda, X.T)
    y32 = direct.matvec_mixed(x32).cpu().numpy().T.astype(np.float64)
    y64 = direct.matvec(x32.double()).cpu().numpy().T
    rel = np.linalg.norm(y32 - y64) / np.linalg.norm(y64)
    print(f"mixed b={tlr.b} (all plan options): hi {info['hiCols']} lo {info['loCols']} rel vs FP64 {rel:.3e}")
    assert rel <= target
    direct.H.set_plan_options(0, 0, 0)
    info = direct.make_mixed(target)
    y32 = direct.matvec_mixed(x32).cpu().numpy().T.astype(np.float64)
    rel = np.linalg.norm(y32 - y64) / np.linalg.norm(y64)
    print(f"mixed b={tlr.b} (default plan): rel vs FP64 {rel:.3e}")
    assert rel <= target


def test_direct_ranks_vs_oracle_ara(trio, ctx):
    """Upper-level tiles of the direct build against the oracle's ARA run on the
    dense kernel tile with the same random stream (seed, operation id
    ((level+1) << 40) | tile, column index): the same ranks on every sampled
    tile, products within the level tolerance."""
    tlr, direct, p, K = trio
    n, b = tlr.n, tlr.b
    L = tlr.structure.num_levels()
    _, off, _, _ = direct.tree.copy(ctx, n, direct.d)
    rng = np.random.default_rng(5)
    diffs = []
    for l in range(1, L - 1):
        info = direct.H.level_info(l)
        if info["numTiles"] == 0:
            continue
        ranks, Us, Vs = direct.H.level(ctx, l)
        bus = 1 << (L - 1 - l)
        m = bus * b
        tiles = direct.structure.tiles(l)
        for t in rng.choice(len(tiles), min(6, len(tiles)), replace=False):
            u, v = O._mo_to_xy(int(tiles[t]))
            A = K[v * m:(v + 1) * m, u * m:(u + 1) * m]  # full leaves: padded space == KD order
            Uo, Vo, k, _ = O.ara_dense(A, info["tol"], bs=16, r=10, max_rank=info["maxRank"], seed=1234,
                                       opid=((l + 1) << 40) | int(t))
            diffs.append(abs(int(ranks[t]) - k))
            if ranks[t] == k:
                assert np.linalg.norm(Us[t] @ Vs[t].T - Uo @ Vo.T, 2) <= 2 * info["tol"]
    diffs = np.array(diffs)
    print(f"direct vs oracle ARA: {len(diffs)} tiles, rank mismatches {(diffs != 0).sum()}, max {diffs.max()}")
    assert diffs.max() == 0
