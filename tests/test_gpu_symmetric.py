"""Direct assembly with symmetricTiles = 1: only the tiles (u, v) with u < v
(column cluster before row cluster) are compressed; tile (v, u) is their
transpose and takes the same rank with U and V exchanged.

Bars:
  * every mirrored tile equals its partner's factors exchanged, bit for bit,
    and has the same rank, on every level including the leaf level;
  * the diagonal blocks, and every compressed tile of the leaf level, equal
    the plain direct build's bit for bit (same kernel, same random stream);
  * every compressed tile's rank equals the oracle's ARA on the dense tile
    with the tile's own random stream (the plain build's bar), every tile --
    compressed or mirrored -- meets the ARA bound on its dense kernel tile
    (spectral error <= the level tolerance, rank between the singular-value
    counts above tol and tol/1000), and the orthonormal factor is U for a
    compressed tile, V for a mirrored one;
  * the expansion error is within 1.5x of the plain direct build's, and the
    sampling work (kernel evaluations) is at most 0.6 of it;
  * FP64 matvecs equal the expansion of the mirrored factors within 1e-13 and
    the mixed product stays within its target of the FP64 product;
  * a batched build with the option equals each cloud's own build: the same
    ranks on every tile, mirrored tiles exactly exchanged.
"""
import numpy as np
import pytest

import hmx
import oracle as O
from test_gpu_batch import cloud_specs
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


CASES = [("weak", 2, 32, 4096, 1e-6), ("maxDiam", 3, 64, 4096, 1e-5), ("boxCenter", 2, 32, 4096, 1e-6)]


@pytest.fixture(scope="module", params=CASES, ids=lambda c: f"{c[0]}-d{c[1]}-b{c[2]}-n{c[3]}")
def pair(request, torch_cuda, ctx):
    adm, d, b, n, tol = request.param
    print(f"[symmetric] building {adm} d={d} b={b} n={n} tol={tol}", flush=True)
    p = O.generate_points(n, d, 1234)
    pts = to_dev(torch_cuda, p)
    plain = hmx.Pipeline(n, d, b, tol, adm=adm, max_rank_policy=hmx.MAXRANK_FULL, points=pts, ctx=ctx, direct=True)
    sym = hmx.Pipeline(n, d, b, tol, adm=adm, max_rank_policy=hmx.MAXRANK_FULL, points=pts, ctx=ctx, direct=True,
                       symmetric=True)
    K = O.dense_matrix(p, n, d, O.kdtree(p, n, d, b)[0], sym.ell)
    return plain, sym, p, K


def partners(tiles):
    """index of tile (v, u) for every tile (u, v), and whether (u, v) is mirrored (u > v)"""
    xy = [O._mo_to_xy(int(c)) for c in tiles]
    pos = {uv: i for i, uv in enumerate(xy)}
    part = np.array([pos[(v, u)] for (u, v) in xy])
    mirrored = np.array([u > v for (u, v) in xy])
    return part, mirrored


def test_mirrored_tiles_are_exchanged_partners(pair, ctx):
    plain, sym, p, K = pair
    L = sym.structure.num_levels()
    mirrored_total = 0
    for l in range(1, L):  # the leaf level too
        tiles = sym.structure.tiles(l)
        if len(tiles) == 0:
            continue
        ranks, Us, Vs = sym.H.level(ctx, l)
        part, mirrored = partners(tiles)
        for t in np.nonzero(mirrored)[0]:
            q = part[t]
            assert not mirrored[q]
            assert ranks[t] == ranks[q], (l, t)
            assert np.array_equal(Us[t], Vs[q]) and np.array_equal(Vs[t], Us[q]), (l, t)
        mirrored_total += int(mirrored.sum())
        assert sym.H.level_info(l)["rankSum"] == int(ranks.sum())
    assert mirrored_total > 0


def test_diag_and_compressed_leaf_tiles_unchanged(pair, ctx):
    """The diagonal blocks, and the compressed (u < v) leaf tiles, equal the plain
    build's bit for bit (same kernel, same random stream); their partners are
    checked as mirrors above."""
    plain, sym, p, K = pair
    S = sym.tree.info()[0]
    assert np.array_equal(plain.H.diag(ctx, S), sym.H.diag(ctx, S))
    L = sym.structure.num_levels()
    rp, Up, Vp = plain.H.level(ctx, L - 1)
    rs, Us, Vs = sym.H.level(ctx, L - 1)
    _, mirrored = partners(sym.structure.tiles(L - 1))
    for t in np.nonzero(~mirrored)[0]:
        assert rp[t] == rs[t], t
        assert np.array_equal(Up[t], Us[t]) and np.array_equal(Vp[t], Vs[t]), t


def test_compressed_ranks_vs_oracle_ara_and_bound(pair, ctx):
    plain, sym, p, K = pair
    n, b = sym.n, sym.b
    L = sym.structure.num_levels()
    _, off, _, _ = sym.tree.copy(ctx, n, sym.d)
    rng = np.random.default_rng(7)
    diffs, checked = [], 0
    for l in range(1, L - 1):
        info = sym.H.level_info(l)
        if info["numTiles"] == 0:
            continue
        ranks, Us, Vs = sym.H.level(ctx, l)
        tiles = sym.structure.tiles(l)
        part, mirrored = partners(tiles)
        bus = 1 << (L - 1 - l)
        m = bus * b
        tol = info["tol"]
        for t in rng.choice(len(tiles), min(6, len(tiles)), replace=False):
            u, v = O._mo_to_xy(int(tiles[t]))
            rows = O._cluster_rows(off, v, bus, b)
            cols = O._cluster_rows(off, u, bus, b)
            A = K[np.ix_(rows[1], cols[1])]
            blk = (Us[t] @ Vs[t].T)[np.ix_(rows[0], cols[0])]
            err = np.linalg.norm(A - blk, 2)
            sv = np.linalg.svd(A, compute_uv=False)
            r = int(ranks[t])
            assert err <= tol, (l, t, err, tol)
            assert int((sv > tol).sum()) <= r <= int((sv > tol / 1000).sum()), (l, t, r)
            if r:
                orth = Vs[t] if mirrored[t] else Us[t]
                assert np.abs(orth.T @ orth - np.eye(r)).max() < 1e-11
            # the compressed tile of the pair against the oracle's ARA with its own stream
            q = part[t] if mirrored[t] else t
            uq, vq = O._mo_to_xy(int(tiles[q]))
            Aq = K[vq * m:(vq + 1) * m, uq * m:(uq + 1) * m]  # full leaves: padded space == KD order
            Uo, Vo, k, _ = O.ara_dense(Aq, tol, bs=16, r=10, max_rank=info["maxRank"], seed=1234,
                                       opid=((l + 1) << 40) | int(q))
            diffs.append(abs(int(ranks[q]) - k))
            checked += 1
    diffs = np.array(diffs)
    print(f"symmetric build vs oracle ARA: {len(diffs)} tiles, rank mismatches {(diffs != 0).sum()}")
    assert checked > 0 and diffs.max() == 0


def test_expansion_error_and_work(pair, ctx):
    plain, sym, p, K = pair
    ep = np.linalg.norm(expand_gpu(ctx, plain) - K) / np.linalg.norm(K)
    es = np.linalg.norm(expand_gpu(ctx, sym) - K) / np.linalg.norm(K)
    wp, ws = plain.H.build_info(), sym.H.build_info()
    print(f"expansion error: plain {ep:.3e}, symmetric {es:.3e}; kernel evaluations {ws['kernelEvals']:.3e} "
          f"vs {wp['kernelEvals']:.3e}")
    assert es <= 1.5 * ep + 1e-9
    assert ws["kernelEvals"] <= 0.6 * wp["kernelEvals"]


@pytest.mark.parametrize("k", [1, 16, 64])
def test_matvecs_on_mirrored_factors(pair, ctx, torch_cuda, k):
    plain, sym, p, K = pair
    n = sym.n
    X = O.generate_rhs(n, k, 41 + k)
    x = to_dev(torch_cuda, X.T)
    E = expand_gpu(ctx, sym)
    y = sym.matvec(x)
    Y = y.cpu().numpy().T
    assert np.linalg.norm(Y - E @ X) / np.linalg.norm(E @ X) < 1e-13
    yt = sym.matvec(x, transpose=True).cpu().numpy().T
    assert np.linalg.norm(yt - E.T @ X) / np.linalg.norm(E.T @ X) < 1e-13
    if k >= 16:
        target = 1e-5
        sym.make_mixed(target)
        x32 = x.float()
        y32 = torch_cuda.empty_like(x32)
        sym.matvec_mixed(x32, y32)
        rel = float(torch_cuda.linalg.norm(y32.double() - y) / torch_cuda.linalg.norm(y))
        print(f"mixed vs FP64 on the symmetric build: {rel:.3e}")
        assert rel <= target


def test_batched_symmetric_build_matches_instances(torch_cuda, ctx):
    specs = cloud_specs(5, 31, nmin=2048, nmax=8192)
    pts = [to_dev(torch_cuda, O.generate_points(n, d, 700 + i)) for i, (n, d, ell) in enumerate(specs)]
    bp = hmx.BatchPipeline(specs, b=32, tol=1e-6, ctx=ctx, points=pts, symmetric=True)
    for i, ((n, d, ell), H) in enumerate(zip(specs, bp.H)):
        single = hmx.Pipeline(n, d, 32, 1e-6, ell=ell, max_rank_policy=hmx.MAXRANK_FULL, points=pts[i], ctx=ctx,
                              direct=True, symmetric=True)
        S = single.tree.info()[0]
        assert np.array_equal(H.diag(ctx, S), single.H.diag(ctx, S))
        L = single.structure.num_levels()
        for l in range(1, L):
            rb, Ub, Vb = H.level(ctx, l)
            rs, _, _ = single.H.level(ctx, l)
            assert np.array_equal(rb, rs), (i, l)
            if l == L - 1 or len(rb) == 0:
                continue
            part, mirrored = partners(single.structure.tiles(l))
            for t in np.nonzero(mirrored)[0]:
                assert np.array_equal(Ub[t], Vb[part[t]]) and np.array_equal(Vb[t], Ub[part[t]]), (i, l, t)
