"""GPU parity: every stage of the HIP pipeline, called through the C ABI, against
the pinned CPU oracle on the same inputs.

Bars: bit-exact for the point cloud, KD-tree (permutation, offsets, boxes),
structure and JDS; identical ARA ranks on every tile and level, and
factor products equal to the oracle's within the ARA tolerance; matvec equal
to the dense expansion of the GPU's own factors within 1e-13 (FP64) and to the
dense kernel product within 10*eps.
"""
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
def ctx(torch_cuda):
    return hmx.Context()


def to_dev(torch, a, dtype=None):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def expand_gpu(ctx, pipe):
    """Dense expansion (KD order) of the GPU H-matrix from its own factors."""
    H, n, b = pipe.H, pipe.n, pipe.b
    S = pipe.tree.info()[0]
    _, off, _, _ = pipe.tree.copy(ctx, n, pipe.d)
    D = H.diag(ctx, S).reshape(S, b, b)
    E = np.zeros((n, n))
    for leaf in range(S):
        a, e = off[leaf], off[leaf + 1]
        E[a:e, a:e] = D[leaf].T[:e - a, :e - a]
    L = pipe.structure.num_levels()
    for l in range(1, L):
        ranks, Us, Vs = H.level(ctx, l)
        bus = 1 << (L - 1 - l)
        for t, ti in enumerate(pipe.structure.tiles(l)):
            u, v = O._mo_to_xy(int(ti))
            rows = O._cluster_rows(off, v, bus, b)
            cols = O._cluster_rows(off, u, bus, b)
            blk = Us[t] @ Vs[t].T
            E[np.ix_(rows[1], cols[1])] += blk[np.ix_(rows[0], cols[0])]
    return E


# ----------------------------------------------------------------- dataset --
@pytest.mark.parametrize("n,d", [(1, 2), (1000, 2), (4097, 3), (65536, 2)])
def test_points_bitwise(torch_cuda, ctx, n, d):
    pts = torch_cuda.empty(n * d, dtype=torch_cuda.float64, device="cuda")
    hmx.generateDataset(ctx, n, d, 1234, pts)
    assert np.array_equal(pts.cpu().numpy(), O.generate_points(n, d, 1234))


def test_random_vector(torch_cuda, ctx):
    n, k = 5000, 3
    x = torch_cuda.empty(k * n, dtype=torch_cuda.float64, device="cuda")
    hmx.generateRandomVector(ctx, k, n, 99, x)
    ref = O.generate_rhs(n, k, 99).T.reshape(-1)
    assert np.allclose(x.cpu().numpy(), ref, rtol=1e-14, atol=1e-14)


# ----------------------------------------------------------------- KD-tree --
@pytest.mark.parametrize("n,d,b,quant", [(13, 2, 2, 0), (100, 2, 8, 4), (1000, 3, 16, 0), (4096, 2, 32, 0),
                                         (5000, 3, 32, 32), (65536, 2, 32, 0), (65536, 3, 32, 64),
                                         (100000, 3, 64, 0)])
def test_kdtree_bitwise(torch_cuda, ctx, n, d, b, quant):
    p = O.generate_points(n, d, 7 + n)
    if quant:
        p = np.floor(p * quant) / quant  # ties
    pts = to_dev(torch_cuda, p)
    tree = hmx.constructKDTree(ctx, n, d, b, pts)
    perm, off, bmin, bmax = tree.copy(ctx, n, d)
    rperm, roff, rmin, rmax, _ = O.kdtree(p, n, d, b)
    assert np.array_equal(perm, rperm)
    assert np.array_equal(off, roff)
    L = O.kd_num_levels(n, b)
    inner = (1 << L) - 1  # leaf-level boxes are never filled (as in the reference)
    assert np.array_equal(bmin[:inner], rmin[:inner])
    assert np.array_equal(bmax[:inner], rmax[:inner])


@pytest.mark.parametrize("adm,d", [("weak", 2), ("boxCenter", 2), ("boxCenter", 3)])
def test_structure_from_gpu_tree(torch_cuda, ctx, adm, d):
    n, b = 8192, 32
    p = O.generate_points(n, d, 3)
    tree = hmx.constructKDTree(ctx, n, d, b, to_dev(torch_cuda, p))
    s = hmx.constructHMatrixStructure(ctx, n, d, b, adm, 1.0, tree)
    _, _, rmin, rmax, _ = O.kdtree(p, n, d, b)
    ref = O.structure(n, d, b, adm, 1.0, rmin, rmax)
    for l in range(1, s.num_levels()):
        assert np.array_equal(s.tiles(l), ref[l - 1])


@pytest.mark.parametrize("adm,d,b,n", [("weak", 2, 32, 8192), ("boxCenter", 2, 32, 8192), ("maxDiam", 2, 32, 8192),
                                       ("boxCenter", 3, 64, 5000), ("maxDiam", 3, 32, 70000), ("maxDiam", 1, 32, 4096),
                                       ("weak", 3, 64, 64), ("maxDiam", 3, 64, 1 << 20)])
def test_structure_device_pass_equals_host_pass(torch_cuda, ctx, adm, d, b, n):
    """The level-synchronous GPU pass gives the host recursion's tile lists (same
    order) and JDS, bit for bit; at small sizes both equal the oracle's."""
    pts = torch_cuda.empty(n * d, dtype=torch_cuda.float64, device="cuda")
    hmx.generateDataset(ctx, n, d, 77 + d, pts)
    tree = hmx.constructKDTree(ctx, n, d, b, pts)
    import time
    t0 = time.perf_counter()
    sh = hmx.constructHMatrixStructure(ctx, n, d, b, adm, 1.0, tree)
    t1 = time.perf_counter()
    sd = hmx.constructHMatrixStructure(ctx, n, d, b, adm, 1.0, tree, device=True)
    t2 = time.perf_counter()
    print(f"host pass {1e3 * (t1 - t0):.1f} ms, device pass {1e3 * (t2 - t1):.1f} ms")
    L = sh.num_levels()
    assert sd.num_levels() == L and np.array_equal(sd.num_tiles(), sh.num_tiles())
    total = 0
    for l in range(1, L):
        th, td = sh.tiles(l), sd.tiles(l)
        assert np.array_equal(th, td), (l, len(th), len(td))
        for a, c in zip(sh.jds(l), sd.jds(l)):
            assert np.array_equal(a, c), l
        total += len(th)
    print(f"{adm} d={d} b={b} n={n}: {L} levels, {total} tiles")
    if 1 < n <= 8192:
        _, _, rmin, rmax, _ = O.kdtree(pts.cpu().numpy(), n, d, b)
        ref = O.structure(n, d, b, adm, 1.0, rmin, rmax)
        for l in range(1, L):
            assert np.array_equal(sd.tiles(l), ref[l - 1])


# ---------------------------------------------------------- TLR + H build --
@pytest.fixture(scope="module", params=[("weak", 1), ("weak", 0), ("boxCenter", 1)])
def pair(request, torch_cuda, ctx):
    adm, policy = request.param
    n, d, b, tol = 4096, 2, 32, 1e-6
    p = O.generate_points(n, d, 1234)
    pipe = hmx.Pipeline(n, d, b, tol, adm=adm, max_rank_policy=policy, points=to_dev(torch_cuda, p), ctx=ctx)
    ref = O.OracleHMatrix(p, n, d, b, adm=adm, tol=tol, max_rank_policy=policy, seed=1234)
    return pipe, ref, p


def test_tlr_ranks_and_factors(pair, ctx):
    pipe, ref, p = pair
    S, b = ref.S, ref.b
    ranks, offs, U, V, D = pipe.tlr_cm.copy(ctx)
    rref = ref.tlr_ranks(False)
    diff = np.abs(ranks - rref)
    print(f"TLR rank mismatches: {(diff != 0).sum()} of {diff.size}, max {diff.max()}")
    assert diff.max() == 0
    assert np.allclose(D, ref.diag(), rtol=1e-14, atol=1e-15)
    assert pipe.tlr_cm.info()["capHits"] == O.lib().or_hm_get_int(ref.h, 3)
    # factor products agree within the ARA tolerance where ranks agree
    Ur = np.ctypeslib.as_array(O.lib().or_hm_tlr_U(ref.h, 0), shape=(int(rref.sum()) * b,))
    Vr = np.ctypeslib.as_array(O.lib().or_hm_tlr_V(ref.h, 0), shape=(int(rref.sum()) * b,))
    oref = np.concatenate([[0], np.cumsum(rref)])
    rng = np.random.default_rng(0)
    for t in rng.choice(S * S, 300, replace=False):
        if t // S == t % S or ranks[t] != rref[t]:
            continue
        r = ranks[t]
        Ug = U[offs[t] * b:(offs[t] + r) * b].reshape(r, b).T
        Vg = V[offs[t] * b:(offs[t] + r) * b].reshape(r, b).T
        Uo = Ur[oref[t] * b:(oref[t] + r) * b].reshape(r, b).T
        Vo = Vr[oref[t] * b:(oref[t] + r) * b].reshape(r, b).T
        assert np.linalg.norm(Ug @ Vg.T - Uo @ Vo.T, 2) <= 2e-6
        if r:
            assert np.abs(Ug.T @ Ug - np.eye(r)).max() < 1e-12


def test_morton_conversion_moves_factors_exactly(pair, ctx):
    pipe, ref, _ = pair
    S, b = ref.S, ref.b
    rc, oc, Uc, Vc, _ = pipe.tlr_cm.copy(ctx)
    rm, om, Um, Vm, _ = pipe.tlr_mo.copy(ctx)
    assert pipe.tlr_mo.info()["ordering"] == hmx.MORTON
    L = O.lib()
    for i in range(0, S * S, 7):
        mo = L.or_cm_to_mo(S, i)
        assert rm[mo] == rc[i]
        r = rc[i]
        assert np.array_equal(Um[om[mo] * b:(om[mo] + r) * b], Uc[oc[i] * b:(oc[i] + r) * b])
        assert np.array_equal(Vm[om[mo] * b:(om[mo] + r) * b], Vc[oc[i] * b:(oc[i] + r) * b])


@pytest.mark.parametrize("pieces", [1, 2, 4])
def test_tlr_pieces_tile_the_whole_matrix_exactly(pair, ctx, pieces):
    """Every piece of a p x p partition holds the whole TLR matrix's tiles bit
    for bit (ranks, U, V) and a diagonal piece its diagonal blocks."""
    pipe, ref, _ = pair
    S, b = ref.S, ref.b
    rf, of, Uf, Vf, Df = pipe.tlr_cm.copy(ctx)
    Sp = S // pieces
    rank_total = 0
    for i in range(pieces):
        for j in range(pieces):
            m, rs = hmx.createColumnMajorLRMatrixPiece(ctx, pipe.cfg, pipe.tree, pipe.points, pieces, i, j)
            info = m.info()
            assert info["numSegments"] == Sp and info["ordering"] == hmx.COLUMN_MAJOR
            rp, op, Up, Vp, Dp = m.copy(ctx)
            assert rs == info["rankSum"] == int(rp.sum())
            rank_total += rs
            for uu in range(Sp):
                for vv in range(Sp):
                    tp = uu * Sp + vv
                    t = (j * Sp + uu) * S + (i * Sp + vv)
                    r = rf[t]
                    assert rp[tp] == r, (i, j, uu, vv)
                    if r:
                        assert np.array_equal(Up[op[tp] * b:(op[tp] + r) * b], Uf[of[t] * b:(of[t] + r) * b])
                        assert np.array_equal(Vp[op[tp] * b:(op[tp] + r) * b], Vf[of[t] * b:(of[t] + r) * b])
            if i == j:
                assert np.array_equal(Dp, Df[i * Sp * b * b:(i + 1) * Sp * b * b])
            m.free()
    assert rank_total == int(rf.sum())


def test_tlr_piece_rejections(pair, ctx):
    pipe, ref, _ = pair
    for pieces, i, j in [(3, 0, 0), (2, 2, 0), (2, 0, 2), (0, 0, 0)]:
        with pytest.raises(hmx.HmxError) as ei:
            hmx.createColumnMajorLRMatrixPiece(ctx, pipe.cfg, pipe.tree, pipe.points, pieces, i, j)
        assert ei.value.status == hmx.HMX_EINVAL
    m, _ = hmx.createColumnMajorLRMatrixPiece(ctx, pipe.cfg, pipe.tree, pipe.points, 2, 0, 0)
    with pytest.raises(hmx.HmxError) as ei:
        hmx.convertColumnMajorToMorton(ctx, m)
    assert ei.value.status == hmx.HMX_EINVAL
    with pytest.raises(hmx.HmxError) as ei:
        hmx.generateHMatrixFromStruct(ctx, pipe.cfg, pipe.tree, pipe.structure, m)
    assert ei.value.status == hmx.HMX_EINVAL
    m.free()


def test_hmatrix_levels_vs_oracle(pair, ctx):
    pipe, ref, p = pair
    L = ref.num_levels
    for l in range(1, L):
        info = pipe.H.level_info(l)
        mr, tol, caps = ref.level_info(l)
        rg = pipe.H.level(ctx, l)[0]
        ro = ref.level(l)[0]
        assert info["numTiles"] == len(ro)
        if len(ro) == 0:
            continue
        assert info["maxRank"] == mr and abs(info["tol"] - tol) <= 1e-12 * max(tol, 1)
        dd = np.abs(rg.astype(int) - ro.astype(int))
        print(f"level {l}: rank mismatches {(dd != 0).sum()} of {len(dd)}, max {dd.max()}")
        assert dd.max() == 0, l
    K = O.dense_matrix(p, ref.n, ref.d, ref.perm, ref.ell)
    E = expand_gpu(ctx, pipe)
    Eo = ref.expand()
    eg = np.linalg.norm(E - K) / np.linalg.norm(K)
    eo = np.linalg.norm(Eo - K) / np.linalg.norm(K)
    # the GPU H-matrix is as accurate as the oracle's (same algorithm, same Omega)
    assert eg <= 1.5 * eo + 1e-9


@pytest.mark.parametrize("k", [1, 4, 8, 16])
def test_matvec_exact_vs_own_expansion(pair, ctx, torch_cuda, k):
    pipe, ref, p = pair
    n = ref.n
    X = O.generate_rhs(n, k, 5 + k)
    x = to_dev(torch_cuda, X.T)  # (k, n) row-major == n x k column-major
    y = pipe.matvec(x)
    Y = y.cpu().numpy().T
    E = expand_gpu(ctx, pipe)
    assert np.linalg.norm(Y - E @ X) / np.linalg.norm(E @ X) < 1e-13
    yt = pipe.matvec(x, transpose=True).cpu().numpy().T
    assert np.linalg.norm(yt - E.T @ X) / np.linalg.norm(E.T @ X) < 1e-13
    # same answer as the oracle's own H-matrix within the compression error
    Yo = ref.matvec(X)
    K = O.dense_matrix(p, n, ref.d, ref.perm, ref.ell)
    err_g = np.linalg.norm(Y - K @ X) / np.linalg.norm(K @ X)
    err_o = np.linalg.norm(Yo - K @ X) / np.linalg.norm(K @ X)
    assert err_g <= 2 * err_o + 1e-9


def test_matvec_deterministic(pair, torch_cuda):
    pipe, ref, _ = pair
    x = torch_cuda.randn(4, ref.n, dtype=torch_cuda.float64, device="cuda")
    y1 = pipe.matvec(x).clone()
    y2 = pipe.matvec(x)
    assert torch_cuda.equal(y1, y2)


def test_matvec_in_original_order(pair, ctx, torch_cuda):
    """Vectors in the caller's point order: gathered into the KD order through
    the tree's permutation and scattered back; equal, bit for bit, to the
    KD-order product with the permutation applied by hand (numpy)."""
    pipe, ref, _ = pair
    n = ref.n
    perm, _, _, _ = pipe.tree.copy(ctx, n, ref.d)
    x = torch_cuda.randn(4, n, dtype=torch_cuda.float64, device="cuda")
    xk = pipe.tree.reorder(ctx, x, False)
    assert np.array_equal(xk.cpu().numpy(), x.cpu().numpy()[:, perm])
    assert torch_cuda.equal(pipe.tree.reorder(ctx, xk, True), x)
    y = pipe.matvec(x, original_order=True)
    yk = pipe.matvec(torch_cuda.from_numpy(x.cpu().numpy()[:, perm].copy()).cuda())
    expect = np.empty_like(yk.cpu().numpy())
    expect[:, perm] = yk.cpu().numpy()
    assert np.array_equal(y.cpu().numpy(), expect)


def test_dense_rows_vs_oracle(pair, ctx, torch_cuda):
    pipe, ref, p = pair
    n = ref.n
    X = O.generate_rhs(n, 3, 1)
    rows = np.array([0, 1, 77, 2048, n - 1], dtype=np.int32)
    Y = torch_cuda.zeros(3, len(rows), dtype=torch_cuda.float64, device="cuda")
    hmx.denseRowsXVector(ctx, n, ref.d, pipe.points, pipe.tree.device_perm(), ref.ell, to_dev(torch_cuda, rows), 3,
                         to_dev(torch_cuda, X.T), Y)
    Yo = O.sampled_rows(p, n, ref.d, ref.perm, ref.ell, rows, X)
    assert np.allclose(Y.cpu().numpy().T, Yo, rtol=1e-12, atol=1e-12)


def test_bad_config_rejected(torch_cuda, ctx):
    with pytest.raises(hmx.HmxError) as ei:
        hmx.Pipeline(1024, 2, 32, 1e-5, ctx=ctx, adm=5)
    assert ei.value.status == hmx.HMX_EINVAL


@pytest.mark.parametrize("n,d,b,quant", [(1000, 3, 32, 0), (5000, 2, 24, 16), (65536, 2, 32, 0), (100000, 3, 64, 0)])
def test_kdtree_division_methods(torch_cuda, ctx, n, d, b, quant):
    """POWER_OF_TWO_ON_LEFT bitwise against the oracle; DIVIDE_IN_HALF is the
    FULL_TREE construction, as in the reference."""
    p = O.generate_points(n, d, 21 + n)
    if quant:
        p = np.floor(p * quant) / quant
    pts = to_dev(torch_cuda, p)
    tree = hmx.constructKDTree(ctx, n, d, b, pts, hmx.POWER_OF_TWO_ON_LEFT)
    perm, off, bmin, bmax = tree.copy(ctx, n, d)
    rperm, roff, rmin, rmax, _, L = O.kdtree_pow2left(p, n, d, b)
    assert np.array_equal(perm, rperm) and np.array_equal(off, roff)
    size = 1 << L
    top = 1 << max(0, (n - 1).bit_length())
    for l in range(L):
        lsize = top >> l
        for s in range(-(-n // lsize)):
            q = (1 << l) - 1 + s
            assert np.array_equal(bmin[q], rmin[q]) and np.array_equal(bmax[q], rmax[q])
    half = hmx.constructKDTree(ctx, n, d, b, pts, hmx.DIVIDE_IN_HALF)
    full = hmx.constructKDTree(ctx, n, d, b, pts, hmx.FULL_TREE)
    hp, ho, hmn, hmx_ = half.copy(ctx, n, d)
    fp, fo, fmn, fmx = full.copy(ctx, n, d)
    inner = (1 << O.kd_num_levels(n, b)) - 1  # leaf-level boxes are never filled
    assert np.array_equal(hp, fp) and np.array_equal(ho, fo)
    assert np.array_equal(hmn[:inner], fmn[:inner]) and np.array_equal(hmx_[:inner], fmx[:inner])


def test_pow2left_tree_structure_guard(torch_cuda, ctx):
    # 1000 points, leaves of 32: 32 = 2^5 leaves -> a structure can be built
    p = O.generate_points(1000, 2, 5)
    t = hmx.constructKDTree(ctx, 1000, 2, 32, to_dev(torch_cuda, p), hmx.POWER_OF_TWO_ON_LEFT)
    s = hmx.constructHMatrixStructure(ctx, 1000, 2, 32, "weak", 1.0, t)
    assert s.num_levels() == 6
    # 900 points: 29 leaves, not 2^levels -> refused
    p = O.generate_points(900, 2, 5)
    t = hmx.constructKDTree(ctx, 900, 2, 32, to_dev(torch_cuda, p), hmx.POWER_OF_TWO_ON_LEFT)
    with pytest.raises(hmx.HmxError) as ei:
        hmx.constructHMatrixStructure(ctx, 900, 2, 32, "weak", 1.0, t)
    assert ei.value.status == hmx.HMX_EUNSUPPORTED
