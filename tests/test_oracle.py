"""Pins the CPU oracle (oracle/hmx_oracle.c) before it is used as the checker.

Every check here is independent of the oracle's own code: published
known-answer vectors, brute-force numpy restatements, closed forms, numpy
SVD, and the reference's own fixture test/batchedMatrix.txt.
"""
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---------------------------------------------------------------- Philox ----
def test_philox_known_answers():
    # Random123 kat_vectors for philox4x32_10
    assert [int(x) for x in O.philox([0, 0, 0, 0], [0, 0])] == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert [int(x) for x in O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)] == [0x408F276D, 0x41C83B0E, 0xA20BC7C6,
                                                                               0x6D5451FD]
    assert [int(x) for x in O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0])] == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_points_uniform_and_deterministic():
    p = O.generate_points(20000, 3, 1234)
    assert p.min() >= 0.0 and p.max() < 1.0
    assert abs(p.mean() - 0.5) < 0.01 and abs(p.var() - 1 / 12) < 0.003
    assert np.array_equal(p, O.generate_points(20000, 3, 1234))
    assert not np.array_equal(p, O.generate_points(20000, 3, 1235))
    # element e=j*n+i uses counter e: a prefix of a larger cloud's first dim
    q = O.generate_points(40000, 1, 1234)
    assert np.array_equal(p[:20000], q[:20000])


def test_rhs_gaussian_moments():
    x = O.generate_rhs(50000, 2, 7)
    assert abs(x.mean()) < 0.02 and abs(x.std() - 1) < 0.02
    assert abs(np.mean(x ** 4) - 3) < 0.15  # kurtosis of a normal


# --------------------------------------------------------------- KD-tree ----
def kd_brute(p, n, d, b):
    """Independent numpy restatement of src/kDTreeConstruction.cu (FULL_TREE)."""
    P = p.reshape(d, n)
    perm = np.arange(n)
    off = [0, n]
    size = n
    boxes, dims = [], []
    while size > b:
        noff, lv_box, lv_dim = [], [], []
        for s in range(len(off) - 1):
            a, e = off[s], off[s + 1]
            seg = perm[a:e].copy()
            if e > a:
                mn, mx = P[:, seg].min(1), P[:, seg].max(1)
            else:
                mn, mx = np.full(d, np.finfo(float).max), np.full(d, -np.finfo(float).max)
            dim = int(np.argmax(mx - mn))
            perm[a:e] = seg[np.argsort(P[dim, seg], kind="stable")]
            noff += [a, a + (e - a + 1) // 2]
            lv_box.append((mn, mx))
            lv_dim.append(dim)
        noff.append(n)
        off = noff
        size = (size + 1) // 2
        boxes.append(lv_box)
        dims.append(lv_dim)
    return perm, np.array(off), boxes, dims


@pytest.mark.parametrize("n,d,b,quant", [(8, 2, 2, 0), (13, 2, 2, 0), (64, 3, 4, 0), (100, 2, 8, 4), (257, 3, 16, 8),
                                         (1000, 2, 32, 16), (4096, 2, 32, 0), (5000, 3, 32, 32)])
def test_kdtree_matches_bruteforce(n, d, b, quant):
    p = O.generate_points(n, d, 99 + n)
    if quant:  # float-quantised coordinates force many ties
        p = np.floor(p * quant) / quant
    perm, off, bmin, bmax, sd = O.kdtree(p, n, d, b)
    rp, roff, boxes, dims = kd_brute(p, n, d, b)
    assert np.array_equal(perm, rp)
    assert np.array_equal(off, roff)
    for l, lv in enumerate(boxes):
        base = (1 << l) - 1
        for s, (mn, mx) in enumerate(lv):
            assert np.array_equal(bmin[base + s], mn) and np.array_equal(bmax[base + s], mx)
            assert sd[base + s] == dims[l][s]


def test_kdtree_invariants_large():
    n, d, b = 65536, 2, 32
    p = O.generate_points(n, d, 1234)
    perm, off, bmin, bmax, sd = O.kdtree(p, n, d, b)
    assert np.array_equal(np.sort(perm), np.arange(n))  # bijection
    assert np.all(np.diff(off) == b)  # power-of-two n: full leaves
    P = p.reshape(d, n)
    L = O.kd_num_levels(n, b)
    # at every split the left child's max on the split dim <= right child's min
    for l in range(L):
        step = n >> l
        for s in range(1 << l):
            a, e = s * step, (s + 1) * step
            h = a + step // 2
            dim = sd[(1 << l) - 1 + s]
            assert P[dim, perm[a:h]].max() <= P[dim, perm[h:e]].min()


def kd_pow2left_np(p, n, d, b):
    """numpy restatement of the POWER_OF_TWO_ON_LEFT division."""
    P = p.reshape(d, n)
    perm = np.arange(n)
    size = 1 << max(0, (n - 1).bit_length())
    dims = []
    while size > b:
        lv = []
        for s in range(-(-n // size)):
            a, e = s * size, min((s + 1) * size, n)
            seg = perm[a:e]
            spans = P[:, seg].max(1) - P[:, seg].min(1)
            dim = int(np.argmax(spans))  # first maximum
            perm[a:e] = seg[np.argsort(P[dim, seg], kind="stable")]
            lv.append(dim)
        dims.append(lv)
        size >>= 1
    nseg = -(-n // size)
    off = np.array([s * size for s in range(nseg)] + [n])
    return perm, off, dims


@pytest.mark.parametrize("n,d,b,quant", [(13, 2, 2, 0), (1000, 3, 32, 0), (1000, 2, 24, 16), (4096, 2, 32, 0),
                                         (5000, 3, 64, 8)])
def test_kdtree_pow2left_matches_numpy(n, d, b, quant):
    p = O.generate_points(n, d, 7 + n)
    if quant:
        p = np.floor(p * quant) / quant
    perm, off, bmin, bmax, sd, L = O.kdtree_pow2left(p, n, d, b)
    rp, roff, dims = kd_pow2left_np(p, n, d, b)
    assert np.array_equal(perm, rp) and np.array_equal(off, roff)
    for l, lv in enumerate(dims):
        for s, dim in enumerate(lv):
            assert sd[(1 << l) - 1 + s] == dim
    # a power-of-two n gives the FULL_TREE tree
    if n & (n - 1) == 0:
        fp, foff, _, _, _ = O.kdtree(p, n, d, b)
        assert np.array_equal(perm, fp) and np.array_equal(off, foff)


# --------------------------------------------------------------- Morton ----
def morton_np(x, y):
    def spread(v):
        r = np.zeros_like(v)
        for bit in range(16):
            r |= ((v >> bit) & 1) << (2 * bit)
        return r
    return spread(x) | (spread(y) << 1)


def test_morton_exhaustive():
    xs, ys = np.meshgrid(np.arange(256, dtype=np.int64), np.arange(256, dtype=np.int64), indexing="ij")
    ref = morton_np(xs, ys)
    L = O.lib()
    for x in range(0, 256, 7):
        for y in range(256):
            assert L.or_morton_from_xy(x, y) == ref[x, y]
    # CMIndextoMOIndex: n -> (row = n % S, col = n / S)
    S = 64
    for nidx in range(0, S * S, 37):
        assert L.or_cm_to_mo(S, nidx) == morton_np(np.int64(nidx // S), np.int64(nidx % S))


# ------------------------------------------------------------ structure ----
def test_weak_structure_is_hodlr_closed_form():
    n, d, b = 4096, 2, 32
    p = O.generate_points(n, d, 5)
    perm, off, bmin, bmax, _ = O.kdtree(p, n, d, b)
    lv = O.structure(n, d, b, "weak", 1.0, bmin, bmax)
    for l in range(1, len(lv) + 1):
        exp = []
        for q in range(1 << (l - 1)):
            exp += [int(morton_np(np.int64(2 * q + 1), np.int64(2 * q))), int(morton_np(np.int64(2 * q), np.int64(2 * q + 1)))]
        assert lv[l - 1].tolist() == exp


def _xy(t):
    return O._mo_to_xy(int(t))


@pytest.mark.parametrize("adm,d", [("boxCenter", 2), ("boxCenter", 3), ("maxDiam", 2)])
def test_admissible_structure_coverage(adm, d):
    n, b = 4096, 32
    p = O.generate_points(n, d, 17)
    perm, off, bmin, bmax, _ = O.kdtree(p, n, d, b)
    lv = O.structure(n, d, b, adm, 1.0, bmin, bmax)
    L = len(lv) + 1  # numLevels
    S = 1 << (L - 1)
    cover = np.zeros((S, S), dtype=np.int32)

    def adm_np(level, u, v):
        base = (1 << level) - 1
        mxu, mnu, mxv, mnv = bmax[base + u], bmin[base + u], bmax[base + v], bmin[base + v]
        dist = np.sqrt(np.sum((0.5 * ((mxu + mnu) - (mxv + mnv))) ** 2))
        du, dv = np.sqrt(np.sum((mxu - mnu) ** 2)), np.sqrt(np.sum((mxv - mnv) ** 2))
        if du == 0 or dv == 0:
            return False
        return (0.5 * (du + dv) if adm == "boxCenter" else max(du, dv)) <= 1.0 * dist

    for l in range(1, L):
        bus = 1 << (L - 1 - l)
        for t in lv[l - 1]:
            u, v = _xy(t)
            cover[v * bus:(v + 1) * bus, u * bus:(u + 1) * bus] += 1
            if l < L - 1:
                assert u != v and adm_np(l, u, v)
            # the parent pair was refined, so it was not admissible
            if l > 1 and (u >> 1) != (v >> 1):
                assert not adm_np(l - 1, u >> 1, v >> 1)
    assert np.all(np.diag(cover) == 0)
    off_diag = cover + np.eye(S, dtype=np.int32)
    assert np.all(off_diag == 1)


def test_jds_closed_form_and_consistency():
    n, d, b = 4096, 2, 32
    p = O.generate_points(n, d, 5)
    perm, off, bmin, bmax, _ = O.kdtree(p, n, d, b)
    weak = O.structure(n, d, b, "weak", 1.0, bmin, bmax)
    for l in range(1, len(weak) + 1):
        ri, ci, tp, ip = O.jds(l, weak[l - 1])
        assert np.array_equal(ri, np.arange(1 << l))
        assert np.array_equal(ci, np.arange(1 << l) ^ 1)
        assert ip.tolist() == [0, 1 << l]
    box = O.structure(n, d, b, "boxCenter", 1.0, bmin, bmax)
    for l in range(1, len(box) + 1):
        tiles = box[l - 1]
        ri, ci, tp, ip = O.jds(l, tiles)
        assert sorted(tp.tolist()) == list(range(len(tiles)))
        rows = [(_xy(t)[1]) for t in tiles]
        counts = np.bincount(rows, minlength=1 << l)
        assert np.all(np.diff(counts[ri]) <= 0)  # descending tile counts
        for j in range(len(ip) - 1):
            for i in range(ip[j + 1] - ip[j]):
                t = tiles[tp[ip[j] + i]]
                assert _xy(t) == (ci[ip[j] + i], ri[i])


# ------------------------------------------------------------------ ARA ----
def load_fixture():
    lines = open(os.path.join(GOLDEN, "batchedMatrix.txt")).read().split("\n")
    ranks = list(map(int, lines[2].split()))
    blocks, cur = [], []
    for line in lines[3:]:
        if not line.strip():
            if cur:
                blocks.append(np.array(cur))
                cur = []
        else:
            cur.append(float(line))
    if cur:
        blocks.append(np.array(cur))
    A = [blocks[i].reshape(ranks[i], 32).T for i in range(8)]
    B = [blocks[8 + i].reshape(ranks[i], 32).T for i in range(8)]
    return ranks, A, B


def test_fixture_convention_and_ranks():
    """test/batchedMatrix.txt (KBLAS ARA, 32x32 tiles, tol 1e-5): A orthonormal,
    B = M^T A; the oracle's ARA reproduces the convention and reconstructs M."""
    ranks, A, B = load_fixture()
    tol = 1e-5
    for i in range(8):
        assert np.abs(A[i].T @ A[i] - np.eye(ranks[i])).max() < 5e-6  # printed with 6 decimals
        M = A[i] @ B[i].T
        U, V, k, capped = O.ara_dense(M, tol, bs=16, r=10, seed=0, opid=i)
        s = np.linalg.svd(M, compute_uv=False)
        assert not capped
        assert np.abs(U.T @ U - np.eye(k)).max() < 1e-12  # A = Q orthonormal
        assert np.allclose(V, M.T @ U, atol=1e-13)  # B = M^T Q
        assert np.linalg.norm(M - U @ V.T, 2) <= tol
        assert int((s > tol).sum()) <= k <= ranks[i]


@pytest.mark.parametrize("m,n,decay,tol", [(32, 32, 0.5, 1e-6), (64, 48, 0.7, 1e-8), (128, 128, 0.8, 1e-5)])
def test_ara_rank_vs_svd(m, n, decay, tol):
    rng = np.random.default_rng(m + n)
    Qa, _ = np.linalg.qr(rng.standard_normal((m, m)))
    Qb, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = decay ** np.arange(min(m, n))
    M = (Qa[:, :len(s)] * s) @ Qb[:, :len(s)].T
    U, V, k, capped = O.ara_dense(M, tol, seed=3)
    exact = int((s > tol).sum())
    # the HMT stopping rule compares Gaussian sample norms with tol/(10 sqrt(2/pi)),
    # so the rank lies between the exact tol-rank and the (tol/1000)-rank
    assert exact <= k <= int((s > tol / 1000).sum())
    assert np.linalg.norm(M - U @ V.T, 2) <= tol


def test_ara_zero_and_capped():
    U, V, k, capped = O.ara_dense(np.zeros((32, 32)), 1e-6)
    assert k == 0 and not capped
    rng = np.random.default_rng(0)
    U, V, k, capped = O.ara_dense(rng.standard_normal((32, 32)), 1e-6, max_rank=16)
    assert k == 16 and capped


# ------------------------------------------------------------- pipeline ----
@pytest.fixture(scope="module")
def small_h():
    n, d, b, tol = 1024, 2, 32, 1e-6
    p = O.generate_points(n, d, 1234)
    return O.OracleHMatrix(p, n, d, b, tol=tol, max_rank_policy=1), p


def test_pipeline_expand_vs_dense(small_h):
    H, p = small_h
    K = O.dense_matrix(p, H.n, H.d, H.perm, H.ell)
    # the dense kernel matrix in KD order, checked against numpy
    P = p.reshape(H.d, H.n)[:, H.perm]
    Kn = np.exp(-np.sqrt(((P[:, :, None] - P[:, None, :]) ** 2).sum(0)) / H.ell)
    assert np.allclose(K, Kn, rtol=1e-14, atol=0)
    He = H.expand()
    err = np.linalg.norm(He - K) / np.linalg.norm(K)
    assert err < 10 * 1e-6
    # every level's tiles meet their (doubled) tolerance in the spectral norm
    L = H.num_levels
    for l in range(1, L):
        ranks, Us, Vs = H.level(l)
        mr, tol, caps = H.level_info(l)
        assert caps == 0
        bus = 1 << (L - 1 - l)
        for t, ti in enumerate(H.tiles[l][:4]):
            u, v = _xy(ti)
            blk = K[v * bus * 32:(v + 1) * bus * 32, u * bus * 32:(u + 1) * bus * 32]
            assert np.linalg.norm(blk - Us[t] @ Vs[t].T, 2) <= 4 * tol


def test_pipeline_tolerance_and_rank_schedule():
    n, d, b = 2048, 2, 32
    p = O.generate_points(n, d, 1)
    H = O.OracleHMatrix(p, n, d, b, tol=1e-5, max_rank_policy=0)
    L = H.num_levels
    mr = np.zeros(L - 2, dtype=np.uint32)
    O.lib().or_generate_max_ranks(L, b, mr.ctypes.data_as(__import__("ctypes").c_void_p))
    assert mr.tolist() == [b * (1 << i) // (4 if i > 5 else 1) for i in range(L - 2)]
    for l in range(1, L - 1):
        maxRank, tol, caps = H.level_info(l)
        assert maxRank == min(int(mr[L - l - 2]), (1 << (L - 1 - l)) * b)
        assert abs(tol - np.float32(1e-5) * 2 ** (L - 1 - l)) <= 1e-12
        assert H.level(l)[0].max() <= maxRank


def test_morton_reorder_round_trip(small_h):
    H, _ = small_h
    S, b = H.S, H.b
    rc, rm = H.tlr_ranks(False), H.tlr_ranks(True)
    L = O.lib()
    for i in range(S * S):
        assert rm[L.or_cm_to_mo(S, i)] == rc[i]


def test_matvec_matches_expansion(small_h):
    H, _ = small_h
    He = H.expand()
    X = O.generate_rhs(H.n, 3, 11)
    Y = H.matvec(X)
    assert np.linalg.norm(Y - He @ X) / np.linalg.norm(He @ X) < 1e-13
    YT = H.matvec(X, transpose=True)
    assert np.linalg.norm(YT - He.T @ X) / np.linalg.norm(He.T @ X) < 1e-13


def test_sampled_rows(small_h):
    H, p = small_h
    X = O.generate_rhs(H.n, 2, 3)
    rows = np.array([0, 5, 17, 1023, 511], dtype=np.int32)
    K = O.dense_matrix(p, H.n, H.d, H.perm, H.ell)
    Y = O.sampled_rows(p, H.n, H.d, H.perm, H.ell, rows, X)
    assert np.allclose(Y, K[rows] @ X, rtol=1e-12, atol=1e-12)


def test_boxcenter_pipeline_accuracy():
    n, d, b, tol = 2048, 3, 32, 1e-5
    p = O.generate_points(n, d, 21)
    H = O.OracleHMatrix(p, n, d, b, adm="boxCenter", tol=tol, max_rank_policy=1)
    X = O.generate_rhs(n, 1, 2)
    K = O.dense_matrix(p, n, d, H.perm, H.ell)
    err = np.linalg.norm(H.matvec(X) - K @ X) / np.linalg.norm(K @ X)
    assert err < 1e-5


def test_matvec_on_supplied_factors_matches(small_h):
    """or_matvec_factors (the CPU baseline's matvec) == or_hm_matvec."""
    H, _ = small_h
    L = H.num_levels
    tiles, ranks, offs, Us, Vs = [], [], [], [], []
    for l in range(1, L):
        r, U, V = H.level(l)
        m = (1 << (L - 1 - l)) * H.b
        tiles.append(H.tiles[l])
        ranks.append(r)
        o = np.zeros(len(r) + 1, dtype=np.int64)
        o[1:] = np.cumsum(r)
        offs.append(o)
        Us.append(np.concatenate([u.T.reshape(-1) for u in U]) if len(U) else np.zeros(1))
        Vs.append(np.concatenate([v.T.reshape(-1) for v in V]) if len(V) else np.zeros(1))
    X = O.generate_rhs(H.n, 2, 8)
    Y = O.matvec_factors(H.n, H.b, L, tiles, ranks, offs, Us, Vs, H.diag(), X)
    Yr = H.matvec(X)
    assert np.linalg.norm(Y - Yr) / np.linalg.norm(Yr) < 1e-14
    # the OpenBLAS CPU baseline computes the same product (other summation order)
    if O.openblas_path() is not None:
        for k in (1, 5):
            Xk = O.generate_rhs(H.n, k, 9 + k)
            Yb = O.blas_matvec_factors(H.n, H.b, L, tiles, ranks, offs, Us, Vs, H.diag(), Xk)
            Yk = H.matvec(Xk)
            assert np.linalg.norm(Yb - Yk) / np.linalg.norm(Yk) < 1e-13
