"""ctypes wrapper of the CPU oracle (oracle/hmx_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The product library never uses it.

What pins each oracle function (checked in tests/test_oracle.py):
  or_philox4x32_10   Random123 known-answer vectors for Philox4x32-10.
  or_generate_points range [0,1), moments, determinism.
  or_kdtree          brute-force restatement with numpy.lexsort on tiny inputs
                     (including forced ties and float-quantised coordinates),
                     bijection / leaf-size / split-order invariants.
  or_structure       weak admissibility closed form (HODLR tile lists); for
                     boxCenter a brute-force O(S^2) coverage check and a numpy
                     re-evaluation of every recorded admissibility decision.
  or_jds             HODLR closed form (rowIndices identity, column = row ^ 1).
  morton helpers     numpy bit interleave over an exhaustive 2^8 x 2^8 grid.
  or_ara_dense       test/batchedMatrix.txt fixture (output convention A
                     orthonormal, B = M^T A, ranks); numpy SVD: spectral error
                     <= tol and rank >= exact count of singular values > tol.
  TLR / Morton / H   expand(H) vs dense kernel matrix (numpy) relative error,
                     Morton reorder round trip, per-level rank caps.
  or_hm_matvec       expand(H) @ X with numpy (rel 1e-13), transpose identity.
  or_sampled_rows    numpy dense kernel rows.
"""
import ctypes
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

ADM = {"weak": 0, "boxCenter": 1, "maxDiam": 2}


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        P = ctypes.POINTER
        c_d, c_i, c_f, c_u64, c_u32 = ctypes.c_double, ctypes.c_int, ctypes.c_float, ctypes.c_uint64, ctypes.c_uint32
        vp = ctypes.c_void_p
        L.or_philox4x32_10.argtypes = [vp, vp, vp]
        L.or_generate_points.argtypes = [c_i, c_i, c_u64, vp]
        L.or_generate_rhs.argtypes = [c_i, c_i, c_u64, vp]
        L.or_gauss.argtypes = [c_u64, c_u32, c_u64, c_u32, c_u32]
        L.or_gauss.restype = c_d
        L.or_kd_num_levels.argtypes = [c_i, c_i]
        L.or_upper_pow2.argtypes = [c_i]
        L.or_kdtree.argtypes = [c_i, c_i, c_i, vp, vp, vp, vp, vp, vp]
        L.or_kdtree_pow2left.argtypes = [c_i, c_i, c_i, vp, vp, vp, vp, vp, vp, vp]
        L.or_morton_from_xy.argtypes = [c_u32, c_u32]
        L.or_cm_to_mo.argtypes = [c_i, c_i]
        L.or_structure.argtypes = [c_i, c_i, c_i, c_i, c_f, vp, vp, vp, vp, vp]
        L.or_jds.argtypes = [c_i, c_i, vp, vp, vp, vp, vp]
        L.or_ara_threshold.argtypes = [c_d]
        L.or_ara_threshold.restype = c_d
        L.or_ara_dense.argtypes = [c_i, c_i, vp, c_d, c_i, c_i, c_i, c_u64, c_u64, vp, vp, vp]
        L.or_hm_create.argtypes = [c_i, c_i, c_i, c_i, c_f, c_f, c_d, c_i, c_i, c_i, c_u64, vp]
        L.or_hm_create.restype = vp
        L.or_hm_destroy.argtypes = [vp]
        for f in ("or_hm_kdtree", "or_hm_structure"):
            getattr(L, f).argtypes = [vp]
        L.or_hm_tlr.argtypes = [vp]
        L.or_hm_tlr.restype = ctypes.c_longlong
        L.or_hm_to_morton.argtypes = [vp]
        L.or_hm_hmatrix.argtypes = [vp]
        L.or_hm_get_int.argtypes = [vp, c_i]
        for f in ("or_hm_perm", "or_hm_offsets", "or_hm_split_dims", "or_hm_num_tiles"):
            getattr(L, f).argtypes = [vp]
            getattr(L, f).restype = P(c_i)
        for f in ("or_hm_bbmin", "or_hm_bbmax", "or_hm_diag"):
            getattr(L, f).argtypes = [vp]
            getattr(L, f).restype = P(c_d)
        L.or_hm_tiles.argtypes = [vp, c_i]
        L.or_hm_tiles.restype = P(c_i)
        L.or_hm_tlr_ranks.argtypes = [vp, c_i]
        L.or_hm_tlr_ranks.restype = P(c_i)
        L.or_hm_tlr_offsets.argtypes = [vp, c_i]
        L.or_hm_tlr_offsets.restype = P(ctypes.c_longlong)
        L.or_hm_tlr_U.argtypes = [vp, c_i]
        L.or_hm_tlr_U.restype = P(c_d)
        L.or_hm_tlr_V.argtypes = [vp, c_i]
        L.or_hm_tlr_V.restype = P(c_d)
        L.or_hm_level_ranks.argtypes = [vp, c_i]
        L.or_hm_level_ranks.restype = P(c_i)
        L.or_hm_level_offsets.argtypes = [vp, c_i]
        L.or_hm_level_offsets.restype = P(ctypes.c_longlong)
        L.or_hm_level_U.argtypes = [vp, c_i]
        L.or_hm_level_U.restype = P(c_d)
        L.or_hm_level_V.argtypes = [vp, c_i]
        L.or_hm_level_V.restype = P(c_d)
        L.or_hm_level_max_rank.argtypes = [vp, c_i]
        L.or_hm_level_tol.argtypes = [vp, c_i]
        L.or_hm_level_tol.restype = c_f
        L.or_hm_level_cap_hits.argtypes = [vp, c_i]
        L.or_hm_matvec.argtypes = [vp, c_i, vp, vp, c_i]
        L.or_sampled_rows.argtypes = [c_i, c_i, vp, vp, c_d, c_i, vp, c_i, vp, vp]
        L.or_dense_matrix.argtypes = [c_i, c_i, vp, vp, c_d, vp]
        L.or_matvec_factors.argtypes = [c_i, c_i, c_i, vp, vp, vp, vp, vp, vp, vp, c_i, vp, vp]
        L.or_num_threads.restype = c_i
        L.cb_open.argtypes = [ctypes.c_char_p]
        L.cb_open.restype = c_i
        L.cb_matvec_factors.argtypes = [c_i, c_i, c_i, vp, vp, vp, vp, vp, vp, vp, c_i, vp, vp]
        L.cb_matvec_factors.restype = c_i
        L.or_default_ell.argtypes = [c_i]
        L.or_default_ell.restype = c_d
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().or_philox4x32_10(_p(c), _p(k), _p(o))
    return o


def generate_points(n, d, seed=1234):
    p = np.zeros(n * d, dtype=np.float64)
    lib().or_generate_points(n, d, seed, _p(p))
    return p


def generate_rhs(n, k, seed=1234):
    x = np.zeros(n * k, dtype=np.float64)
    lib().or_generate_rhs(n, k, seed, _p(x))
    return x.reshape(k, n).T.copy(order="F")  # n x k column-major


def kd_num_levels(n, b):
    return lib().or_kd_num_levels(n, b)


def kdtree(points, n, d, b):
    lv = kd_num_levels(n, b)
    S = 1 << lv
    nb = (1 << (lv + 1)) - 1
    perm = np.zeros(max(n, 1), dtype=np.int32)
    off = np.zeros(S + 1, dtype=np.int32)
    bmin = np.zeros(nb * d)
    bmax = np.zeros(nb * d)
    sd = np.zeros(nb, dtype=np.int32)
    p = np.ascontiguousarray(points, dtype=np.float64)
    lib().or_kdtree(n, d, b, _p(p), _p(perm), _p(off), _p(bmin), _p(bmax), _p(sd))
    return perm[:n], off, bmin.reshape(nb, d), bmax.reshape(nb, d), sd


def kdtree_pow2left(points, n, d, b):
    """POWER_OF_TWO_ON_LEFT tree: (perm, offsets, bbmin, bbmax, split_dims, levels)."""
    p = np.ascontiguousarray(points, dtype=np.float64)
    size, L = 1 << max(0, int(n - 1).bit_length()), 0
    while size > b:
        size >>= 1
        L += 1
    nbox = (2 << L) - 1
    perm = np.zeros(n, np.int32)
    nseg = (n + size - 1) // size
    off = np.zeros(nseg + 1, np.int32)
    bmin = np.zeros(nbox * d)
    bmax = np.zeros(nbox * d)
    sd = np.zeros(nbox, np.int32)
    ns = ctypes.c_int(0)
    lv = lib().or_kdtree_pow2left(n, d, b, _p(p), _p(perm), _p(off), _p(bmin), _p(bmax), _p(sd), ctypes.byref(ns))
    assert ns.value == nseg and lv == L
    return perm, off, bmin.reshape(nbox, d), bmax.reshape(nbox, d), sd, L


def structure(n, d, b, adm, eta, bbmin, bbmax):
    n2 = lib().or_upper_pow2(n)
    L = (n2 // b).bit_length()  # ctz(n2/b)+1 for a power of two
    caps = [(1 << (l + 1)) * ((1 << (l + 1)) - 1) for l in range(L)]
    tiles = np.zeros(max(sum(caps), 1), dtype=np.int32)
    nt = np.zeros(L, dtype=np.int32)
    base = np.zeros(L, dtype=np.int64)
    bmin = np.ascontiguousarray(bbmin, dtype=np.float64)
    bmax = np.ascontiguousarray(bbmax, dtype=np.float64)
    got = lib().or_structure(n, d, b, ADM[adm] if isinstance(adm, str) else adm, eta, _p(bmin), _p(bmax), _p(nt),
                             _p(tiles), _p(base))
    assert got == L
    return [tiles[base[l - 1]:base[l - 1] + nt[l - 1]].copy() for l in range(1, L)]


def jds(level, tiles):
    nr = 1 << level
    t = np.ascontiguousarray(tiles, dtype=np.int32)
    ri = np.zeros(nr, dtype=np.int32)
    ci = np.zeros(max(len(t), 1), dtype=np.int32)
    tp = np.zeros(max(len(t), 1), dtype=np.int32)
    ip = np.zeros(nr + 1, dtype=np.int32)
    nd = lib().or_jds(level, len(t), _p(t), _p(ri), _p(ci), _p(tp), _p(ip))
    return ri, ci[:len(t)], tp[:len(t)], ip[:nd + 1]


def morton(x, y):
    return lib().or_morton_from_xy(x, y)


def ara_threshold(tol):
    return lib().or_ara_threshold(tol)


def ara_dense(A, tol, bs=16, r=10, max_rank=None, seed=0, opid=0):
    A = np.asfortranarray(A, dtype=np.float64)
    m, n = A.shape
    mr = max_rank if max_rank is not None else min(m, n)
    U = np.zeros((mr, m))
    V = np.zeros((mr, n))
    capped = ctypes.c_int(0)
    k = lib().or_ara_dense(m, n, _p(A), tol, bs, r, mr, seed, opid, _p(U), _p(V), ctypes.byref(capped))
    return U[:k].T.copy(), V[:k].T.copy(), k, capped.value


class OracleHMatrix:
    """Whole-pipeline CPU oracle: points -> KD-tree -> structure -> TLR -> Morton -> H."""

    def __init__(self, points, n, d, b, adm="weak", eta=1.0, tol=1e-6, ell=None, bs=16, r=10, max_rank_policy=0,
                 seed=0, stages="all"):
        self.n, self.d, self.b = n, d, b
        self.ell = lib().or_default_ell(d) if ell is None else ell
        p = np.ascontiguousarray(points, dtype=np.float64)
        self.points = p
        self.h = lib().or_hm_create(n, d, b, ADM[adm] if isinstance(adm, str) else adm, eta, tol, self.ell, bs, r,
                                    max_rank_policy, seed, _p(p))
        L = lib()
        L.or_hm_kdtree(self.h)
        self.S = L.or_hm_get_int(self.h, 0)
        self.perm = np.ctypeslib.as_array(L.or_hm_perm(self.h), shape=(n,)).copy()
        self.offsets = np.ctypeslib.as_array(L.or_hm_offsets(self.h), shape=(self.S + 1,)).copy()
        L.or_hm_structure(self.h)
        self.num_levels = L.or_hm_get_int(self.h, 2)
        nt = np.ctypeslib.as_array(L.or_hm_num_tiles(self.h), shape=(self.num_levels,))
        self.tiles = {l: np.ctypeslib.as_array(L.or_hm_tiles(self.h, l), shape=(max(nt[l - 1], 1),))[:nt[l - 1]].copy()
                      for l in range(1, self.num_levels)}
        if stages == "structure":
            return
        self.rank_sum = L.or_hm_tlr(self.h)
        L.or_hm_to_morton(self.h)
        L.or_hm_hmatrix(self.h)

    def __del__(self):
        try:
            lib().or_hm_destroy(self.h)
        except Exception:
            pass

    def tlr_ranks(self, morton=False):
        S = self.S
        return np.ctypeslib.as_array(lib().or_hm_tlr_ranks(self.h, int(morton)), shape=(S * S,)).copy()

    def diag(self):
        return np.ctypeslib.as_array(lib().or_hm_diag(self.h), shape=(self.S * self.b * self.b,)).copy()

    def level(self, l):
        """(ranks, U list, V list) for H level l (1..num_levels-1); U_t is m x r_t."""
        if not 1 <= l < self.num_levels:
            raise IndexError(f"level {l} outside 1..{self.num_levels - 1}")
        L = lib()
        nt = len(self.tiles[l])
        ranks = np.ctypeslib.as_array(L.or_hm_level_ranks(self.h, l), shape=(max(nt, 1),))[:nt].copy()
        m = (1 << (self.num_levels - 1 - l)) * self.b
        tot = int(ranks.sum())
        if tot == 0:
            return ranks, [np.zeros((m, 0)) for _ in ranks], [np.zeros((m, 0)) for _ in ranks]
        Uf = np.ctypeslib.as_array(L.or_hm_level_U(self.h, l), shape=(tot * m,)).copy()
        Vf = np.ctypeslib.as_array(L.or_hm_level_V(self.h, l), shape=(tot * m,)).copy()
        Us, Vs, o = [], [], 0
        for r in ranks:
            Us.append(Uf[o * m:(o + r) * m].reshape(r, m).T)
            Vs.append(Vf[o * m:(o + r) * m].reshape(r, m).T)
            o += r
        return ranks, Us, Vs

    def level_info(self, l):
        if not 1 <= l < self.num_levels:
            raise IndexError(f"level {l} outside 1..{self.num_levels - 1}")
        L = lib()
        return L.or_hm_level_max_rank(self.h, l), L.or_hm_level_tol(self.h, l), L.or_hm_level_cap_hits(self.h, l)

    def matvec(self, X, transpose=False):
        X = np.asfortranarray(X, dtype=np.float64)
        if X.ndim == 1:
            X = X[:, None]
        Y = np.zeros_like(X, order="F")
        lib().or_hm_matvec(self.h, X.shape[1], _p(X), _p(Y), int(transpose))
        return Y

    def expand(self):
        """Dense expansion of H in KD order (n x n), small n only."""
        n, b, L = self.n, self.b, self.num_levels
        H = np.zeros((n, n))
        off = self.offsets
        D = self.diag().reshape(self.S, b, b)
        for leaf in range(self.S):
            a, e = off[leaf], off[leaf + 1]
            H[a:e, a:e] = D[leaf].T[: e - a, : e - a]
        for l in range(1, L):
            ranks, Us, Vs = self.level(l)
            bus = 1 << (L - 1 - l)
            for t, ti in enumerate(self.tiles[l]):
                u, v = _mo_to_xy(int(ti))
                rows = _cluster_rows(off, v, bus, b)
                cols = _cluster_rows(off, u, bus, b)
                blk = Us[t] @ Vs[t].T
                H[np.ix_(rows[1], cols[1])] += blk[np.ix_(rows[0], cols[0])]
        return H


def _compact(x):
    x &= 0x55555555
    x = (x | (x >> 1)) & 0x33333333
    x = (x | (x >> 2)) & 0x0F0F0F0F
    x = (x | (x >> 4)) & 0x00FF00FF
    x = (x | (x >> 8)) & 0x0000FFFF
    return x


def _mo_to_xy(d):
    return _compact(d), _compact(d >> 1)


def _cluster_rows(off, cl, bus, b):
    """(padded positions, point positions) of cluster cl spanning bus leaves."""
    pad, pos = [], []
    for j in range(bus):
        leaf = cl * bus + j
        a, e = off[leaf], off[leaf + 1]
        for q in range(e - a):
            pad.append(j * b + q)
            pos.append(a + q)
    return np.array(pad, dtype=np.int64), np.array(pos, dtype=np.int64)


def sampled_rows(points, n, d, perm, ell, rows, X):
    X = np.asfortranarray(X, dtype=np.float64)
    if X.ndim == 1:
        X = X[:, None]
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    perm = np.ascontiguousarray(perm, dtype=np.int32)
    p = np.ascontiguousarray(points, dtype=np.float64)
    Y = np.zeros((len(rows), X.shape[1]), order="F")
    lib().or_sampled_rows(n, d, _p(p), _p(perm), ell, len(rows), _p(rows), X.shape[1], _p(X), _p(Y))
    return Y


def dense_matrix(points, n, d, perm, ell):
    K = np.zeros((n, n), order="F")
    p = np.ascontiguousarray(points, dtype=np.float64)
    perm = np.ascontiguousarray(perm, dtype=np.int32)
    lib().or_dense_matrix(n, d, _p(p), _p(perm), ell, _p(K))
    return K


def matvec_factors(n, b, num_levels, tiles, ranks, offs, U, V, diag, X):
    """CPU H-matvec on externally built factors (lists indexed by level-1)."""
    nl = num_levels - 1
    keep = []

    def arr(lst, ctype, dtype):
        a = [np.ascontiguousarray(x, dtype=dtype) for x in lst]
        keep.extend(a)
        return (ctypes.c_void_p * nl)(*[x.ctypes.data for x in a])

    X = np.asfortranarray(X, dtype=np.float64)
    if X.ndim == 1:
        X = X[:, None]
    Y = np.zeros_like(X, order="F")
    nt = np.array([len(t) for t in tiles], dtype=np.int32)
    d = np.ascontiguousarray(diag, dtype=np.float64)
    lib().or_matvec_factors(n, b, num_levels, _p(nt), arr(tiles, None, np.int32), arr(ranks, None, np.int32),
                            arr(offs, None, np.int64), arr(U, None, np.float64), arr(V, None, np.float64), _p(d),
                            X.shape[1], _p(X), _p(Y))
    return Y


def num_threads():
    return lib().or_num_threads()


def openblas_path():
    """numpy's bundled ILP64 OpenBLAS (None when absent)."""
    import glob
    base = os.path.join(os.path.dirname(os.path.dirname(np.__file__)), "numpy.libs")
    hits = sorted(glob.glob(os.path.join(base, "libscipy_openblas64_*.so")))
    return hits[0] if hits else None


def blas_matvec_factors(n, b, num_levels, tiles, ranks, offs, U, V, diag, X):
    """matvec_factors with every block product in OpenBLAS dgemm (OpenMP over
    tiles and leaves): the CPU baseline of the benchmark."""
    path = openblas_path()
    if path is None or lib().cb_open(path.encode()) != 0:
        raise RuntimeError("OpenBLAS not found")
    nl = num_levels - 1
    keep = []

    def arr(lst, dtype):
        a = [np.ascontiguousarray(x, dtype=dtype) for x in lst]
        keep.extend(a)
        return (ctypes.c_void_p * nl)(*[x.ctypes.data for x in a])

    X = np.asfortranarray(X, dtype=np.float64)
    if X.ndim == 1:
        X = X[:, None]
    Y = np.zeros_like(X, order="F")
    nt = np.array([len(t) for t in tiles], dtype=np.int32)
    d = np.ascontiguousarray(diag, dtype=np.float64)
    st = lib().cb_matvec_factors(n, b, num_levels, _p(nt), arr(tiles, np.int32), arr(ranks, np.int32),
                                 arr(offs, np.int64), arr(U, np.float64), arr(V, np.float64), _p(d), X.shape[1],
                                 _p(X), _p(Y))
    if st != 0:
        raise RuntimeError("OpenBLAS not loaded")
    return Y
