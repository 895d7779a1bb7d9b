"""Python binding of libhmx (include/hmx.h) -- the MI355X H-matrix library.

Thin ctypes layer: device memory comes from torch tensors (plumbing only), all
compute runs in the gfx950 kernels of libhmx.so.  Function names follow the
reference entry points (src/main.cu:52-173):

    generateDataset -> constructKDTree -> constructHMatrixStructure ->
    createColumnMajorLRMatrix -> convertColumnMajorToMorton ->
    generateHMatrixFromStruct -> HMatrixVecMult

Loading fails loudly when libhmx.so is missing: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libhmx.so")
HEADER_PATH = os.path.join(os.path.dirname(os.path.dirname(_HERE)), "include", "hmx.h")

HMX_OK, HMX_EINVAL, HMX_ENOMEM, HMX_EDEVICE, HMX_ERANKCAP, HMX_EUNSUPPORTED, HMX_ESTATE = range(7)
POWER_OF_TWO_ON_LEFT, DIVIDE_IN_HALF, FULL_TREE = 0, 1, 2
WEAK, BOX_CENTER, MAX_DIAM = 0, 1, 2
ADMISSIBILITY = {"weak": WEAK, "boxCenter": BOX_CENTER, "maxDiam": MAX_DIAM}
MAXRANK_REF, MAXRANK_FULL = 0, 1
COLUMN_MAJOR, MORTON = 0, 1


class HmxError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"hmx status {status}: {msg}")
        self.status = status


class Config(ctypes.Structure):
    """hmx_config (src/config.h:109-117 plus the reference's hard-coded parameters)."""
    _fields_ = [
        ("divMethod", ctypes.c_int32),
        ("admissibilityCondition", ctypes.c_int32),
        ("numberOfInputPoints", ctypes.c_uint32),
        ("dimensionOfInputPoints", ctypes.c_uint32),
        ("bucketSize", ctypes.c_uint32),
        ("vectorWidth", ctypes.c_uint32),
        ("lowestLevelTolerance", ctypes.c_float),
        ("eta", ctypes.c_float),
        ("correlationLength", ctypes.c_double),
        ("araBlockSize", ctypes.c_int32),
        ("araR", ctypes.c_int32),
        ("maxRankPolicy", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
        ("strictRankCap", ctypes.c_int32),
        ("symmetricTiles", ctypes.c_int32),
    ]


_lib = None

# (name, restype, argtypes)
_vp, _u32, _i32, _u64, _i64, _f32, _f64 = (ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint64,
                                           ctypes.c_int64, ctypes.c_float, ctypes.c_double)
_P = ctypes.POINTER
_SIGS = [
    ("hmx_create", _i32, [_P(_vp), _vp]),
    ("hmx_destroy", _i32, [_vp]),
    ("hmx_last_error", ctypes.c_char_p, [_vp]),
    ("hmx_config_default", None, [_P(Config)]),
    ("hmx_version", ctypes.c_char_p, []),
    ("hmx_generateDataset", _i32, [_vp, _u32, _u32, _u64, _vp]),
    ("hmx_generateRandomVector", _i32, [_vp, _u32, _u32, _u64, _vp]),
    ("hmx_constructKDTree", _i32, [_vp, _u32, _u32, _u32, _i32, _vp, _P(_vp)]),
    ("hmx_generateDatasetBatch", _i32, [_vp, _i32, _P(_u32), _P(_u32), _P(_u64), _P(_vp)]),
    ("hmx_constructKDTreeBatch", _i32, [_vp, _i32, _P(_u32), _P(_u32), _u32, _i32, _P(_vp), _P(_vp)]),
    ("hmx_kdtree_info", _i32, [_vp, _P(_u32), _P(_u32), _P(_u32)]),
    ("hmx_kdtree_copy", _i32, [_vp, _vp, _vp, _vp, _vp, _vp]),
    ("hmx_kdtree_device_perm", _vp, [_vp]),
    ("hmx_reorderVectors", _i32, [_vp, _vp, _u32, _i32, _vp, _vp]),
    ("hmx_freeKDTree", _i32, [_vp]),
    ("hmx_constructHMatrixStructure", _i32, [_vp, _u32, _u32, _u32, _i32, _f32, _vp, _P(_vp)]),
    ("hmx_constructHMatrixStructureDevice", _i32, [_vp, _u32, _u32, _u32, _i32, _f32, _vp, _P(_vp)]),
    ("hmx_structure_from_boxes", _i32, [_u32, _u32, _u32, _i32, _f32, _vp, _vp, _P(_vp)]),
    ("hmx_structure_info", _i32, [_vp, _P(_i32), _vp]),
    ("hmx_structure_tiles", _i32, [_vp, _i32, _vp]),
    ("hmx_structure_jds", _i32, [_vp, _i32, _vp, _vp, _vp, _vp, _P(_i32)]),
    ("hmx_freeHMatrixStructure", _i32, [_vp]),
    ("hmx_createColumnMajorLRMatrix", _i32, [_vp, _P(Config), _vp, _vp, _P(_vp), _P(_u64)]),
    ("hmx_convertColumnMajorToMorton", _i32, [_vp, _vp, _P(_vp)]),
    ("hmx_createColumnMajorLRMatrixPiece", _i32, [_vp, _P(Config), _vp, _vp, _u32, _u32, _u32, _P(_vp), _P(_u64)]),
    ("hmx_tlr_info", _i32, [_vp, _P(_u32), _P(_u32), _P(_i32), _P(_u64), _P(_u64)]),
    ("hmx_tlr_copy", _i32, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("hmx_freeMatrix", _i32, [_vp]),
    ("hmx_generateMaxRanks", None, [_u32, _u32, _vp]),
    ("hmx_generateHMatrixFromStruct", _i32, [_vp, _P(Config), _vp, _vp, _vp, _P(_vp)]),
    ("hmx_generateHMatrixDirect", _i32, [_vp, _P(Config), _vp, _vp, _vp, _P(_vp)]),
    ("hmx_hmatrix_level_info", _i32, [_vp, _i32, _P(_i32), _P(_i64), _P(_i32), _P(_f32), _P(_i32)]),
    ("hmx_hmatrix_level_copy", _i32, [_vp, _vp, _i32, _vp, _vp, _vp]),
    ("hmx_hmatrix_diag_copy", _i32, [_vp, _vp, _vp]),
    ("hmx_hmatrix_bytes", _i32, [_vp, _P(_u64), _P(_u64), _P(_u64)]),
    ("hmx_freeHMatrix", _i32, [_vp]),
    ("hmx_HMatrixVecMult", _i32, [_vp, _vp, _u32, _vp, _vp]),
    ("hmx_HMatrixVecMultTranspose", _i32, [_vp, _vp, _u32, _vp, _vp]),
    ("hmx_HMatrixVecMultNoAlloc", _i32, [_vp, _vp, _u32, _vp, _vp]),
    ("hmx_matvec_profile", _i32, [_vp, _vp, _u32, _vp, _vp, _i32, _vp]),
    ("hmx_denseRowsXVector", _i32, [_vp, _u32, _u32, _vp, _vp, _f64, _u32, _vp, _u32, _vp, _vp]),
    ("hmx_matvec_profile_multi", _i32, [_vp, _vp, _u32, _vp, _vp, _i32, _i32, _vp]),
    ("hmx_hmatrix_make_mixed", _i32, [_vp, _vp, _f64]),
    ("hmx_hmatrix_mixed_info", _i32, [_vp, _P(_i64), _P(_i64), _P(_f64), _P(_f64), _P(_u64)]),
    ("hmx_HMatrixVecMultMixed", _i32, [_vp, _vp, _u32, _vp, _vp]),
    ("hmx_selftest_mfma", _i32, [_vp, _i32, _vp, _vp, _vp]),
    ("hmx_stream_read_probe", _i32, [_vp, _vp, _u64, _i32, _P(_f32)]),
    ("hmx_fp64_peak_probe", _i32, [_vp, _i32, _i32, _P(_f32)]),
    ("hmx_hmatvec_grad_factors", _i32, [_vp, _vp, _u32, _vp, _vp, _vp, _vp]),
    ("hmx_hmatrix_level_get", _i32, [_vp, _vp, _i32, _vp, _vp]),
    ("hmx_hmatrix_set_level", _i32, [_vp, _vp, _i32, _vp, _vp]),
    ("hmx_hmatrix_plan_info", _i32, [_vp, _i32, _P(_u64), _P(_i64), _P(_i64)]),
    ("hmx_hmatrix_plan_phase_bytes", _i32, [_vp, _i32, _P(_u64), _P(_u64)]),
    ("hmx_hmatrix_set_plan_options", _i32, [_vp, _i32, _i32, _i32]),
    ("hmx_hmatrix_set_fp64_layout", _i32, [_vp, _i32]),
    ("hmx_hmatrix_save", _i32, [_vp, _vp, ctypes.c_char_p]),
    ("hmx_hmatrix_load", _i32, [_vp, ctypes.c_char_p, _P(_vp)]),
    ("hmx_generateHMatrixDirectBatch", _i32, [_vp, _i32, _vp, _vp, _vp, _vp, _vp]),
    ("hmx_batch_create", _i32, [_vp, _i32, _vp, _P(_vp)]),
    ("hmx_batch_info", _i32, [_vp, _P(_i64), _P(_i64), _P(_i32)]),
    ("hmx_batch_matvec", _i32, [_vp, _vp, _u32, _vp, _vp, _i32]),
    ("hmx_batch_tune", _i32, [_vp, _vp, _u32, _P(_i32), _P(_f32)]),
    ("hmx_batch_free", _i32, [_vp]),
    ("hmx_hmatrix_build_info", _i32, [_vp, _P(_f64), _P(_f64), _P(_f64)]),
    ("hmx_debug_set_alloc_fault", None, [_i64]),
    ("hmx_debug_alloc_count", _i64, []),
    ("hmx_debug_live_bytes", _i64, []),
]


def lib():
    """Load libhmx.so (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libhmx.so not found at {LIB_PATH}: build it (python __graft_entry__.py build)")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return [s[0] for s in _SIGS]


def default_config(**kw):
    cfg = Config()
    lib().hmx_config_default(ctypes.byref(cfg))
    for k, v in kw.items():
        if not hasattr(cfg, k):
            raise AttributeError(k)
        setattr(cfg, k, v)
    return cfg


def _check(st, ctx=None):
    if st != HMX_OK:
        msg = lib().hmx_last_error(ctx.handle if ctx is not None else None)
        raise HmxError(st, (msg or b"").decode())


def _ptr(t):
    """Device pointer of a torch tensor (or an int address)."""
    if t is None:
        return None
    if isinstance(t, int):
        return ctypes.c_void_p(t)
    return ctypes.c_void_p(t.data_ptr())


def _np(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Context:
    def __init__(self, stream=None):
        h = ctypes.c_void_p()
        s = None
        if stream is not None:
            s = ctypes.c_void_p(stream if isinstance(stream, int) else stream.cuda_stream)
        _check(lib().hmx_create(ctypes.byref(h), s))
        self.handle = h

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                lib().hmx_destroy(self.handle)
                self.handle = None
        except Exception:  # interpreter shutdown
            pass


class _Obj:
    _free = None

    def __init__(self, handle):
        self.handle = handle

    def free(self):
        """Releases the library object now (the reference's free* call); later use raises."""
        if getattr(self, "handle", None) and self._free:
            h, self.handle = self.handle, None
            getattr(lib(), self._free)(h)

    def __del__(self):
        try:
            self.free()
        except Exception:  # interpreter shutdown
            pass


class KDTree(_Obj):
    _free = "hmx_freeKDTree"

    def info(self):
        S, b, L = _u32(), _u32(), _u32()
        _check(lib().hmx_kdtree_info(self.handle, ctypes.byref(S), ctypes.byref(b), ctypes.byref(L)))
        return S.value, b.value, L.value

    def copy(self, ctx, n, d):
        S, _, L = self.info()
        perm = np.zeros(n, dtype=np.int32)
        off = np.zeros(S + 1, dtype=np.int32)
        nb = (2 << L) - 1
        bmin = np.zeros(nb * d)
        bmax = np.zeros(nb * d)
        _check(lib().hmx_kdtree_copy(ctx.handle, self.handle, _np(perm), _np(off), _np(bmin), _np(bmax)), ctx)
        return perm, off, bmin.reshape(nb, d), bmax.reshape(nb, d)

    def device_perm(self):
        return lib().hmx_kdtree_device_perm(self.handle)

    def reorder(self, ctx, x, to_original, out=None):
        """k vectors ((n,) or (k, n) FP64 on the GPU) from the original point
        order to the KD order (to_original=False) or back."""
        import torch
        k = 1 if x.dim() == 1 else x.shape[0]
        if out is None:
            out = torch.empty_like(x)
        _check(lib().hmx_reorderVectors(ctx.handle, self.handle, k, 1 if to_original else 0, _ptr(x), _ptr(out)), ctx)
        return out


class Structure(_Obj):
    _free = "hmx_freeHMatrixStructure"

    def num_levels(self):
        L = _i32()
        _check(lib().hmx_structure_info(self.handle, ctypes.byref(L), None))
        return L.value

    def num_tiles(self):
        L = self.num_levels()
        nt = np.zeros(L, dtype=np.int32)
        _check(lib().hmx_structure_info(self.handle, None, _np(nt)))
        return nt

    def tiles(self, level):
        nt = self.num_tiles()[level - 1]
        out = np.zeros(max(nt, 1), dtype=np.int32)
        _check(lib().hmx_structure_tiles(self.handle, level, _np(out)))
        return out[:nt]

    def jds(self, level):
        nt = self.num_tiles()[level - 1]
        nr = 1 << level
        ri = np.zeros(nr, dtype=np.int32)
        ci = np.zeros(max(nt, 1), dtype=np.int32)
        tp = np.zeros(max(nt, 1), dtype=np.int32)
        ip = np.zeros(nr + 1, dtype=np.int32)
        nd = _i32()
        _check(lib().hmx_structure_jds(self.handle, level, _np(ri), _np(ci), _np(tp), _np(ip), ctypes.byref(nd)))
        return ri, ci[:nt], tp[:nt], ip[:nd.value + 1]


class TLRMatrix(_Obj):
    _free = "hmx_freeMatrix"

    def info(self):
        S, b, o, rs, ch = _u32(), _u32(), _i32(), _u64(), _u64()
        _check(lib().hmx_tlr_info(self.handle, ctypes.byref(S), ctypes.byref(b), ctypes.byref(o), ctypes.byref(rs),
                                  ctypes.byref(ch)))
        return dict(numSegments=S.value, segmentSize=b.value, ordering=o.value, rankSum=rs.value, capHits=ch.value)

    def copy(self, ctx, factors=True):
        i = self.info()
        S, b, rs = i["numSegments"], i["segmentSize"], i["rankSum"]
        ranks = np.zeros(S * S, dtype=np.int32)
        offs = np.zeros(S * S + 1, dtype=np.int64)
        U = np.zeros(max(rs * b, 1)) if factors else None
        V = np.zeros(max(rs * b, 1)) if factors else None
        D = np.zeros(S * b * b)
        _check(lib().hmx_tlr_copy(ctx.handle, self.handle, _np(ranks), _np(offs), _np(U) if factors else None,
                                  _np(V) if factors else None, _np(D)), ctx)
        return ranks, offs, U, V, D


class HMatrix(_Obj):
    _free = "hmx_freeHMatrix"

    def __init__(self, handle, num_levels, n, b):
        super().__init__(handle)
        self.num_levels, self.n, self.b = num_levels, n, b

    def build_info(self):
        """Direct assembly: sampler flops, kernel evaluations, sampler kernel ms."""
        f, e, ms = _f64(), _f64(), _f64()
        _check(lib().hmx_hmatrix_build_info(self.handle, ctypes.byref(f), ctypes.byref(e), ctypes.byref(ms)))
        return dict(samplerFlops=f.value, kernelEvals=e.value, samplerMs=ms.value)

    def level_info(self, level):
        nt, rs, mr, tol, ch = _i32(), _i64(), _i32(), _f32(), _i32()
        _check(lib().hmx_hmatrix_level_info(self.handle, level, ctypes.byref(nt), ctypes.byref(rs), ctypes.byref(mr),
                                            ctypes.byref(tol), ctypes.byref(ch)))
        return dict(numTiles=nt.value, rankSum=rs.value, maxRank=mr.value, tol=tol.value, capHits=ch.value)

    def level(self, ctx, level):
        """(ranks, [U_t], [V_t]) with U_t, V_t of shape m x r_t."""
        li = self.level_info(level)
        m = (1 << (self.num_levels - 1 - level)) * self.b
        ranks = np.zeros(max(li["numTiles"], 1), dtype=np.int32)
        U = np.zeros(max(li["rankSum"] * m, 1))
        V = np.zeros(max(li["rankSum"] * m, 1))
        _check(lib().hmx_hmatrix_level_copy(ctx.handle, self.handle, level, _np(ranks), _np(U), _np(V)), ctx)
        ranks = ranks[:li["numTiles"]]
        Us, Vs, o = [], [], 0
        for r in ranks:
            Us.append(U[o * m:(o + r) * m].reshape(r, m).T)
            Vs.append(V[o * m:(o + r) * m].reshape(r, m).T)
            o += r
        return ranks, Us, Vs

    def level_ranks(self, ctx, level):
        """Tile ranks of one level (host array, no factor copy)."""
        li = self.level_info(level)
        ranks = np.zeros(max(li["numTiles"], 1), dtype=np.int32)
        _check(lib().hmx_hmatrix_level_copy(ctx.handle, self.handle, level, _np(ranks), None, None), ctx)
        return ranks[:li["numTiles"]]

    def level_device(self, ctx, level):
        """Device copies (torch float64, flat) of one level's U and V."""
        import torch
        li = self.level_info(level)
        m = (1 << (self.num_levels - 1 - level)) * self.b
        u = torch.empty(max(li["rankSum"] * m, 1), dtype=torch.float64, device="cuda")
        v = torch.empty_like(u)
        _check(lib().hmx_hmatrix_level_get(ctx.handle, self.handle, level, _ptr(u), _ptr(v)), ctx)
        return u, v

    def level_raw(self, ctx, level):
        """(ranks, exclusive column offsets, flat U, flat V) of one level."""
        li = self.level_info(level)
        m = (1 << (self.num_levels - 1 - level)) * self.b
        ranks = np.zeros(max(li["numTiles"], 1), dtype=np.int32)
        U = np.zeros(max(li["rankSum"] * m, 1))
        V = np.zeros(max(li["rankSum"] * m, 1))
        _check(lib().hmx_hmatrix_level_copy(ctx.handle, self.handle, level, _np(ranks), _np(U), _np(V)), ctx)
        ranks = ranks[:li["numTiles"]]
        offs = np.zeros(len(ranks) + 1, dtype=np.int64)
        offs[1:] = np.cumsum(ranks)
        return ranks, offs, U, V

    def diag(self, ctx, S):
        D = np.zeros(S * self.b * self.b)
        _check(lib().hmx_hmatrix_diag_copy(ctx.handle, self.handle, _np(D)), ctx)
        return D

    def save(self, ctx, path):
        _check(lib().hmx_hmatrix_save(ctx.handle, self.handle, os.fsencode(path)), ctx)

    @staticmethod
    def load(ctx, path):
        h = ctypes.c_void_p()
        _check(lib().hmx_hmatrix_load(ctx.handle, os.fsencode(path), ctypes.byref(h)), ctx)
        # magic, version, n, d, b, S, numLevels
        with open(path, "rb") as f:
            hdr = np.frombuffer(f.read(28), dtype=np.uint32)
        return HMatrix(h, int(hdr[6]), int(hdr[2]), int(hdr[4]))

    def set_plan_options(self, dense_near=0, direct_z=0, locality=0):
        _check(lib().hmx_hmatrix_set_plan_options(self.handle, int(dense_near), int(direct_z), int(locality)))

    def set_fp64_layout(self, layout):
        """FP64 matrix-core plan factors: 0 copy when it fits (default), 1 in place, 2 copy."""
        _check(lib().hmx_hmatrix_set_fp64_layout(self.handle, int(layout)))

    def plan_info(self, plan):
        """plan 0: FP64 matrix-core plan, 1: mixed plan, 2: FP64 streaming plan (bytes per call or pass,
        dense blocks, low-rank columns)."""
        by, nd, nc = _u64(), _i64(), _i64()
        _check(lib().hmx_hmatrix_plan_info(self.handle, plan, ctypes.byref(by), ctypes.byref(nd), ctypes.byref(nc)))
        return dict(bytes=by.value, denseBlocks=nd.value, lowRankColumns=nc.value)

    def plan_phase_bytes(self, plan):
        """The plan's bytes by phase: (V factors, U factors + dense/near blocks or points + scales)."""
        p1, p2 = _u64(), _u64()
        _check(lib().hmx_hmatrix_plan_phase_bytes(self.handle, plan, ctypes.byref(p1), ctypes.byref(p2)))
        return p1.value, p2.value

    def bytes(self):
        lr, dn, fe = _u64(), _u64(), _u64()
        _check(lib().hmx_hmatrix_bytes(self.handle, ctypes.byref(lr), ctypes.byref(dn), ctypes.byref(fe)))
        return dict(lowRankBytes=lr.value, denseBytes=dn.value, factorEntries=fe.value)


class Batch(_Obj):
    """Many H-matrices applied as one block-diagonal operator (one launch per
    phase, optional hipGraph replay).  Keeps the H-matrices alive."""
    _free = "hmx_batch_free"

    def __init__(self, ctx, hmatrices):
        arr = (ctypes.c_void_p * len(hmatrices))(*[h.handle.value for h in hmatrices])
        h = ctypes.c_void_p()
        _check(lib().hmx_batch_create(ctx.handle, len(hmatrices), arr, ctypes.byref(h)), ctx)
        super().__init__(h)
        self.ctx = ctx
        self._keep = list(hmatrices)
        self.sizes = [hm.n for hm in hmatrices]
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)

    def info(self):
        n, l, c = _i64(), _i64(), _i32()
        _check(lib().hmx_batch_info(self.handle, ctypes.byref(n), ctypes.byref(l), ctypes.byref(c)))
        return dict(totalPoints=n.value, totalLeaves=l.value, chunk=c.value)

    def matvec(self, x, y=None, graph=False):
        """x: (k, total) or (total,) float64 CUDA tensor."""
        import torch
        k = 1 if x.dim() == 1 else x.shape[0]
        if y is None:
            y = torch.empty_like(x)
        _check(lib().hmx_batch_matvec(self.ctx.handle, self.handle, k, _ptr(x), _ptr(y), int(graph)), self.ctx)
        return y

    def tune(self, k):
        c, ms = _i32(), _f32()
        _check(lib().hmx_batch_tune(self.ctx.handle, self.handle, k, ctypes.byref(c), ctypes.byref(ms)), self.ctx)
        return c.value, ms.value


# ---- reference-named entry points --------------------------------------------
def generateDataset(ctx, n, d, seed, points):
    _check(lib().hmx_generateDataset(ctx.handle, n, d, seed, _ptr(points)), ctx)


def generateRandomVector(ctx, k, n, seed, out):
    _check(lib().hmx_generateRandomVector(ctx.handle, k, n, seed, _ptr(out)), ctx)


def constructKDTree(ctx, n, d, b, points, divMethod=FULL_TREE):
    h = ctypes.c_void_p()
    _check(lib().hmx_constructKDTree(ctx.handle, n, d, b, divMethod, _ptr(points), ctypes.byref(h)), ctx)
    return KDTree(h)


def generateDatasetBatch(ctx, ns, ds, seeds, points):
    """Points of many clouds in one launch; cloud k as generateDataset(ns[k], ds[k], seeds[k])."""
    c = len(ns)
    na, da = (_u32 * c)(*ns), (_u32 * c)(*ds)
    sa = (_u64 * c)(*seeds)
    pa = (_vp * c)(*[_ptr(p) for p in points])
    _check(lib().hmx_generateDatasetBatch(ctx.handle, c, na, da, sa, pa), ctx)


def constructKDTreeBatch(ctx, ns, ds, b, points, divMethod=FULL_TREE):
    """KD-trees of many clouds, one forest per point dimension; tree k equals constructKDTree on cloud k."""
    c = len(ns)
    na, da = (_u32 * c)(*ns), (_u32 * c)(*ds)
    pa = (_vp * c)(*[_ptr(p) for p in points])
    ha = (_vp * c)()
    _check(lib().hmx_constructKDTreeBatch(ctx.handle, c, na, da, b, divMethod, pa, ha), ctx)
    return [KDTree(ctypes.c_void_p(ha[k])) for k in range(c)]


def constructHMatrixStructure(ctx, n, d, b, adm, eta, tree, device=False):
    """device=True: the level-synchronous pass on the GPU (the same lists, bit for bit)."""
    h = ctypes.c_void_p()
    a = ADMISSIBILITY[adm] if isinstance(adm, str) else adm
    f = lib().hmx_constructHMatrixStructureDevice if device else lib().hmx_constructHMatrixStructure
    _check(f(ctx.handle, n, d, b, a, eta, tree.handle, ctypes.byref(h)), ctx)
    return Structure(h)


def structure_from_boxes(n, d, b, adm, eta, bbmin, bbmax):
    h = ctypes.c_void_p()
    a = ADMISSIBILITY[adm] if isinstance(adm, str) else adm
    bmin = np.ascontiguousarray(bbmin, dtype=np.float64)
    bmax = np.ascontiguousarray(bbmax, dtype=np.float64)
    _check(lib().hmx_structure_from_boxes(n, d, b, a, eta, _np(bmin), _np(bmax), ctypes.byref(h)))
    return Structure(h)


def createColumnMajorLRMatrix(ctx, cfg, tree, points):
    h = ctypes.c_void_p()
    rs = _u64()
    _check(lib().hmx_createColumnMajorLRMatrix(ctx.handle, ctypes.byref(cfg), tree.handle, _ptr(points),
                                               ctypes.byref(h), ctypes.byref(rs)), ctx)
    return TLRMatrix(h), rs.value


def createColumnMajorLRMatrixPiece(ctx, cfg, tree, points, pieces, row_piece, col_piece):
    """Piece (row_piece, col_piece) of a pieces x pieces partition of the TLR matrix."""
    h = ctypes.c_void_p()
    rs = _u64()
    _check(lib().hmx_createColumnMajorLRMatrixPiece(ctx.handle, ctypes.byref(cfg), tree.handle, _ptr(points), pieces,
                                                    row_piece, col_piece, ctypes.byref(h), ctypes.byref(rs)), ctx)
    return TLRMatrix(h), rs.value


def convertColumnMajorToMorton(ctx, tlr):
    h = ctypes.c_void_p()
    _check(lib().hmx_convertColumnMajorToMorton(ctx.handle, tlr.handle, ctypes.byref(h)), ctx)
    return TLRMatrix(h)


def generateMaxRanks(numLevels, b):
    out = np.zeros(max(numLevels - 2, 1), dtype=np.uint32)
    lib().hmx_generateMaxRanks(numLevels, b, _np(out))
    return out[:max(numLevels - 2, 0)]


def generateHMatrixFromStruct(ctx, cfg, tree, structure, morton_tlr):
    h = ctypes.c_void_p()
    _check(lib().hmx_generateHMatrixFromStruct(ctx.handle, ctypes.byref(cfg), tree.handle, structure.handle,
                                               morton_tlr.handle, ctypes.byref(h)), ctx)
    return HMatrix(h, structure.num_levels(), cfg.numberOfInputPoints, cfg.bucketSize)


def generateHMatrixDirect(ctx, cfg, tree, structure, points):
    """Whole H-matrix without the S x S TLR matrix (diagonal and leaf level as
    the TLR path produces them, upper levels sampled from the kernel)."""
    h = ctypes.c_void_p()
    _check(lib().hmx_generateHMatrixDirect(ctx.handle, ctypes.byref(cfg), tree.handle, structure.handle,
                                           _ptr(points), ctypes.byref(h)), ctx)
    return HMatrix(h, structure.num_levels(), cfg.numberOfInputPoints, cfg.bucketSize)


def generateHMatrixDirectBatch(ctx, cfgs, trees, structures, points):
    """Direct assembly of many H-matrices, upper levels compressed together."""
    n = len(cfgs)
    carr = (Config * n)(*cfgs)
    tarr = (ctypes.c_void_p * n)(*[t.handle.value for t in trees])
    sarr = (ctypes.c_void_p * n)(*[s.handle.value for s in structures])
    parr = (ctypes.c_void_p * n)(*[p.data_ptr() for p in points])
    harr = (ctypes.c_void_p * n)()
    _check(lib().hmx_generateHMatrixDirectBatch(ctx.handle, n, carr, tarr, sarr, parr, harr), ctx)
    return [HMatrix(ctypes.c_void_p(harr[i]), structures[i].num_levels(), cfgs[i].numberOfInputPoints,
                    cfgs[i].bucketSize) for i in range(n)]


class BatchPipeline:
    """Many point clouds through the whole path: the datasets in one launch,
    the KD-trees as one forest per point dimension, the structures on the
    host, then one batched direct assembly and one batched matvec operator.
    clouds: list of (n, d, ell)."""

    def __init__(self, clouds, b=32, tol=1e-6, adm="weak", max_rank_policy=MAXRANK_FULL, seed=1234, ctx=None,
                 points=None, symmetric=False):
        import time
        import torch
        self.ctx = ctx or Context()
        self.clouds = list(clouds)
        t = {}
        t0 = time.perf_counter()
        ns = [n for n, _, _ in self.clouds]
        ds = [d for _, d, _ in self.clouds]
        if points is None:
            self.points = [torch.empty(n * d, dtype=torch.float64, device="cuda") for n, d in zip(ns, ds)]
            generateDatasetBatch(self.ctx, ns, ds, [seed + i for i in range(len(ns))], self.points)
        else:
            self.points = list(points)
        self.trees = constructKDTreeBatch(self.ctx, ns, ds, b, self.points)
        self.structures, self.cfgs = [], []
        for i, (n, d, ell) in enumerate(self.clouds):
            self.structures.append(constructHMatrixStructure(self.ctx, n, d, b, adm, 1.0, self.trees[i]))
            self.cfgs.append(default_config(numberOfInputPoints=n, dimensionOfInputPoints=d, bucketSize=b,
                                            lowestLevelTolerance=tol, admissibilityCondition=ADMISSIBILITY.get(adm, adm),
                                            correlationLength=ell, maxRankPolicy=max_rank_policy, seed=seed,
                                            symmetricTiles=int(symmetric)))
        torch.cuda.synchronize()
        t["TREES_AND_STRUCTURES"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        self.H = generateHMatrixDirectBatch(self.ctx, self.cfgs, self.trees, self.structures, self.points)
        torch.cuda.synchronize()
        t["HMATRIX_BATCH"] = time.perf_counter() - t0
        self.batch = Batch(self.ctx, self.H)
        self.timings = t

    def matvec(self, x, y=None, graph=False):
        return self.batch.matvec(x, y, graph)


def HMatrixVecMult(ctx, H, k, x, y, transpose=False):
    f = lib().hmx_HMatrixVecMultTranspose if transpose else lib().hmx_HMatrixVecMult
    _check(f(ctx.handle, H.handle, k, _ptr(x), _ptr(y)), ctx)


def matvec_profile(ctx, H, k, x, y, iters=20):
    ms = np.zeros(4, dtype=np.float32)
    _check(lib().hmx_matvec_profile(ctx.handle, H.handle, k, _ptr(x), _ptr(y), iters, _np(ms)), ctx)
    return ms


def matvec_profile_multi(ctx, H, k, x, y, iters=20, mixed=False):
    ms = np.zeros(4, dtype=np.float32)
    _check(lib().hmx_matvec_profile_multi(ctx.handle, H.handle, k, _ptr(x), _ptr(y), iters, int(mixed), _np(ms)), ctx)
    return ms


def make_mixed(ctx, H, target=3e-6):
    _check(lib().hmx_hmatrix_make_mixed(ctx.handle, H.handle, target), ctx)
    hi, lo, fr, le, by = _i64(), _i64(), _f64(), _f64(), _u64()
    _check(lib().hmx_hmatrix_mixed_info(H.handle, ctypes.byref(hi), ctypes.byref(lo), ctypes.byref(fr),
                                        ctypes.byref(le), ctypes.byref(by)))
    return dict(hiCols=hi.value, loCols=lo.value, frobH2=fr.value, loEnergy=le.value, bytes=by.value)


def HMatrixVecMultMixed(ctx, H, k, x, y):
    _check(lib().hmx_HMatrixVecMultMixed(ctx.handle, H.handle, k, _ptr(x), _ptr(y)), ctx)


def fp64_peak_probe(ctx, unit=0, iters=20000):
    """Measured FP64 TFLOP/s of the matrix cores (unit 0) or vector FMA (unit 1)."""
    t = ctypes.c_float(0)
    _check(lib().hmx_fp64_peak_probe(ctx.handle, unit, iters, ctypes.byref(t)), ctx)
    return t.value


def stream_read_probe(ctx, buf, iters=5):
    """GB/s of a pass that reads the device tensor buf once (16-byte
    non-temporal loads): the measured streaming-read ceiling."""
    ms = ctypes.c_float(0)
    nbytes = buf.numel() * buf.element_size()
    _check(lib().hmx_stream_read_probe(ctx.handle, _ptr(buf), nbytes, iters, ctypes.byref(ms)), ctx)
    return nbytes / (ms.value * 1e-3) / 1e9


def selftest_mfma(ctx, type_, A, B, D):
    _check(lib().hmx_selftest_mfma(ctx.handle, type_, _ptr(A), _ptr(B), _ptr(D)), ctx)


def denseRowsXVector(ctx, n, d, points, perm_ptr, ell, rows, k, X, Y):
    _check(lib().hmx_denseRowsXVector(ctx.handle, n, d, _ptr(points), ctypes.c_void_p(perm_ptr), ell, rows.numel(),
                                      _ptr(rows), k, _ptr(X), _ptr(Y)), ctx)


class Pipeline:
    """The reference driver sequence (src/main.cu:46-173) on the GPU."""

    def __init__(self, n, d, b, tol, adm="weak", eta=1.0, ell=0.0, max_rank_policy=MAXRANK_REF, seed=1234,
                 points=None, ctx=None, timings=None, direct=False, symmetric=False):
        import time
        import torch
        self.ctx = ctx or Context()
        self.n, self.d, self.b = n, d, b
        self.cfg = default_config(numberOfInputPoints=n, dimensionOfInputPoints=d, bucketSize=b,
                                  lowestLevelTolerance=tol, admissibilityCondition=ADMISSIBILITY.get(adm, adm),
                                  eta=eta, correlationLength=ell, maxRankPolicy=max_rank_policy, seed=seed,
                                  symmetricTiles=int(symmetric))
        self.ell = ell if ell > 0 else (0.2 if d == 3 else 0.1)
        t = {}

        def mark(name, t0):
            torch.cuda.synchronize()
            t[name] = time.perf_counter() - t0
            return time.perf_counter()

        t0 = time.perf_counter()
        if points is None:
            self.points = torch.empty(n * d, dtype=torch.float64, device="cuda")
            generateDataset(self.ctx, n, d, seed, self.points)
        else:
            self.points = points
        t0 = mark("GENERATE_DATASET", t0)
        self.tree = constructKDTree(self.ctx, n, d, b, self.points)
        t0 = mark("KDTREE", t0)
        self.structure = constructHMatrixStructure(self.ctx, n, d, b, adm, eta, self.tree)
        t0 = mark("STRUCTURE", t0)
        self.direct = direct
        if direct:
            # no S x S TLR matrix: diagonal + leaf tiles as the TLR path, upper levels from the kernel
            self.tlr_cm = self.tlr_mo = None
            self.rank_sum = None
            self.H = generateHMatrixDirect(self.ctx, self.cfg, self.tree, self.structure, self.points)
            t0 = mark("HMATRIX_DIRECT", t0)
        else:
            self.tlr_cm, self.rank_sum = createColumnMajorLRMatrix(self.ctx, self.cfg, self.tree, self.points)
            t0 = mark("TLR_MATRIX", t0)
            self.tlr_mo = convertColumnMajorToMorton(self.ctx, self.tlr_cm)
            t0 = mark("CMTOMO", t0)
            self.H = generateHMatrixFromStruct(self.ctx, self.cfg, self.tree, self.structure, self.tlr_mo)
            t0 = mark("HMATRIX", t0)
        self.timings = t
        if timings is not None:
            timings.update(t)

    def make_mixed(self, target=3e-6):
        self.mixed_info = make_mixed(self.ctx, self.H, target)
        return self.mixed_info

    def matvec_mixed(self, x, y=None):
        """x: (n,) or (k, n) float32 contiguous (an n x k column-major block)."""
        import torch
        k = 1 if x.dim() == 1 else x.shape[0]
        if y is None:
            y = torch.empty_like(x)
        HMatrixVecMultMixed(self.ctx, self.H, k, x, y)
        return y

    def matvec(self, x, y=None, transpose=False, original_order=False):
        """y = H x (or H^T x).  x: (n,) or (k, n) contiguous, i.e. an n x k
        column-major block, in KD-tree order like the reference; with
        original_order=True x and y are in the caller's point order and are
        permuted through the tree on the way in and out."""
        import torch
        k = 1 if x.dim() == 1 else x.shape[0]
        if y is None:
            y = torch.empty_like(x)
        if original_order:
            xk = self.tree.reorder(self.ctx, x, False)
            yk = torch.empty_like(xk)
            HMatrixVecMult(self.ctx, self.H, k, xk, yk, transpose)
            self.tree.reorder(self.ctx, yk, True, out=y)
            return y
        HMatrixVecMult(self.ctx, self.H, k, x, y, transpose)
        return y
