"""C-ABI checks that need no GPU: the library loads, exports every entry
point declared in include/hmx.h, and its host-only paths (configuration
defaults, rank schedule, structure + JDS from boxes) agree with the oracle."""
import ctypes
import re

import numpy as np
import pytest

import hmx
import oracle as O


def declared_symbols():
    text = open(hmx.HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hmx_[A-Za-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = hmx.lib()
    names = declared_symbols()
    assert len(names) >= 30
    for name in names:
        assert hasattr(L, name), name
    # the binding declares a signature for every exported entry point
    assert sorted(hmx.exported_symbols()) == names


def test_config_defaults_match_reference_parseargs():
    cfg = hmx.default_config()
    # src/config.h:121-127 and the hard-coded ARA/eta parameters
    assert cfg.divMethod == hmx.FULL_TREE and cfg.admissibilityCondition == hmx.WEAK
    assert (cfg.numberOfInputPoints, cfg.dimensionOfInputPoints, cfg.bucketSize, cfg.vectorWidth) == (1024, 2, 32, 16)
    assert cfg.lowestLevelTolerance == np.float32(1e-5)
    assert (cfg.araBlockSize, cfg.araR, cfg.eta, cfg.seed) == (16, 10, 1.0, 1234)
    # extension: a rank-cap hit is counted, not an error, unless asked for
    assert (cfg.strictRankCap, cfg.symmetricTiles) == (0, 0)


def test_alloc_fault_hook_counts_without_a_device():
    L = hmx.lib()
    L.hmx_debug_set_alloc_fault(-1)
    assert L.hmx_debug_alloc_count() == 0
    L.hmx_debug_set_alloc_fault(5)
    assert L.hmx_debug_alloc_count() == 0
    L.hmx_debug_set_alloc_fault(-1)


def test_generate_max_ranks():
    for L, b in [(3, 32), (8, 32), (12, 32), (15, 64)]:
        got = hmx.generateMaxRanks(L, b)
        exp = np.zeros(L - 2, dtype=np.uint32)
        O.lib().or_generate_max_ranks(L, b, exp.ctypes.data_as(ctypes.c_void_p))
        assert got.tolist() == exp.tolist()


@pytest.mark.parametrize("n,d,adm", [(4096, 2, "weak"), (4096, 2, "boxCenter"), (8192, 3, "boxCenter"),
                                     (5000, 2, "boxCenter"), (4096, 3, "maxDiam")])
def test_structure_from_boxes_matches_oracle(n, d, adm):
    b = 32
    p = O.generate_points(n, d, 77)
    perm, off, bmin, bmax, _ = O.kdtree(p, n, d, b)
    ref = O.structure(n, d, b, adm, 1.0, bmin, bmax)
    s = hmx.structure_from_boxes(n, d, b, adm, 1.0, bmin, bmax)
    assert s.num_levels() == len(ref) + 1
    for l in range(1, s.num_levels()):
        assert np.array_equal(s.tiles(l), ref[l - 1])
        got, exp = s.jds(l), O.jds(l, ref[l - 1])
        for a, e in zip(got, exp):
            assert np.array_equal(a, e)


def test_structure_rejects_bad_arguments():
    bmin = np.zeros((3, 2))
    with pytest.raises(hmx.HmxError) as ei:
        hmx.structure_from_boxes(100, 2, 32, 7, 1.0, bmin, bmin)
    assert ei.value.status == hmx.HMX_EINVAL
    with pytest.raises(hmx.HmxError):
        hmx.structure_from_boxes(100, 2, 24, "weak", 1.0, bmin, bmin)  # upper2(n)/b = 5, not a power of two


def test_null_handles_rejected_without_a_device():
    """Argument checks that run before any device call (CPU-only here)."""
    import ctypes
    L = hmx.lib()
    out = ctypes.c_void_p()
    assert L.hmx_batch_create(None, 1, None, ctypes.byref(out)) == hmx.HMX_EINVAL
    assert L.hmx_batch_info(None, None, None, None) == hmx.HMX_EINVAL
    assert L.hmx_batch_matvec(None, None, 1, None, None, 0) == hmx.HMX_EINVAL
    assert L.hmx_batch_tune(None, None, 1, None, None) == hmx.HMX_EINVAL
    assert L.hmx_batch_free(None) == hmx.HMX_OK
    assert L.hmx_hmatrix_plan_info(None, 0, None, None, None) == hmx.HMX_EINVAL
    assert L.hmx_hmatrix_save(None, None, b"x") == hmx.HMX_EINVAL
    assert L.hmx_hmatrix_load(None, b"x", ctypes.byref(out)) == hmx.HMX_EINVAL
    assert L.hmx_generateHMatrixDirect(None, None, None, None, None, ctypes.byref(out)) == hmx.HMX_EINVAL
