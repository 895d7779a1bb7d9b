"""Error paths, forced from the host so every kernel still runs normally.

  * strictRankCap: with the reference's rank caps (b/2 for TLR tiles) a 3D
    max-diameter structure has leaf tiles whose ARA stops at the cap; the
    default build counts them (capHits), the strict build returns
    HMX_ERANKCAP and no object -- for the TLR matrix, the H-matrix from the
    structure and the direct assembly; with uncapped ranks (which end at the
    tile dimension, where the factors are exact) the strict build succeeds.
  * forced device-allocation failures (hmx_debug_set_alloc_fault): a build
    and a first matvec are failed at allocations spread over all the ones they
    make; every failed call returns HMX_ENOMEM, frees what it had built (device
    bytes the library holds back to where they were), and the next call, without the fault,
    produces the same H-matrix and product bit for bit.
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


def test_strict_rank_cap(torch_cuda, ctx):
    n, d, b = 4096, 3, 32
    pts = to_dev(torch_cuda, O.generate_points(n, d, 1234))
    tree = hmx.constructKDTree(ctx, n, d, b, pts)
    st = hmx.constructHMatrixStructure(ctx, n, d, b, "maxDiam", 1.0, tree)
    kw = dict(numberOfInputPoints=n, dimensionOfInputPoints=d, bucketSize=b, lowestLevelTolerance=1e-6,
              admissibilityCondition=hmx.ADMISSIBILITY["maxDiam"], maxRankPolicy=hmx.MAXRANK_REF)
    tlr, _ = hmx.createColumnMajorLRMatrix(ctx, hmx.default_config(**kw), tree, pts)
    hits = tlr.info()["capHits"]
    print(f"TLR tiles at the b/2 cap: {hits}")
    assert hits > 0
    strict = hmx.default_config(strictRankCap=1, **kw)
    with pytest.raises(hmx.HmxError) as ei:
        hmx.createColumnMajorLRMatrix(ctx, strict, tree, pts)
    assert ei.value.status == hmx.HMX_ERANKCAP and "rank cap" in str(ei.value)
    mo = hmx.convertColumnMajorToMorton(ctx, tlr)
    with pytest.raises(hmx.HmxError) as ei:
        hmx.generateHMatrixFromStruct(ctx, strict, tree, st, mo)
    assert ei.value.status == hmx.HMX_ERANKCAP
    with pytest.raises(hmx.HmxError) as ei:
        hmx.generateHMatrixDirect(ctx, strict, tree, st, pts)
    assert ei.value.status == hmx.HMX_ERANKCAP
    # uncapped ranks end at the tile dimension at most (exact factors): no failure
    full = hmx.default_config(strictRankCap=1, **dict(kw, maxRankPolicy=hmx.MAXRANK_FULL))
    H = hmx.generateHMatrixDirect(ctx, full, tree, st, pts)
    assert H.level_info(H.num_levels - 1)["maxRank"] == b


def test_forced_allocation_failures(torch_cuda, ctx):
    torch = torch_cuda
    L = hmx.lib()
    n, d, b = 3000, 2, 32
    pts = to_dev(torch, O.generate_points(n, d, 77))
    cfg = hmx.default_config(numberOfInputPoints=n, dimensionOfInputPoints=d, bucketSize=b,
                             lowestLevelTolerance=1e-6, maxRankPolicy=hmx.MAXRANK_FULL)
    tree = hmx.constructKDTree(ctx, n, d, b, pts)
    st = hmx.constructHMatrixStructure(ctx, n, d, b, "weak", 1.0, tree)
    x = torch.randn(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)

    def build_and_apply():
        H = hmx.generateHMatrixDirect(ctx, cfg, tree, st, pts)
        try:
            hmx.HMatrixVecMult(ctx, H, 1, x, y)
        except BaseException:
            # the raised error's traceback keeps this frame, and H with it, alive
            H.free()
            raise
        return H

    torch.cuda.synchronize()
    L.hmx_debug_set_alloc_fault(-1)
    ref_H = build_and_apply()
    ref_y = y.clone()
    total = L.hmx_debug_alloc_count()
    print(f"a direct build and its first matvec make {total} device allocations")
    assert total > 10
    print("the H-matrix's factor bytes:", ref_H.bytes())
    del ref_H
    torch.cuda.synchronize()
    live0 = L.hmx_debug_live_bytes()
    points = sorted(set(np.linspace(0, total - 1, 12).astype(int).tolist()))
    try:
        for k in points:
            L.hmx_debug_set_alloc_fault(int(k))
            with pytest.raises(hmx.HmxError) as ei:
                build_and_apply()
            assert ei.value.status == hmx.HMX_ENOMEM and "forced" in str(ei.value), k
    finally:
        L.hmx_debug_set_alloc_fault(-1)
    torch.cuda.synchronize()
    live1 = L.hmx_debug_live_bytes()
    print(f"device bytes the library holds before/after the failed calls: {live0} / {live1}")
    assert live1 == live0
    H = build_and_apply()
    assert torch.equal(y, ref_y)
    assert L.hmx_debug_alloc_count() == total
    del H
