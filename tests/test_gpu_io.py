"""Binary H-matrix files: a saved and reloaded H-matrix has identical levels,
diagonal and structure, and its H-matvec equals the original's bit for bit;
damaged files are refused."""
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


@pytest.mark.parametrize("n,d,b,direct", [(3000, 2, 32, False), (4096, 3, 64, True)])
def test_save_load_roundtrip(torch_cuda, tmp_path, n, d, b, direct):
    ctx = hmx.Context()
    p = O.generate_points(n, d, 9)
    pipe = hmx.Pipeline(n, d, b, 1e-5, adm="maxDiam", max_rank_policy=hmx.MAXRANK_FULL, points=to_dev(torch_cuda, p),
                        ctx=ctx, direct=direct)
    path = tmp_path / "h.hmx"
    pipe.H.save(ctx, str(path))
    H2 = hmx.HMatrix.load(ctx, str(path))
    assert (H2.n, H2.b, H2.num_levels) == (n, b, pipe.H.num_levels)
    S = pipe.tree.info()[0]
    assert np.array_equal(H2.diag(ctx, S), pipe.H.diag(ctx, S))
    for l in range(1, pipe.H.num_levels):
        assert pipe.H.level_info(l) == H2.level_info(l)
        r1, o1, U1, V1 = pipe.H.level_raw(ctx, l)
        r2, o2, U2, V2 = H2.level_raw(ctx, l)
        assert np.array_equal(r1, r2) and np.array_equal(U1, U2) and np.array_equal(V1, V2)
    for k in (1, 16):
        x = torch_cuda.randn(k, n, dtype=torch_cuda.float64, device="cuda")
        y1 = pipe.matvec(x)
        y2 = torch_cuda.empty_like(x)
        hmx.HMatrixVecMult(ctx, H2, k, x, y2)
        assert torch_cuda.equal(y1, y2)
    # damaged files
    raw = path.read_bytes()
    bad = tmp_path / "bad.hmx"
    bad.write_bytes(raw[: len(raw) // 2])
    with pytest.raises(hmx.HmxError) as ei:
        hmx.HMatrix.load(ctx, str(bad))
    assert ei.value.status == hmx.HMX_EINVAL
    bad.write_bytes(b"XXXX" + raw[4:])
    with pytest.raises(hmx.HmxError) as ei:
        hmx.HMatrix.load(ctx, str(bad))
    assert ei.value.status == hmx.HMX_EINVAL
