"""A compiled C++ caller of the C ABI (kaust-hierarchical-matrices_amd/
examples/hmx_main.cpp, built beside libhmx.so): the reference driver's
sequence (src/main.cu:46-186) -- dataset, KD-tree, structure, TLR matrix,
Morton conversion, H-matrix, H-matvec -- or the direct assembly, then the
matvec error on sampled dense rows.  The program exits 0 only when every call
returned HMX_OK and the relative error is below 10 * tolerance."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "kaust-hierarchical-matrices_amd", "lib", "hmx_main")


@pytest.fixture(scope="module")
def gpu_ok():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} not built (make -C kaust-hierarchical-matrices_amd)")


@pytest.mark.parametrize("args", [["16384", "2", "32", "1e-6", "0", "1"],   # TLR path, one RHS
                                  ["16384", "3", "64", "1e-5", "1", "16"],  # direct assembly, matrix cores
                                  ["5001", "3", "32", "1e-6", "0", "5"]],   # partial leaves
                         ids=["tlr-2d", "direct-3d-16rhs", "partial-5rhs"])
def test_cpp_driver(gpu_ok, args):
    r = subprocess.run([EXE] + args, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    m = re.search(r"relErr=([0-9.e+-]+)", r.stdout)
    assert m and float(m.group(1)) < 10 * float(args[3])
