"""AddressSanitizer/UBSan build of the device context's host bookkeeping.

The context (``csrc/context.cpp``) and the host set-up (``csrc/precompute.cpp``)
are compiled with g++ against a host-memory stand-in for the HIP runtime
(``tests/hostcheck/``): device allocations are tracked heap blocks, and each
kernel launch is replaced by a check of every range the real kernel indexes
(chunk and tile tables, streams, partial slots, LDS size), run when the launch
runs -- for a captured graph at every replay.  ``harness.cpp`` drives the C-ABI
through the product's launch shapes (64-wide blocks on persistent grids of 8
per CU with 16-step graphs, fused tiles of one and two chunks, 128- to
512-node tiles), ragged batches whose shells halt one by one, dial-from-ansatz
geometry set under a live graph, the device-run relaxation, and a hipMalloc
failure injected at every allocation of ms_create, ms_set_tuning and
ms_set_zero_dial_geometry.  CPU only.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HC = os.path.join(ROOT, "tests", "hostcheck")
CSRC = os.path.join(ROOT, "morphoshell_amd", "csrc")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("hostcheck") / "harness")
    srcs = [os.path.join(HC, f) for f in ("harness.cpp", "mock_runtime.cpp", "mock_kernels.cpp")]
    srcs += [os.path.join(CSRC, "context.cpp"), os.path.join(CSRC, "precompute.cpp")]
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-fopenmp", f"-I{HC}", f"-I{os.path.join(ROOT, 'include')}",
           f"-I{CSRC}", *srcs, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return out


def test_context_bookkeeping_under_sanitizers(harness):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", OMP_NUM_THREADS="2")
    r = subprocess.run([harness], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and "HOSTCHECK OK" in r.stdout, r.stdout[-3000:] + r.stderr[-3000:]
