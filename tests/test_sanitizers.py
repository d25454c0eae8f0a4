"""AddressSanitizer runs on the CPU: the C-ABI library's host code (ASan
build, every descriptor-validation path of every entry point) and the FP64
C port (ASan + UBSan build, two workload cells, both back-substitution
forms, results equal to the plain build's)."""

import glob
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "asan_driver.py")


def _stale(out, sources):
    return not os.path.exists(out) or any(os.path.getmtime(s) > os.path.getmtime(out) for s in sources)


def _run(preload, args):
    env = dict(os.environ, LD_PRELOAD=preload, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1")
    p = subprocess.run([sys.executable, DRIVER, *args], env=env, capture_output=True, text=True, timeout=600)
    text = p.stdout + p.stderr
    assert p.returncode == 0, text[-4000:]
    assert "ERROR: AddressSanitizer" not in text and "runtime error:" not in text, text[-4000:]
    return text


def test_abi_host_code_under_asan():
    rt = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    if not rt:
        pytest.skip("clang ASan runtime not present")
    from braincell_amd import build as B

    srcs = B.sources() + [os.path.join(B.INCLUDE, "braincell_amd.h"), os.path.join(B.CSRC, "bc_channels.h")]
    if _stale(B.asan_library_path(), srcs):
        B.build_asan()
    out = _run(rt[-1], ["abi"])
    assert "abi return codes" in out


def test_c_port_under_asan_and_ubsan(tmp_path):
    libasan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not os.path.exists(libasan):
        pytest.skip("gcc ASan runtime not present")
    from oracle import build_cport as BC

    if _stale(BC.ASAN_OUT, [BC.SRC]):
        BC.build_asan()
    if not os.path.exists(BC.OUT):
        BC.build(verbose=False)
    npz = str(tmp_path / "port.npz")
    _run(libasan, ["port", npz])
    got = np.load(npz)
    # the plain build gives the same traces
    import oracle.cport_run as CR
    from braincell_amd import workloads as W
    from tests.helpers import soma_cv

    models = [W.uneven_tree_cell(i) for i in (27, 49)]
    ref = CR.PortBatch(models, 0.025, [soma_cv(m) for m in models]).run(400, sequential=True, threads=2)
    # (the plain build contracts to FMA instructions; the sanitizer build does not)
    np.testing.assert_allclose(got["seq"], ref, rtol=0, atol=1e-9)
    np.testing.assert_allclose(got["rd"], ref, rtol=0, atol=1e-9)
