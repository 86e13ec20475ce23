"""The host C++ (file parsers, JSON, .marvin, net setup, argument checks) and
the C oracle under AddressSanitizer + UndefinedBehaviorSanitizer: builds the
sanitizer variants (`make asan` in the package and in oracle/) and runs the
CPU tests that drive that code -- formats, the C ABI, the malformed-input
mutations and the oracle tests -- in a child process with clang's ASan
runtime preloaded.  Any out-of-bounds access, use after free, overflow or
undefined shift aborts the child and fails this test."""
import glob
import os
import subprocess
import sys

import pytest

from conftest import PKG, REPO

FILES = ["tests/test_formats.py", "tests/test_capi.py", "tests/test_malformed_inputs.py",
         "tests/test_oracle_tdf.py", "tests/test_oracle_tsdf.py",
         "tests/test_oracle_registration.py", "tests/test_oracle_conv.py"]


def _asan_runtime():
    hits = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return hits[-1] if hits else None


@pytest.mark.timeout(1500)
def test_cpu_suite_under_asan_ubsan(tmp_path):
    rt = _asan_runtime()
    if rt is None:
        pytest.skip("clang's ASan runtime is not installed")
    jobs = str(min(8, os.cpu_count() or 2))
    subprocess.run(["make", "-s", "-j", jobs, "-C", PKG, "asan"], check=True)
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "asan"], check=True)
    lib = os.path.join(PKG, "lib3dmatch_amd_asan.so")
    orc = os.path.join(REPO, "oracle", "liboracle_asan.so")
    log = str(tmp_path / "san")
    env = dict(os.environ, LD_PRELOAD=rt, TDM_LIB=lib, TDM_ORACLE_LIB=orc,
               ASAN_OPTIONS=f"detect_leaks=0:halt_on_error=1:log_path={log}",
               UBSAN_OPTIONS=f"halt_on_error=1:print_stacktrace=1:log_path={log}")
    # the child loads the sanitizer builds, not the ordinary ones
    probe = ("import sys; sys.path[:0] = ['3dmatch-toolbox_amd', 'oracle'];"
             "import ctypes, tdm, oracle;"
             "assert tdm.lib()._name.endswith('_asan.so'), tdm.lib()._name;"
             "assert oracle.lib()._name.endswith('_asan.so'), oracle.lib()._name;"
             "ctypes.CDLL(None).__asan_report_load4")
    r = subprocess.run([sys.executable, "-c", probe], cwd=REPO, env=env, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu",
                        "-p", "no:cacheprovider", *FILES], cwd=REPO, env=env,
                       capture_output=True, text=True, timeout=1400)
    reports = "".join(open(p).read() for p in glob.glob(log + "*"))
    assert r.returncode == 0 and not reports, r.stdout[-4000:] + r.stderr[-2000:] + reports[-6000:]
