"""The host readers and writers (ExoMol .states/.trans host side, HITRAN .par, TIPS tables,
text and opacity-table output, the C ABI's argument checks) run under AddressSanitizer and
UndefinedBehaviorSanitizer: the CPU tests of those files, in a child process that loads the
sanitizer build of the library (make -C gpu_cross_amd/csrc asan) with the clang ASan runtime
preloaded.  Any report aborts the child."""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ASAN_LIB = os.path.join(ROOT, "gpu_cross_amd", "libgx_asan.so")
RUNTIME = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))


@pytest.mark.skipif(not RUNTIME, reason="clang ASan runtime not found")
def test_host_code_under_asan_and_ubsan():
    if not os.path.exists(ASAN_LIB):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "gpu_cross_amd", "csrc"), "asan", "-j8"], check=True)
    nm = subprocess.run(["nm", "-D", ASAN_LIB], capture_output=True, text=True).stdout
    assert "__asan_" in nm and "__ubsan_" in nm
    env = dict(os.environ, GX_LIB=ASAN_LIB, LD_PRELOAD=RUNTIME[-1],
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    files = [os.path.join(ROOT, "tests", f) for f in
             ("test_ingest_cpu.py", "test_output_cpu.py", "test_tips_cpu.py", "test_capi_cpu.py")]
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-m", "not gpu", "-p", "no:cacheprovider",
                        "-k", "not sanitizer", *files], cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=900)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert " passed" in r.stdout
