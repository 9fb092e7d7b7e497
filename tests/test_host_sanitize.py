"""Host AddressSanitizer + UndefinedBehaviorSanitizer runs of the library's
host-only code: the threaded jump-ahead generator behind the hashing pattern
(src/sparse_helpers.cpp:90-116 draws, split over threads) and the serving
loop.  The driver (ski-lls_amd/drivers/host_sanitize.cpp) links the
instrumented host sources with the regular kernel objects; a sanitizer
finding aborts it with a non-zero status."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "ski-lls_amd", "drivers", "bin", "host_sanitize")

pytestmark = pytest.mark.skipif(not os.path.exists(EXE),
                                reason="host_sanitize not built (make -C ski-lls_amd sanitize)")


def _run(mode, asan_options, timeout):
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = asan_options
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    return subprocess.run([EXE, mode], capture_output=True, text=True, timeout=timeout, env=env)


def test_generator_under_address_and_ub_sanitizers():
    # leak checking on: no HIP runtime in this mode
    r = _run("gen", "detect_leaks=1:halt_on_error=1", 120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 mismatching concurrent patterns, 0 out-of-range" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr


@pytest.mark.gpu
def test_server_under_address_and_ub_sanitizers():
    # the GPU runtime maps memory in the shadow gap and keeps allocations to
    # the end of the process, so those two checks are off
    r = _run("server", "protect_shadow_gap=0:detect_leaks=0:halt_on_error=1", 300)
    print(r.stdout[-2000:], r.stderr[-4000:], flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert ", 0 failures," in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
