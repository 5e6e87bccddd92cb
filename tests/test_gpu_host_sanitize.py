"""The library's host code under AddressSanitizer and
UndefinedBehaviorSanitizer (host side only; the GPU code is the ordinary
build): tests/cpp/host_sanitize drives the planners, device tables,
host staging, the serving arena and scheduler (several requests, an injected
failure, a refused budget, a mixed program), the DLTC and RRR paths, the
digit splits and the UHF CCSD step through the C ABI on small problems and
checks each result.  A sanitizer report ends it with a non-zero status."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_code_under_asan_and_ubsan(dev):
    exe = os.path.join(ROOT, "tests", "cpp", "host_sanitize")
    assert os.path.exists(exe), "build first (dtcf_amd/csrc/Makefile)"
    env = dict(os.environ)
    # leaks: the HIP runtime keeps allocations for the process's life
    env["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1:abort_on_error=0:exitcode=23"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1:exitcode=24"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240, env=env)
    print(r.stdout[-4000:], r.stderr[-8000:], flush=True)
    assert r.returncode == 0, f"exit status {r.returncode}"
    assert "host sanitize: ok" in r.stdout
