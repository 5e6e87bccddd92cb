"""The C++ drop-in symbols, called from C++ the way the reference calls them
(tests/cpp/dropin_selftest.cpp), and the header drop-ins expanded in one
translation unit against stand-in reference declarations
(tests/cpp/dropin_macros.cpp), both built with the libraries."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_dropin_selftest(dev):
    exe = os.path.join(ROOT, "tests", "cpp", "dropin_selftest")
    assert os.path.exists(exe), "build first (dtcf_amd/csrc/Makefile)"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and "PASS" in r.stdout


def test_dropin_header_macros(dev):
    exe = os.path.join(ROOT, "tests", "cpp", "dropin_macros")
    assert os.path.exists(exe), "build first (dtcf_amd/csrc/Makefile)"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and "PASS" in r.stdout
