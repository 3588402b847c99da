"""Every extern "C" entry point of the oracle, built with AddressSanitizer and UndefinedBehaviorSanitizer
(oracle/sanitize_main.cpp, `make -C oracle sanitize`): out-of-bounds accesses, leaks, signed overflow and the
other undefined behaviour the sanitizers detect abort the run.  CPU only (about 20 s)."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None, reason="needs g++ and make")
def test_oracle_under_address_and_undefined_sanitizers():
    r = subprocess.run(["make", "-C", ORACLE, "sanitize"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert "every oracle entry point ran" in r.stdout
