"""The oracle built with AddressSanitizer and UndefinedBehaviorSanitizer
(leak checks on), driven through every entry point on small PML, Bloch,
nonuniform, component-last and boundft = HH grids by
oracle/sanitize_driver.c.  Host code only; needs gcc with libasan."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_oracle_under_asan_ubsan():
    r = subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "sanitize driver ok" in r.stdout
