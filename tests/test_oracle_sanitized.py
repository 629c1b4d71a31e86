"""The oracle built with AddressSanitizer and UndefinedBehaviorSanitizer
(`make -C oracle selftest`): its driver runs every model in 2D and 3D, with
and without walls, through fast marching, the iteration (first and second
order, factored), the tile solver, the flow export, the adjoint and the TTI
factorization, and must finish without a sanitizer report."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(os.path.dirname(HERE), "oracle")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ with the sanitizer runtimes")
def test_oracle_selftest_under_sanitizers():
    res = subprocess.run(["make", "-s", "-C", ORACLE, "selftest"], stdout=subprocess.PIPE,
                         stderr=subprocess.STDOUT, text=True, timeout=600)
    out = res.stdout
    assert res.returncode == 0, out[-4000:]
    assert "selftest: ok" in out
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-4000:]
