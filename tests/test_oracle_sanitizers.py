"""The oracle's C loops under AddressSanitizer and UndefinedBehaviorSanitizer
(oracle/sanitize_check.c, `make -C oracle sanitize`): every flux method
through the flat-index compflux/vortexforce/innerproduct loops over closed,
periodic and island grids with the derived masks and stencil orders, the
index map and the Laplacian assembly.  Those loops read neighbours without
an upper bound check, as numba does: the stencil orders keep them in bounds,
and the sanitizers check it.  A planted order one step too high at the end
of the array must be reported as a heap overflow."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("make") is None,
                                reason="needs gcc and make")


def _make(target):
    return subprocess.run(["make", "-s", "-C", ORACLE, target], capture_output=True, text=True,
                          timeout=300)


def test_oracle_loops_clean_under_asan_ubsan():
    r = _make("sanitize")
    out = r.stdout + r.stderr
    assert r.returncode == 0, out
    assert "grids clean" in r.stdout
    assert "runtime error" not in out and "AddressSanitizer" not in out


def test_planted_out_of_bounds_order_is_caught():
    r = _make("planted")
    out = r.stdout + r.stderr
    assert r.returncode != 0
    assert "heap-buffer-overflow" in out and "orc_compflux" in out
