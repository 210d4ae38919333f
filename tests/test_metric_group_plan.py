"""The grouped metric pipeline's plan (rimp2_gpu_amd/csrc/metric_group.h),
run on the CPU by a host emulation (tests/cpu/metric_group_emu.cpp): every
access of every planned launch is bounds-checked, no tile of a launch may read
what another tile of it writes, and L, L^{-1} and B = L^{-1} X of fragments
from 1 to 320 auxiliary functions (below, at and above the 64-wide panel,
full and ragged doubling groups) must match a direct Cholesky, forward
substitution and product to 1e-12; the status words of a bad pivot, a NaN
input and a denominator >= 0 must flag exactly their fragment.  The plan
without the transform (the exact-slice modes, which transform on int8 MFMA
after it) is run too: L and L^{-1} as above, and no packing or transform step."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_grouped_metric_plan_emulated(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    exe = tmp_path / "metric_group_emu"
    src = os.path.join(ROOT, "tests", "cpu", "metric_group_emu.cpp")
    subprocess.run([gxx, "-O2", "-std=c++17", "-Wall", "-Wextra", "-o", str(exe), src], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("OK"), r.stdout
