"""The host-side C and C++ code under AddressSanitizer and
UndefinedBehaviorSanitizer (gcc / g++ builds, CPU only):

  * the CPU oracle (oracle/rimp2_oracle.c, oracle/mlp_oracle.c), every entry
    point on small ragged shapes with exactly sized buffers
    (tests/cpu/oracle_san_main.c);
  * the grouped metric pipeline's plan (rimp2_gpu_amd/csrc/metric_group.h) in
    its host emulation (tests/cpu/metric_group_emu.cpp, fragments up to 129
    auxiliary functions: every panel and doubling case).

A sanitizer report (out-of-bounds access, use after free, leak, undefined
arithmetic) aborts the program, so a zero exit and the "OK" line are the
pass condition."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1",
           OMP_NUM_THREADS="4")


def _run(exe, *args):
    r = subprocess.run([str(exe), *args], capture_output=True, text=True, timeout=300, env=ENV)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert r.stdout.startswith("OK"), r.stdout
    assert "runtime error" not in r.stderr, r.stderr[-4000:]


def test_oracle_under_address_and_undefined_sanitizers(tmp_path):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not found")
    exe = tmp_path / "oracle_san"
    srcs = [os.path.join(ROOT, "tests", "cpu", "oracle_san_main.c"),
            os.path.join(ROOT, "oracle", "rimp2_oracle.c"), os.path.join(ROOT, "oracle", "mlp_oracle.c")]
    subprocess.run([gcc, *SAN, "-std=c11", "-fopenmp", "-Wall", "-o", str(exe), *srcs, "-lm"], check=True)
    _run(exe)


def test_grouped_metric_plan_under_address_and_undefined_sanitizers(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    exe = tmp_path / "metric_group_emu_san"
    src = os.path.join(ROOT, "tests", "cpu", "metric_group_emu.cpp")
    subprocess.run([gxx, *SAN, "-std=c++17", "-Wall", "-Wextra", "-o", str(exe), src], check=True)
    _run(exe, "--small")
