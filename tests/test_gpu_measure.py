"""The benchmark's probe of the int8 MFMA rate the card sustains
(rimp2_gpu_amd/measure/mfma_ceiling.hip), in its three forms (registers only,
LDS reads, the pair kernel's operand flow with DMA): it runs, counts the
operations it issues, and reports a rate at or below the published dense INT8
peak."""
import ctypes
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lds_reads", [0, 1, 2])
def test_sustained_rate_probe(lds_reads):
    f = ctypes.CDLL(os.path.join(ROOT, "rimp2_gpu_amd", "librimp2_measure.so")).rimp2_measure_i8_mfma
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int] + [ctypes.POINTER(ctypes.c_double)] * 3
    s, o, m = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    iters = 100001  # rounded up to an even count
    assert f(0, iters, lds_reads, ctypes.byref(s), ctypes.byref(o), ctypes.byref(m)) == 0
    per_wave = (iters + 1) * 15 * 65536.0
    assert o.value % per_wave == 0 and o.value / per_wave % 8 == 0  # whole 8-wave workgroups
    assert m.value == (iters + 1) * 15 * 2  # two waves per SIMD
    tops = o.value / s.value / 1e12
    assert 300 < tops <= 5000 * 1.02, tops  # never above the dense 5.0 POP/s peak
