"""The measurement probes at the C ABI (pq3d_probe_stream, pq3d_probe_fma64):
their rates lie between a tenth of the published MI355X peaks and those
peaks (a reading above a published peak would mean the probe's byte or flop
count is wrong), and bad arguments are refused with status -2."""
import ctypes

import pytest

pytestmark = pytest.mark.gpu


def _lib():
    from pyquake3d_amd import _lib
    L = _lib.lib()
    fp, dp = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double)
    _lib._sig(L, "pq3d_probe_stream", [ctypes.c_longlong, ctypes.c_int, fp, dp])
    _lib._sig(L, "pq3d_probe_fma64", [ctypes.c_int, ctypes.c_int, fp, dp])
    return L


def test_stream_probe_rate():
    L = _lib()
    ms, gbs = ctypes.c_float(), ctypes.c_double()
    assert L.pq3d_probe_stream(1 << 30, 3, ctypes.byref(ms), ctypes.byref(gbs)) == 0
    print(f"read stream {gbs.value:.0f} GB/s", flush=True)
    assert 800.0 < gbs.value <= 8000.0 and ms.value > 0
    assert L.pq3d_probe_stream(8, 3, ctypes.byref(ms), ctypes.byref(gbs)) == -2


def test_fma64_probe_rate():
    L = _lib()
    ms, tf = ctypes.c_float(), ctypes.c_double()
    assert L.pq3d_probe_fma64(1024, 3, ctypes.byref(ms), ctypes.byref(tf)) == 0
    print(f"FP64 FMA {tf.value:.1f} TFLOP/s", flush=True)
    assert 7.86 < tf.value <= 78.6 and ms.value > 0
    assert L.pq3d_probe_fma64(0, 3, ctypes.byref(ms), ctypes.byref(tf)) == -2
