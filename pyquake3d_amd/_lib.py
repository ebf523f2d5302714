"""ctypes binding to libpq3d.so (the HIP implementation).

The library is built in-tree (``pyquake3d_amd/libpq3d.so``) by
``__graft_entry__.build()``.  There is no fallback: if the library is missing
or no GPU is visible, every entry point raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpq3d.so")

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_i64p = ctypes.POINTER(ctypes.c_longlong)

MAT_F64, MAT_F32, MAT_BF16, MAT_NONE = 0, 1, 2, 3
_lib = None


class Pq3dError(RuntimeError):
    pass


def _sig(L, name, argtypes, restype=ctypes.c_int):
    f = getattr(L, name)
    f.argtypes = argtypes
    f.restype = restype
    return f


def lib():
    """Load libpq3d.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise Pq3dError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                        "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    d, sz, b = ctypes.c_double, ctypes.c_size_t, ctypes.c_bool
    _sig(L, "TDstressFS_C", [_dp, _dp, _dp, sz, _dp, _dp, _dp, d, d, d, d, d, b, _dp, _dp], None)
    _sig(L, "TDstressEachSourceAtReceiver_C",
         [d, d, d, _dp, _dp, _dp, sz, d, d, d, d, d, b, _dp, _dp], None)
    _sig(L, "TDstressFS_C1", [_dp, _dp, _dp, sz, _dp, _dp, _dp, d, d, d, d, d] + [_dp] * 12, None)
    _sig(L, "pq3d_last_error", [], ctypes.c_char_p)
    _sig(L, "pq3d_device_count", [ctypes.POINTER(ctypes.c_int)])
    _sig(L, "pq3d_device_info", [ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                 ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int),
                                 ctypes.POINTER(ctypes.c_int)])
    _sig(L, "pq3d_malloc", [ctypes.POINTER(_vp), sz])
    _sig(L, "pq3d_free", [_vp])
    _sig(L, "pq3d_memcpy_h2d", [_vp, _vp, sz])
    _sig(L, "pq3d_memcpy_d2h", [_vp, _vp, sz])
    _sig(L, "pq3d_memcpy_d2d", [_vp, _vp, sz])
    _sig(L, "pq3d_memset", [_vp, ctypes.c_int, sz])
    _sig(L, "pq3d_synchronize", [])
    _sig(L, "pq3d_event_create", [ctypes.POINTER(_vp)])
    _sig(L, "pq3d_event_destroy", [_vp])
    _sig(L, "pq3d_event_record", [_vp])
    _sig(L, "pq3d_event_elapsed_ms", [_vp, _vp, ctypes.POINTER(ctypes.c_float)])
    ll = ctypes.c_longlong
    _sig(L, "pq3d_gf_assemble", [_dp, _dp, _dp, ll, _dp, _dp, ll, d, d, ctypes.c_int, ctypes.c_int,
                                 ctypes.POINTER(_vp), ll, _i64p, ll, _i64p, ll, ctypes.c_int])
    _lib = L
    _declare_extra(L)
    return L


_extra_decls = []


def register(decl):
    """Modules add their own entry-point declarations through this hook."""
    _extra_decls.append(decl)
    if _lib is not None:
        decl(_lib)


def _declare_extra(L):
    for decl in _extra_decls:
        decl(L)


def check(rc, what=""):
    if rc != 0:
        msg = lib().pq3d_last_error()
        raise Pq3dError(f"{what}: {msg.decode() if msg else 'error'} (status {rc})")


def ptr(a):
    """ctypes double* of a C-contiguous float64 array."""
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def i64ptr(a):
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(_i64p)


def device_count():
    n = ctypes.c_int(0)
    check(lib().pq3d_device_count(ctypes.byref(n)), "pq3d_device_count")
    return n.value


def device_info(dev=0):
    name = ctypes.create_string_buffer(256)
    cu = ctypes.c_int()
    hbm = ctypes.c_double()
    ma = ctypes.c_int()
    mi = ctypes.c_int()
    check(lib().pq3d_device_info(dev, name, 256, ctypes.byref(cu), ctypes.byref(hbm), ctypes.byref(ma),
                                 ctypes.byref(mi)), "pq3d_device_info")
    return {"name": name.value.decode(), "cu": cu.value, "hbm_bytes": hbm.value}


class DeviceBuffer:
    """Owning device allocation (freed on garbage collection)."""

    def __init__(self, nbytes):
        p = _vp()
        check(lib().pq3d_malloc(ctypes.byref(p), int(nbytes)), "pq3d_malloc")
        self.ptr = p.value
        self.nbytes = int(nbytes)

    def upload(self, arr, offset=0):
        arr = np.ascontiguousarray(arr)
        assert offset + arr.nbytes <= self.nbytes
        check(lib().pq3d_memcpy_h2d(_vp(self.ptr + offset), arr.ctypes.data_as(_vp), arr.nbytes), "h2d")

    def download(self, arr, offset=0):
        assert arr.flags.c_contiguous and offset + arr.nbytes <= self.nbytes
        check(lib().pq3d_memcpy_d2h(arr.ctypes.data_as(_vp), _vp(self.ptr + offset), arr.nbytes), "d2h")
        return arr

    def free(self):
        if self.ptr and _lib is not None:
            _lib.pq3d_free(_vp(self.ptr))
        self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Timer:
    """Elapsed device time between two events on the library stream."""

    def __init__(self):
        L = lib()
        self.a, self.b = _vp(), _vp()
        check(L.pq3d_event_create(ctypes.byref(self.a)), "event")
        check(L.pq3d_event_create(ctypes.byref(self.b)), "event")

    def start(self):
        check(lib().pq3d_event_record(self.a), "record")

    def stop(self):
        check(lib().pq3d_event_record(self.b), "record")

    def ms(self):
        f = ctypes.c_float()
        check(lib().pq3d_event_elapsed_ms(self.a, self.b, ctypes.byref(f)), "elapsed")
        return f.value
