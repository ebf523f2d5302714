"""ctypes binding to the reference's TD stress library built from source.

TEST INFRASTRUCTURE ONLY.  ``oracle/Makefile`` compiles
``src/TDstressFS_C.cpp`` of the reference (with its own flags, src/Makefile:11)
into ``oracle/_ref/libTDstressFS_ref.so``.  The argument types mirror the
reference's own binding (src/Hmatrix.py:32-56).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
REF_SO = os.path.join(_HERE, "_ref", "libTDstressFS_ref.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


def available():
    return os.path.exists(REF_SO)


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(REF_SO)
        L.TDstressFS_C.argtypes = [_dp, _dp, _dp, ctypes.c_size_t, _dp, _dp, _dp,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_bool, _dp, _dp]
        L.TDstressFS_C.restype = None
        L.TDstressEachSourceAtReceiver_C.argtypes = [
            ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp, ctypes.c_size_t,
            ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
            ctypes.c_bool, _dp, _dp]
        L.TDstressEachSourceAtReceiver_C.restype = None
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def td_stress_fs_c(X, Y, Z, P1, P2, P3, Ss, Ds, Ts, mu, lam, halfspace):
    """n receivers x 1 source -> stress, strain (n, 6) from the reference C++."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    Z = np.ascontiguousarray(Z, dtype=np.float64)
    P1 = np.ascontiguousarray(P1, dtype=np.float64)
    P2 = np.ascontiguousarray(P2, dtype=np.float64)
    P3 = np.ascontiguousarray(P3, dtype=np.float64)
    n = X.size
    st = np.zeros(n * 6)
    sn = np.zeros(n * 6)
    lib().TDstressFS_C(_p(X), _p(Y), _p(Z), n, _p(P1), _p(P2), _p(P3),
                       Ss, Ds, Ts, mu, lam, bool(halfspace), _p(st), _p(sn))
    return st.reshape(n, 6), sn.reshape(n, 6)


def td_stress_each_source(x, y, z, P1, P2, P3, Ss, Ds, Ts, mu, lam, halfspace):
    """1 receiver x n sources (P* are n x 3) -> stress, strain (n, 6)."""
    P1 = np.ascontiguousarray(P1, dtype=np.float64)
    P2 = np.ascontiguousarray(P2, dtype=np.float64)
    P3 = np.ascontiguousarray(P3, dtype=np.float64)
    n = P1.shape[0]
    st = np.zeros(n * 6)
    sn = np.zeros(n * 6)
    lib().TDstressEachSourceAtReceiver_C(x, y, z, _p(P1), _p(P2), _p(P3), n,
                                         Ss, Ds, Ts, mu, lam, bool(halfspace), _p(st), _p(sn))
    return st.reshape(n, 6), sn.reshape(n, 6)


def td_stress_pairs(R, P1, P2, P3, Ss, Ds, Ts, mu, lam, halfspace):
    """Pairwise evaluation (R[i] with TD i) through the reference library."""
    R = np.asarray(R, dtype=np.float64)
    m = R.shape[0]
    st = np.zeros((m, 6))
    sn = np.zeros((m, 6))
    for i in range(m):
        a, b = td_stress_fs_c(R[i:i + 1, 0], R[i:i + 1, 1], R[i:i + 1, 2],
                              P1[i], P2[i], P3[i], Ss, Ds, Ts, mu, lam, halfspace)
        st[i] = a[0]
        sn[i] = b[0]
    return st, sn
