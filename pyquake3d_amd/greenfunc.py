"""Triangular-dislocation stress and traction Green's functions on the GPU.

Python-level mirror of the reference's stress API, backed by libpq3d.so:

* ``TDstressFS`` / ``TDstressHS``  -- src/SH_greenfunction.py:200-318, :861-889
  (returns ``(Stress, Strain)`` with shape (6, n), rows [xx, yy, zz, xy, xz, yz]).
  Unlike the reference, the vertex arrays are not modified in place
  (SH_greenfunction.py:873-875 negates P*[2]).
* ``lib`` -- the ctypes handle exposing ``TDstressFS_C`` and
  ``TDstressEachSourceAtReceiver_C`` with the reference's signatures
  (src/Hmatrix.py:28-58).
* ``traction_gf`` -- the six dense traction Green's functions
  (src/QDsim_gpu.py:1745-1878), optionally restricted to row/column subsets.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, ptr, i64ptr

MATRIX_NAMES = ("A1s", "A2s", "Bs", "A1d", "A2d", "Bd")


class _LibProxy:
    """Attribute access to the loaded library (loaded on first use)."""

    def __getattr__(self, name):
        return getattr(_lib.lib(), name)


lib = _LibProxy()


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel())


def _stress(X, Y, Z, P1, P2, P3, Ss, Ds, Ts, mu, lam, halfspace):
    X, Y, Z = _f64(X), _f64(Y), _f64(Z)
    P1, P2, P3 = _f64(P1), _f64(P2), _f64(P3)
    n = X.size
    st = np.zeros(n * 6)
    sn = np.zeros(n * 6)
    _lib.lib().TDstressFS_C(ptr(X), ptr(Y), ptr(Z), n, ptr(P1), ptr(P2), ptr(P3), float(Ss), float(Ds),
                            float(Ts), float(mu), float(lam), bool(halfspace), ptr(st), ptr(sn))
    if n and np.isnan(st[0]) and np.isnan(st).all():
        raise _lib.Pq3dError("TDstressFS_C failed: " + _lib.lib().pq3d_last_error().decode())
    return st.reshape(n, 6).T.copy(), sn.reshape(n, 6).T.copy()


def TDstressFS(X, Y, Z, P1, P2, P3, Ss, Ds, Ts, mu, lambda_):
    """Full-space TD stress/strain at n receivers (SH_greenfunction.py:200)."""
    return _stress(X, Y, Z, P1, P2, P3, Ss, Ds, Ts, mu, lambda_, False)


def TDstressHS(X, Y, Z, P1, P2, P3, Ss, Ds, Ts, mu, lambda_):
    """Half-space TD stress/strain at n receivers (SH_greenfunction.py:861)."""
    return _stress(X, Y, Z, P1, P2, P3, Ss, Ds, Ts, mu, lambda_, True)


def TDstressEachSourceAtReceiver(x, y, z, P1, P2, P3, Ss, Ds, Ts, mu, lam, halfspace=True):
    """One receiver x n TDs (P* are n x 3); returns stress, strain (n, 6)."""
    P1 = np.ascontiguousarray(P1, dtype=np.float64)
    P2 = np.ascontiguousarray(P2, dtype=np.float64)
    P3 = np.ascontiguousarray(P3, dtype=np.float64)
    n = P1.shape[0]
    st = np.zeros(n * 6)
    sn = np.zeros(n * 6)
    _lib.lib().TDstressEachSourceAtReceiver_C(float(x), float(y), float(z), ptr(P1), ptr(P2), ptr(P3), n,
                                              float(Ss), float(Ds), float(Ts), float(mu), float(lam),
                                              bool(halfspace), ptr(st), ptr(sn))
    return st.reshape(n, 6), sn.reshape(n, 6)


def source_vertices(nodelst, elelst):
    """P1, P2, P3 (N x 3) of every TD from 1-based connectivity."""
    e = np.asarray(elelst, dtype=np.int64) - 1
    nodes = np.asarray(nodelst, dtype=np.float64)
    return (np.ascontiguousarray(nodes[e[:, 0]]), np.ascontiguousarray(nodes[e[:, 1]]),
            np.ascontiguousarray(nodes[e[:, 2]]))


def assemble_device(P1, P2, P3, xg, ev, mu, lam, halfspace, d_out, ld, out_type=_lib.MAT_F64, rows=None,
                    cols=None, src_per_block=0):
    """Write the six traction Green's functions into device buffers d_out
    (six device pointers); element (i, j) at j*ld + i (source-major)."""
    P1, P2, P3 = (np.ascontiguousarray(p, dtype=np.float64) for p in (P1, P2, P3))
    xg = np.ascontiguousarray(xg, dtype=np.float64)
    ev = np.ascontiguousarray(ev, dtype=np.float64)
    r = None if rows is None else np.ascontiguousarray(rows, dtype=np.int64)
    c = None if cols is None else np.ascontiguousarray(cols, dtype=np.int64)
    outs = (ctypes.c_void_p * 6)(*[ctypes.c_void_p(int(p)) for p in d_out])
    nr = len(xg) if r is None else len(r)
    nc = len(P1) if c is None else len(c)
    rc = _lib.lib().pq3d_gf_assemble(ptr(P1), ptr(P2), ptr(P3), len(P1), ptr(xg), ptr(ev), len(xg), float(mu),
                                     float(lam), int(bool(halfspace)), int(out_type), outs, int(ld),
                                     i64ptr(r) if r is not None else None, nr,
                                     i64ptr(c) if c is not None else None, nc, int(src_per_block))
    check(rc, "pq3d_gf_assemble")


def traction_gf(P1, P2, P3, xg, ev, mu, lam, halfspace=True, rows=None, cols=None, dtype=np.float64):
    """Host copies of (A1s, A2s, Bs, A1d, A2d, Bd) restricted to rows x cols,
    each C-ordered (row = receiver, column = source) as the reference stores
    them (QDsim_gpu.py:236-244)."""
    nr = len(xg) if rows is None else len(rows)
    nc = len(P1) if cols is None else len(cols)
    out_type = _lib.MAT_F64 if dtype == np.float64 else _lib.MAT_F32
    esz = 8 if out_type == _lib.MAT_F64 else 4
    ld = max(nr, 1)
    bufs = [_lib.DeviceBuffer(max(nr * nc * esz, 8)) for _ in range(6)]
    assemble_device(P1, P2, P3, xg, ev, mu, lam, halfspace, [b.ptr for b in bufs], ld, out_type, rows, cols)
    out = []
    for b in bufs:
        h = np.empty((nc, nr), dtype=np.float64 if esz == 8 else np.float32)
        b.download(h)
        out.append(np.ascontiguousarray(h.T))
        b.free()
    return tuple(out)
