"""GPU test of pq3d_gf_assemble_noalloc, the assembly entry point that
allocates nothing (its staging lives in the library's fixed device scratch),
against pq3d_gf_assemble through the C ABI: the same kernels on the same
inputs, so the results must be bit-identical.

Two calls in a row share the scratch, the second with a smaller input (and
with row/column index lists and FP32 output), so stale contents of the first
call must not leak into the second; a third call whose staging exceeds the
scratch is rejected on the host with a message.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MU = LAM = 32038120320.0


def _call(name, P, xg, ev, out_type, rows=None, cols=None):
    """Six matrices (nr x nc, receiver x source) from entry point ``name``."""
    from pyquake3d_amd import _lib
    L = _lib.lib()
    f = getattr(L, name)
    dp, ll, i, vp = ctypes.POINTER(ctypes.c_double), ctypes.c_longlong, ctypes.c_int, ctypes.c_void_p
    lp = ctypes.POINTER(ctypes.c_longlong)
    f.argtypes = [dp, dp, dp, ll, dp, dp, ll, ctypes.c_double, ctypes.c_double, i, i, ctypes.POINTER(vp), ll, lp,
                  ll, lp, ll, i]
    f.restype = i
    P1, P2, P3 = (np.ascontiguousarray(p, dtype=np.float64) for p in P)
    xg = np.ascontiguousarray(xg, dtype=np.float64)
    ev = np.ascontiguousarray(ev, dtype=np.float64)
    nsrc, nrcv = len(P1), len(xg)
    nr = nrcv if rows is None else len(rows)
    nc = nsrc if cols is None else len(cols)
    esz = 8 if out_type == _lib.MAT_F64 else 4
    ld = (nr + 3) // 4 * 4
    bufs = [_lib.DeviceBuffer(ld * nc * esz) for _ in range(6)]
    for b in bufs:
        _lib.check(L.pq3d_memset(vp(b.ptr), 0xFF, b.nbytes), "memset")   # garbage, so unwritten entries show
    ptrs = (vp * 6)(*[vp(b.ptr) for b in bufs])
    r = None if rows is None else np.ascontiguousarray(rows, dtype=np.int64)
    c = None if cols is None else np.ascontiguousarray(cols, dtype=np.int64)
    rc = f(_lib.ptr(P1), _lib.ptr(P2), _lib.ptr(P3), nsrc, _lib.ptr(xg), _lib.ptr(ev), nrcv, MU, LAM, 1, out_type,
           ptrs, ld, None if r is None else _lib.i64ptr(r), nr, None if c is None else _lib.i64ptr(c), nc, 0)
    _lib.check(rc, name)
    dt = np.float64 if out_type == _lib.MAT_F64 else np.float32
    out = []
    for b in bufs:
        a = b.download(np.empty((nc, ld), dtype=dt))
        out.append(a[:, :nr].T.copy())
    return np.stack(out)


def _mesh(kind, a, b):
    from pyquake3d_amd import greenfunc, readmsh, synth
    if kind == "rough":
        nodes_km, tri = synth.rough_fault_mesh(a, b, length_km=12.0, depth_km=6.0, seed=2)
    else:
        nodes_km, tri = synth.planar_fault_mesh(a, b, 8.0, 4.0)
    nodes = nodes_km * 1e3
    ev, xg = readmsh.get_eleVec(nodes, tri, False)
    return greenfunc.source_vertices(nodes, tri), xg, ev


def test_noalloc_assembly_equals_allocating_entry():
    from pyquake3d_amd import _lib
    # first call: a 1,200-element rough fault, all pairs, FP64 output
    P, xg, ev = _mesh("rough", 30, 20)
    assert len(xg) == 1200
    want = _call("pq3d_gf_assemble", P, xg, ev, _lib.MAT_F64)
    got = _call("pq3d_gf_assemble_noalloc", P, xg, ev, _lib.MAT_F64)
    assert np.isfinite(want).all() and np.abs(want[0]).max() > 0
    assert np.array_equal(got, want)
    # second call: smaller (160 elements), index lists, FP32 output
    P, xg, ev = _mesh("planar", 10, 8)
    rng = np.random.default_rng(1)
    rows = rng.permutation(len(xg))[:57]
    cols = rng.permutation(len(xg))[:91]
    want = _call("pq3d_gf_assemble", P, xg, ev, _lib.MAT_F32, rows, cols)
    got = _call("pq3d_gf_assemble_noalloc", P, xg, ev, _lib.MAT_F32, rows, cols)
    assert want.shape == (6, 57, 91) and np.isfinite(want).all()
    assert np.array_equal(got, want)


def test_noalloc_assembly_rejects_oversized_staging():
    """An input whose staging exceeds the scratch is refused before anything
    is read or launched (status -2, message names the limit)."""
    from pyquake3d_amd import _lib
    L = _lib.lib()
    f = L.pq3d_gf_assemble_noalloc
    dp, ll, i, vp = ctypes.POINTER(ctypes.c_double), ctypes.c_longlong, ctypes.c_int, ctypes.c_void_p
    lp = ctypes.POINTER(ctypes.c_longlong)
    f.argtypes = [dp, dp, dp, ll, dp, dp, ll, ctypes.c_double, ctypes.c_double, i, i, ctypes.POINTER(vp), ll, lp,
                  ll, lp, ll, i]
    f.restype = i
    dummy = np.zeros(3)
    ptrs = (vp * 6)()
    n = 1_000_000                  # 1e6 sources: about 1.3 GB of staging
    rc = f(_lib.ptr(dummy), _lib.ptr(dummy), _lib.ptr(dummy), n, _lib.ptr(dummy), _lib.ptr(dummy), 4, MU, LAM, 1,
           _lib.MAT_F64, ptrs, 4, None, 0, None, 0, 0)
    assert rc == -2
    msg = L.pq3d_last_error().decode()
    assert "scratch holds" in msg and "column chunks" in msg
