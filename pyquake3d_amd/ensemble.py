"""Ensemble of simulations on one shared mesh (batched GEMM on MFMA).

``Ensemble(n, B=64, matrices_from=dense_sim)`` holds B members, each with its
own friction parameters and state, sharing one set of FP32 traction Green's
functions (borrowed from a ``DenseSim`` with FP32 storage, or owned).  One
global step = one accepted Dormand-Prince step for every active member, each
with its own step size (the reference's accept/reject rule per member,
QDsim_gpu.py:959-985); members that fail are masked out.

Field arrays are (B, n) float64 with the names of ``QDsim_gpu.FIELDS``.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check
from .QDsim_gpu import FIELD_ID


def _declare(L):
    vp, ll, i, d = ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int, ctypes.c_double
    dp = ctypes.POINTER(ctypes.c_double)
    _lib._sig(L, "pq3d_ens_create", [ctypes.POINTER(vp), ll, i, ctypes.POINTER(vp), ll, i])
    _lib._sig(L, "pq3d_ens_destroy", [vp])
    _lib._sig(L, "pq3d_ens_matrix_ptrs", [vp, ctypes.POINTER(vp), ctypes.POINTER(ll)])
    _lib._sig(L, "pq3d_ens_set_field", [vp, i, dp])
    _lib._sig(L, "pq3d_ens_get_field", [vp, i, dp])
    _lib._sig(L, "pq3d_ens_set_params", [vp, d, d, d])
    _lib._sig(L, "pq3d_ens_members", [vp, dp, ctypes.POINTER(ll)])
    _lib._sig(L, "pq3d_ens_set_members", [vp, dp, ctypes.POINTER(ctypes.c_int)])
    _lib._sig(L, "pq3d_ens_gemm", [vp, i, ctypes.POINTER(ctypes.c_float)])
    _lib._sig(L, "pq3d_ens_gemm_result", [vp, dp])
    _lib._sig(L, "pq3d_ens_run", [vp, ll, i, ctypes.POINTER(ctypes.c_int)])


_lib.register(_declare)


class Ensemble:
    def __init__(self, n, B=64, matrices_from=None, kc=0):
        L = _lib.lib()
        h = ctypes.c_void_p()
        if matrices_from is not None:
            if matrices_from.mat_type != _lib.MAT_F32:
                raise ValueError("the ensemble GEMM reads FP32 matrices")
            ptrs = (ctypes.c_void_p * 6)(*[ctypes.c_void_p(p) for p in matrices_from.mat_ptrs])
            check(L.pq3d_ens_create(ctypes.byref(h), int(n), int(B), ptrs, int(matrices_from.ld), int(kc)),
                  "pq3d_ens_create")
            self._keep = matrices_from  # the borrowed matrices must outlive the ensemble
        else:
            check(L.pq3d_ens_create(ctypes.byref(h), int(n), int(B), None, 0, int(kc)), "pq3d_ens_create")
        self.h = h
        self.n, self.B = int(n), int(B)
        ptrs = (ctypes.c_void_p * 6)()
        ld = ctypes.c_longlong()
        check(L.pq3d_ens_matrix_ptrs(h, ptrs, ctypes.byref(ld)), "matrix_ptrs")
        self.mat_ptrs = [int(p) for p in ptrs]
        self.ld = ld.value

    def set(self, name, arr):
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(arr, dtype=np.float64), (self.B, self.n)))
        check(_lib.lib().pq3d_ens_set_field(self.h, FIELD_ID[name], _lib.ptr(a)), "ens set " + name)

    def get(self, name):
        out = np.empty((self.B, self.n))
        check(_lib.lib().pq3d_ens_get_field(self.h, FIELD_ID[name], _lib.ptr(out)), "ens get " + name)
        return out

    def set_params(self, V0, Vpl, eta):
        check(_lib.lib().pq3d_ens_set_params(self.h, V0, Vpl, eta), "ens params")

    def members(self):
        d = np.zeros((self.B, 5))
        i = np.zeros((self.B, 4), dtype=np.int64)
        check(_lib.lib().pq3d_ens_members(self.h, _lib.ptr(d), i.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))),
              "ens members")
        return {"time": d[:, 0], "dttry": d[:, 1], "h": d[:, 2], "dtnext": d[:, 3], "nrjct": i[:, 0],
                "status": i[:, 1], "active": i[:, 2], "steps": i[:, 3]}

    def set_members(self, dttry=None, active=None):
        dt = None if dttry is None else np.ascontiguousarray(np.broadcast_to(dttry, (self.B,)), dtype=np.float64)
        ac = None if active is None else np.ascontiguousarray(np.broadcast_to(active, (self.B,)), dtype=np.int32)
        check(_lib.lib().pq3d_ens_set_members(self.h, _lib.ptr(dt) if dt is not None else None,
                                              ac.ctypes.data_as(ctypes.POINTER(ctypes.c_int)) if ac is not None
                                              else None), "ens set members")

    def gemm(self, reps=0):
        f = ctypes.c_float()
        check(_lib.lib().pq3d_ens_gemm(self.h, int(reps), ctypes.byref(f)), "ens gemm")
        return f.value

    def gemm_result(self):
        out = np.empty((3, self.B, self.n))
        check(_lib.lib().pq3d_ens_gemm_result(self.h, _lib.ptr(out)), "ens gemm result")
        return out

    def run(self, nsteps, graph_steps=4):
        st = ctypes.c_int()
        check(_lib.lib().pq3d_ens_run(self.h, int(nsteps), int(graph_steps), ctypes.byref(st)), "ens run")
        return st.value

    def close(self):
        if self.h:
            _lib.lib().pq3d_ens_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def init_members(ens, cols_per_member, V0, Vpl, mu, Cs):
    """Steady-state initial condition per member from init-file columns
    (B, n, 10) in the reference's layout (QDsim_gpu.py:714-770)."""
    cols = np.asarray(cols_per_member, dtype=np.float64)
    rake, a, b, dc, f0, _, Tn, V, sl, nl = np.moveaxis(cols, 2, 0)
    Tt = Tn * a * np.arcsinh(V / (2.0 * V0) * np.exp((f0 + b * np.log(V0 / Vpl)) / a))
    state = np.log(np.sinh(Tt / Tn / a) * 2.0 * V0 / V) * a
    x, y = np.cos(rake), np.sin(rake)
    v1 = V * x + 1e-16
    v2 = V * y + 1e-16
    vals = {"Tno": Tn, "Tt1o": Tt * x, "Tt2o": Tt * y, "state": state, "slipv1": v1, "slipv2": v2,
            "slipv": np.sqrt(v1 * v1 + v2 * v2), "a": a, "b": b, "dc": dc, "f0": f0, "shear_loading": sl,
            "normal_loading": nl}
    for k, v in vals.items():
        ens.set(k, v)
    ens.set("rake0", rake)
    ens.set_params(V0, Vpl, mu / (2.0 * Cs))
