"""Ensemble of simulations on one shared mesh (batched GEMM on MFMA).

``Ensemble(n, B=64, matrices_from=dense_sim)`` holds B slots ("members"), each
with its own friction parameters and state, sharing one set of FP32 traction
Green's functions (borrowed from a ``DenseSim`` with FP32 storage, or owned).
One global step = one GEMM for all members, then each active member's
Dormand-Prince attempts with its own step size under the reference's
accept/reject rule (QDsim_gpu.py:959-985); a member that has not accepted
after the step's attempt rounds continues at the next global step, and a
member that fails is masked out.  All work of one ensemble runs on its own
HIP stream (``run(..., wait=False)`` returns immediately), so ensembles of
different meshes overlap on the device.

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
    _lib._sig(L, "pq3d_ens_device_bytes", [vp, dp])
    _lib._sig(L, "pq3d_ens_set_field", [vp, i, dp])
    _lib._sig(L, "pq3d_ens_get_field", [vp, i, dp])
    _lib._sig(L, "pq3d_ens_set_member_field", [vp, i, i, dp])
    _lib._sig(L, "pq3d_ens_get_member_field", [vp, i, i, dp])
    _lib._sig(L, "pq3d_ens_set_params", [vp, d, d, d, i])
    _lib._sig(L, "pq3d_ens_members", [vp, dp, ctypes.POINTER(ll)])
    _lib._sig(L, "pq3d_ens_reset_member", [vp, i, d, d, i])
    _lib._sig(L, "pq3d_ens_gemm", [vp, i, ctypes.POINTER(ctypes.c_float)])
    _lib._sig(L, "pq3d_ens_gemm_result", [vp, dp])
    _lib._sig(L, "pq3d_ens_time_product", [vp, i, ctypes.POINTER(ctypes.c_float)])
    _lib._sig(L, "pq3d_ens_set_kc", [vp, i])
    _lib._sig(L, "pq3d_ens_get_kc", [vp, ctypes.POINTER(i), ctypes.POINTER(i)])
    _lib._sig(L, "pq3d_ens_set_precision", [vp, i])
    _lib._sig(L, "pq3d_ens_get_precision", [vp, ctypes.POINTER(i), ctypes.POINTER(i)])
    _lib._sig(L, "pq3d_ens_scan_structure", [vp, ctypes.POINTER(i)])
    _lib._sig(L, "pq3d_ens_set_structure", [vp, i])
    _lib._sig(L, "pq3d_ens_get_structure", [vp, ctypes.POINTER(i), ctypes.POINTER(i)])
    _lib._sig(L, "pq3d_ens_run", [vp, ll, i, i])
    _lib._sig(L, "pq3d_ens_sync", [vp])
    _lib._sig(L, "pq3d_ens_query", [vp, ctypes.POINTER(ctypes.c_int)])


_lib.register(_declare)


class Ensemble:
    def __init__(self, n, B=64, matrices_from=None, kc=0, precision="bf16x3"):
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
        self.mat_ptrs = [int(p) if p else 0 for p in ptrs]
        self.ld = ld.value
        # the GEMM skips Bs and Bd when both hold only zeros (planar fault);
        # rescanned before the first product after the matrices change
        self.exploit_structure = True
        self.mat_gen = 0
        self._structure_gen = None
        self.fix_tn = False
        if precision != "fp32":
            self.set_precision(precision)

    def device_bytes(self):
        """Device memory the ensemble holds (borrowed matrices not counted)."""
        b = ctypes.c_double()
        check(_lib.lib().pq3d_ens_device_bytes(self.h, ctypes.byref(b)), "ens device bytes")
        return b.value

    def matrices_changed(self):
        """Mark owned matrices as rewritten through ``mat_ptrs``."""
        self.mat_gen += 1

    def _sync_structure(self):
        src = self.__dict__.get("_keep")
        gen = (src.mat_gen if src is not None else self.mat_gen) if self.exploit_structure else -1
        if gen == self._structure_gen:
            return
        if gen == -1:
            check(_lib.lib().pq3d_ens_set_structure(self.h, 0), "ens set structure")
        else:
            check(_lib.lib().pq3d_ens_scan_structure(self.h, None), "ens scan structure")
        self._structure_gen = gen

    @property
    def structure(self):
        """(zero-matrix mask, whether the GEMM skips Bs and Bd)."""
        self._sync_structure()
        m, sk = ctypes.c_int(), ctypes.c_int()
        check(_lib.lib().pq3d_ens_get_structure(self.h, ctypes.byref(m), ctypes.byref(sk)), "ens get structure")
        return m.value, bool(sk.value)

    # ---- fields
    def set(self, name, arr):
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(arr, dtype=np.float64), (self.B, self.n)))
        check(_lib.lib().pq3d_ens_set_field(self.h, FIELD_ID[name], _lib.ptr(a)), "ens set " + name)

    def get(self, name):
        out = np.empty((self.B, self.n))
        check(_lib.lib().pq3d_ens_get_field(self.h, FIELD_ID[name], _lib.ptr(out)), "ens get " + name)
        return out

    def set_member(self, m, name, arr):
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(arr, dtype=np.float64), (self.n,)))
        check(_lib.lib().pq3d_ens_set_member_field(self.h, int(m), FIELD_ID[name], _lib.ptr(a)),
              "ens set member " + name)

    def get_member(self, m, name):
        out = np.empty(self.n)
        check(_lib.lib().pq3d_ens_get_member_field(self.h, int(m), FIELD_ID[name], _lib.ptr(out)),
              "ens get member " + name)
        return out

    def set_params(self, V0, Vpl, eta, fix_tn=False):
        """V0, Vpl, eta = mu / (2 Cs); ``fix_tn``: the reference's Fix_Tn
        (QDsim_gpu.py:950-953), dsigma/dt is the normal loading alone and the
        GEMM leaves out Bs and Bd."""
        check(_lib.lib().pq3d_ens_set_params(self.h, V0, Vpl, eta, int(bool(fix_tn))), "ens params")
        self.fix_tn = bool(fix_tn)

    # ---- member control
    def members(self):
        d = np.zeros((self.B, 6))
        i = np.zeros((self.B, 6), dtype=np.int64)
        check(_lib.lib().pq3d_ens_members(self.h, _lib.ptr(d), i.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))),
              "ens members")
        return {"time": d[:, 0], "dttry": d[:, 1], "h": d[:, 2], "dtnext": d[:, 3], "t_end": d[:, 4],
                "nrjct": i[:, 0], "status": i[:, 1], "active": i[:, 2], "finished": i[:, 3], "steps": i[:, 4],
                "finish_gstep": i[:, 5]}

    def reset_member(self, m, h0=1e-3, t_end=0.0, active=True):
        check(_lib.lib().pq3d_ens_reset_member(self.h, int(m), float(h0), float(t_end), int(bool(active))),
              "ens reset member")

    # ---- execution
    def gemm(self, reps=0):
        self._sync_structure()
        f = ctypes.c_float()
        check(_lib.lib().pq3d_ens_gemm(self.h, int(reps), ctypes.byref(f)), "ens gemm")
        return f.value

    def time_product(self, reps=4):
        """ms per product as a global step uses it (operand preparation, GEMM,
        one pass over the split-K partial sums)."""
        self._sync_structure()
        f = ctypes.c_float()
        check(_lib.lib().pq3d_ens_time_product(self.h, int(reps), ctypes.byref(f)), "ens time product")
        return f.value

    @property
    def kc(self):
        """Split-K width of the GEMM tiles (64 rows x 64 members x kc sources)."""
        kc, nc = ctypes.c_int(), ctypes.c_int()
        check(_lib.lib().pq3d_ens_get_kc(self.h, ctypes.byref(kc), ctypes.byref(nc)), "ens get kc")
        return kc.value

    def set_kc(self, kc):
        check(_lib.lib().pq3d_ens_set_kc(self.h, int(kc)), "ens set kc")

    PRECISIONS = {"fp32": 0, "bf16x3": 1}

    def set_precision(self, mode):
        """GEMM arithmetic: "fp32" (FP32 MFMA) or "bf16x3" (each FP32 operand
        split exactly into three BF16 parts, the six leading products on the
        BF16 MFMA units with FP32 accumulation: a shear pass, and a normal pass
        over Bs and Bd unless both are zero or fix_tn is set).  "bf16x3" rounds
        kc up to a multiple of 32."""
        check(_lib.lib().pq3d_ens_set_precision(self.h, int(self.PRECISIONS[mode])), "ens set precision")

    @property
    def precision(self):
        """(mode as set, whether the BF16-split GEMM is the one launched)."""
        self._sync_structure()
        m, a = ctypes.c_int(), ctypes.c_int()
        check(_lib.lib().pq3d_ens_get_precision(self.h, ctypes.byref(m), ctypes.byref(a)), "ens get precision")
        return ("fp32", "bf16x3")[m.value], bool(a.value)

    def gemm_result(self):
        out = np.empty((3, self.B, self.n))
        check(_lib.lib().pq3d_ens_gemm_result(self.h, _lib.ptr(out)), "ens gemm result")
        return out

    def run(self, nsteps, graph_steps=4, wait=True):
        self._sync_structure()
        check(_lib.lib().pq3d_ens_run(self.h, int(nsteps), int(graph_steps), 0 if wait else 1), "ens run")

    def sync(self):
        check(_lib.lib().pq3d_ens_sync(self.h), "ens sync")

    def done(self):
        d = ctypes.c_int()
        check(_lib.lib().pq3d_ens_query(self.h, ctypes.byref(d)), "ens query")
        return bool(d.value)

    def close(self):
        if self.h:
            _lib.lib().pq3d_ens_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def steady_state_fields(cols, V0, Vpl):
    """Field values of one member from its init-file columns (n, 10) in the
    reference's layout (QDsim_gpu.py:714-770)."""
    rake, a, b, dc, f0, _, Tn, V, sl, nl = np.asarray(cols, dtype=np.float64).T
    Tt = Tn * a * np.arcsinh(V / (2.0 * V0) * np.exp((f0 + b * np.log(V0 / Vpl)) / a))
    state = np.log(np.sinh(Tt / Tn / a) * 2.0 * V0 / V) * a
    x, y = np.cos(rake), np.sin(rake)
    v1 = V * x + 1e-16
    v2 = V * y + 1e-16
    z = np.zeros_like(a)
    return {"Tno": Tn, "Tt1o": Tt * x, "Tt2o": Tt * y, "state": state, "slipv1": v1, "slipv2": v2,
            "slipv": np.sqrt(v1 * v1 + v2 * v2), "slip1": z, "slip2": z, "slip": z, "Tt": Tt, "rake": rake,
            "fric": Tt / Tn, "a": a, "b": b, "dc": dc, "f0": f0, "shear_loading": sl, "normal_loading": nl,
            "P": z, "dPdt0": z, "rake0": rake}


def init_members(ens, cols_per_member, V0, Vpl, mu, Cs, fix_tn=False):
    """Steady-state initial condition for all members from init-file columns
    (B, n, 10)."""
    cols = np.asarray(cols_per_member, dtype=np.float64)
    per = [steady_state_fields(cols[m], V0, Vpl) for m in range(cols.shape[0])]
    for k in per[0]:
        ens.set(k, np.stack([p[k] for p in per]))
    ens.set_params(V0, Vpl, mu / (2.0 * Cs), fix_tn)


def load_member(ens, m, cols, V0, Vpl, h0=1e-3, t_end=0.0):
    """Admit one member into slot m with its init-file columns (n, 10)."""
    for k, v in steady_state_fields(cols, V0, Vpl).items():
        ens.set_member(m, k, v)
    ens.reset_member(m, h0=h0, t_end=t_end, active=True)
