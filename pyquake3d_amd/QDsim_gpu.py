"""Drop-in for the reference's ``QDsim_gpu`` module (src/QDsim_gpu.py).

``QDsim(elelst, nodelst, fnamePara)`` keeps the reference's constructor,
attributes and stepping calls:

* ``initGPUvariable()``                          QDsim_gpu.py:893-939
* ``simu_forwardGPU(dttry) -> (h, dtnext)``      QDsim_gpu.py:942-1027
* ``simu_forward(dttry) -> (h, dtnext)``         QDsim_gpu.py:1029-1122
* attributes ``Tno_gpu``, ``Tt1o_gpu``, ``Tt2o_gpu``, ``state_gpu``,
  ``slipv1_gpu``, ``slipv2_gpu``, ``slipv_gpu``, ``slip_gpu``, ``slip1_gpu``,
  ``slip2_gpu``, ``Tt_gpu``, ``rake_gpu``, ``fric_gpu`` ... with ``.get()``
  (main_gpu.py:120-129).

Everything runs on the GPU through libpq3d.so.  ``simu_forward`` -- the
reference's CPU branch -- is the same device step with the CPU branch's error
estimate (all three traction errors, QDsim_gpu.py:1373), after which the host
attributes (``slipv1``, ``Tt1o``, ...) are refreshed, as that branch leaves them.

The Green's functions are assembled on the GPU in FP64 (or FP32 storage with
the ``GF precision: fp32`` parameter key, which the reference does not have)
and cached as ``A1s.npy ... Bd.npy`` in the ``Corefunc directory``, in the
reference's layout (float64, row = receiver, column = source,
QDsim_gpu.py:188-245).  Dilatancy / pore pressure (``If Dilatancy:
True``) runs on the device after every accepted step (QDsim_gpu.py:611-671).
"""
import ctypes
import os
import sys
from collections import deque
from math import sqrt

import numpy as np

from . import _lib, autotune, greenfunc, outputs, readmsh
from ._lib import check

FIELDS = ("Tno", "Tt1o", "Tt2o", "state", "slipv1", "slipv2", "slipv", "slip1", "slip2", "slip", "Tt", "rake",
          "rake0", "fric", "a", "b", "dc", "f0", "shear_loading", "normal_loading", "P", "dPdt0",
          "AdotV1", "AdotV2", "dsigmadt")
FIELD_ID = {name: i for i, name in enumerate(FIELDS)}


def _declare(L):
    vp, d, ll, i = ctypes.c_void_p, ctypes.c_double, ctypes.c_longlong, ctypes.c_int
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    _lib._sig(L, "pq3d_dense_create", [ctypes.POINTER(vp), ll, i, i])
    _lib._sig(L, "pq3d_dense_destroy", [vp])
    _lib._sig(L, "pq3d_dense_matrix_ptrs", [vp, ctypes.POINTER(vp), ctypes.POINTER(ll)])
    _lib._sig(L, "pq3d_dense_set_field", [vp, i, dp])
    _lib._sig(L, "pq3d_dense_get_field", [vp, i, dp])
    _lib._sig(L, "pq3d_dense_field_ptr", [vp, i, ctypes.POINTER(vp)])
    _lib._sig(L, "pq3d_dense_set_params", [vp, d, d, d, i, i])
    _lib._sig(L, "pq3d_dense_set_time", [vp, d])
    _lib._sig(L, "pq3d_dense_state", [vp, dp, ip, ip, ctypes.POINTER(ll)])
    _lib._sig(L, "pq3d_dense_errors", [vp, dp, dp])
    _lib._sig(L, "pq3d_dense_device_bytes", [vp, dp])
    _lib._sig(L, "pq3d_dense_derivative", [vp, vp, vp, vp, vp, vp])
    _lib._sig(L, "pq3d_dense_dp45", [vp, d, i, vp, dp])
    _lib._sig(L, "pq3d_dense_matvec", [vp])
    _lib._sig(L, "pq3d_dense_time_matvec", [vp, i, ctypes.POINTER(ctypes.c_float)])
    _lib._sig(L, "pq3d_dense_time_product", [vp, i, ctypes.POINTER(ctypes.c_float)])
    _lib._sig(L, "pq3d_dense_set_kc", [vp, i])
    _lib._sig(L, "pq3d_dense_get_kc", [vp, ip, ip])
    _lib._sig(L, "pq3d_dense_scan_structure", [vp, ip])
    _lib._sig(L, "pq3d_dense_set_structure", [vp, i])
    _lib._sig(L, "pq3d_dense_get_structure", [vp, ip, ip])
    _lib._sig(L, "pq3d_dense_step", [vp, d, dp, dp, ip])
    _lib._sig(L, "pq3d_dense_run", [vp, ll, i, dp, ip])
    _lib._sig(L, "pq3d_dense_reset_ctrl", [vp])
    _lib._sig(L, "pq3d_dense_set_dilatancy", [vp, i, dp, dp, d, d, d, d, d, d])
    _lib._sig(L, "pq3d_dense_get_dilatancy", [vp, dp])
    _lib._sig(L, "pq3d_dense_dilatancy_update", [vp, d])


_lib.register(_declare)


def readPara0(fname):
    """``key: value`` parameter file, values kept as strings (QDsim_gpu.py:90-99)."""
    Para0 = {}
    with open(fname, "r") as f:
        for line in f:
            if ":" in line:
                tem = line.split(":")
                Para0[tem[0].strip()] = tem[1].strip()
    return Para0


def find_boundary_edges_and_nodes(triangles):
    """Edges used by one triangle only, and their nodes (QDsim_gpu.py:56-80)."""
    from collections import defaultdict
    edge_count = defaultdict(int)
    for tri in triangles:
        for e in ((tri[0], tri[1]), (tri[1], tri[2]), (tri[2], tri[0])):
            edge_count[tuple(sorted(e))] += 1
    boundary_edges = [e for e, c in edge_count.items() if c == 1]
    nodes = set()
    for e in boundary_edges:
        nodes.update(e)
    return boundary_edges, np.array(list(nodes))


class DenseSim:
    """Device-resident dense simulation (owns the six matrices and all fields)."""

    def __init__(self, n, mat_type=_lib.MAT_F64, kc=256):
        L = _lib.lib()
        h = ctypes.c_void_p()
        check(L.pq3d_dense_create(ctypes.byref(h), int(n), int(mat_type), int(kc)), "pq3d_dense_create")
        self.h = h
        self.n = int(n)
        self.mat_type = mat_type
        self.fix_tn = False
        # matrices that hold only zeros (Bs, Bd of a planar fault) are skipped by
        # the product; a scan runs before the first product after any change
        self.exploit_structure = True
        self.mat_gen = 0
        self._structure_gen = None
        ptrs = (ctypes.c_void_p * 6)()
        ld = ctypes.c_longlong()
        check(L.pq3d_dense_matrix_ptrs(h, ptrs, ctypes.byref(ld)), "matrix_ptrs")
        self.mat_ptrs = [int(p) if p else 0 for p in ptrs]
        self.ld = ld.value

    def set_dilatancy(self, zp, Parr, c_hyd, eps, hs, beta, P0, V0):
        """Enable the pore-pressure update after every accepted step
        (QDsim_gpu.py:628-671); ``Parr`` (n, K) profiles in Pa on the grid zp."""
        zp = np.ascontiguousarray(zp, dtype=np.float64)
        Parr = np.ascontiguousarray(Parr, dtype=np.float64)
        if Parr.shape != (self.n, len(zp)):
            raise ValueError("Parr must be (n, len(zp))")
        check(_lib.lib().pq3d_dense_set_dilatancy(self.h, len(zp), _lib.ptr(zp), _lib.ptr(Parr), float(c_hyd),
                                                  float(eps), float(hs), float(beta), float(P0), float(V0)),
              "set_dilatancy")
        self._dila_K = len(zp)

    def dilatancy_update(self, dt):
        """One pore-pressure update with step dt (Calc_P_implicit, QDsim_gpu.py:628)."""
        check(_lib.lib().pq3d_dense_dilatancy_update(self.h, float(dt)), "dilatancy_update")

    def get_dilatancy(self):
        """Current pore-pressure profiles (n, K) in Pa."""
        out = np.empty((self.n, self._dila_K))
        check(_lib.lib().pq3d_dense_get_dilatancy(self.h, _lib.ptr(out)), "get_dilatancy")
        return out

    def attach_hmatrix(self, hmat):
        """Take the traction rates from a ``Hmatrix.DeviceHmatrix`` (create the
        simulation with ``MAT_NONE`` so that no dense matrices are allocated)."""
        check(_lib.lib().pq3d_dense_attach_hmatrix(self.h, hmat.h), "attach_hmatrix")
        self._hmat = hmat  # must outlive the simulation's use of it

    def set(self, name, arr):
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(arr, dtype=np.float64), (self.n,)))
        check(_lib.lib().pq3d_dense_set_field(self.h, FIELD_ID[name], _lib.ptr(a)), "set_field " + name)

    def get(self, name):
        out = np.empty(self.n)
        check(_lib.lib().pq3d_dense_get_field(self.h, FIELD_ID[name], _lib.ptr(out)), "get_field " + name)
        return out

    @property
    def has_dense(self):
        """True when the traction product runs on this simulation's own dense
        matrices (not MAT_NONE, no H-matrix attached)."""
        return self.mat_type != _lib.MAT_NONE and "_hmat" not in self.__dict__

    def set_params(self, V0, Vpl, eta, fix_tn=False, gpu_semantics=True):
        self.fix_tn = bool(fix_tn)
        check(_lib.lib().pq3d_dense_set_params(self.h, V0, Vpl, eta, int(bool(fix_tn)), int(bool(gpu_semantics))),
              "set_params")

    def set_time(self, t):
        check(_lib.lib().pq3d_dense_set_time(self.h, float(t)), "set_time")

    def upload_matrix(self, k, A):
        """Host (receiver x source) matrix -> device, converted to the storage type."""
        A = np.asarray(A)
        n = self.n
        esz = {_lib.MAT_F64: 8, _lib.MAT_F32: 4, _lib.MAT_BF16: 2}[self.mat_type]
        if self.mat_type == _lib.MAT_F64:
            col = np.zeros((n, self.ld), dtype=np.float64)
            col[:, :n] = A.T
        elif self.mat_type == _lib.MAT_F32:
            col = np.zeros((n, self.ld), dtype=np.float32)
            col[:, :n] = A.T
        else:
            f = np.ascontiguousarray(A.T, dtype=np.float32)
            bits = f.view(np.uint32)
            rnd = ((bits >> 16) & 1) + 0x7FFF  # round to nearest even
            col = np.zeros((n, self.ld), dtype=np.uint16)
            col[:, :n] = ((bits + rnd) >> 16).astype(np.uint16)
        check(_lib.lib().pq3d_memcpy_h2d(ctypes.c_void_p(self.mat_ptrs[k]), col.ctypes.data_as(ctypes.c_void_p),
                                         n * self.ld * esz), "upload matrix")
        self.matrices_changed()

    def download_matrix(self, k):
        """Device matrix -> host float64 (receiver x source)."""
        n = self.n
        dt = {_lib.MAT_F64: np.float64, _lib.MAT_F32: np.float32, _lib.MAT_BF16: np.uint16}[self.mat_type]
        col = np.empty((n, self.ld), dtype=dt)
        check(_lib.lib().pq3d_memcpy_d2h(col.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.mat_ptrs[k]),
                                         col.nbytes), "download matrix")
        if dt == np.uint16:
            col = (col.astype(np.uint32) << 16).view(np.float32)
        return np.ascontiguousarray(col[:, :n].T, dtype=np.float64)

    def assemble(self, P1, P2, P3, xg, ev, mu, lam, halfspace):
        if self.mat_type == _lib.MAT_BF16:
            raise ValueError("assemble into FP64 or FP32 storage")
        greenfunc.assemble_device(P1, P2, P3, xg, ev, mu, lam, halfspace, self.mat_ptrs, self.ld,
                                  self.mat_type)
        self.matrices_changed()

    def matrices_changed(self):
        """Mark the matrices as rewritten (call after writing through
        ``mat_ptrs`` directly); the next product rescans their structure."""
        self.mat_gen += 1

    def _sync_structure(self):
        want = self.mat_gen if (self.exploit_structure and self.has_dense) else -1
        if want == self._structure_gen:
            return
        if want == -1:
            check(_lib.lib().pq3d_dense_set_structure(self.h, 0), "set_structure")
        else:
            check(_lib.lib().pq3d_dense_scan_structure(self.h, None), "scan_structure")
        self._structure_gen = want

    @property
    def structure(self):
        """(zero-matrix mask, whether the product skips Bs and Bd)."""
        self._sync_structure()
        m, sk = ctypes.c_int(), ctypes.c_int()
        check(_lib.lib().pq3d_dense_get_structure(self.h, ctypes.byref(m), ctypes.byref(sk)), "get_structure")
        return m.value, bool(sk.value)

    def matvec(self):
        self._sync_structure()
        check(_lib.lib().pq3d_dense_matvec(self.h), "matvec")

    def time_matvec(self, reps=20):
        self._sync_structure()
        f = ctypes.c_float()
        check(_lib.lib().pq3d_dense_time_matvec(self.h, int(reps), ctypes.byref(f)), "time_matvec")
        return f.value

    def time_product(self, reps=5):
        """ms per traction product including the split-K partial-sum pass."""
        self._sync_structure()
        f = ctypes.c_float()
        check(_lib.lib().pq3d_dense_time_product(self.h, int(reps), ctypes.byref(f)), "time_product")
        return f.value

    @property
    def kc(self):
        """Split-K width of the dense product (source columns per tile)."""
        kc, nc = ctypes.c_int(), ctypes.c_int()
        check(_lib.lib().pq3d_dense_get_kc(self.h, ctypes.byref(kc), ctypes.byref(nc)), "get_kc")
        return kc.value

    def set_kc(self, kc):
        check(_lib.lib().pq3d_dense_set_kc(self.h, int(kc)), "set_kc")

    def step(self, dttry):
        self._sync_structure()
        h = ctypes.c_double()
        dtn = ctypes.c_double()
        nr = ctypes.c_int()
        rc = _lib.lib().pq3d_dense_step(self.h, float(dttry), ctypes.byref(h), ctypes.byref(dtn), ctypes.byref(nr))
        if rc == 1:
            return None, None, nr.value
        check(rc, "pq3d_dense_step")
        return h.value, dtn.value, nr.value

    def run(self, nsteps, graph_steps=8, records=False):
        self._sync_structure()
        st = ctypes.c_int()
        rec = np.zeros((nsteps, 4)) if records else None
        check(_lib.lib().pq3d_dense_run(self.h, int(nsteps), int(graph_steps),
                                        _lib.ptr(rec) if records else None, ctypes.byref(st)), "pq3d_dense_run")
        return st.value, rec

    def state(self):
        out = np.zeros(6)
        nr = ctypes.c_int()
        stt = ctypes.c_int()
        steps = ctypes.c_longlong()
        check(_lib.lib().pq3d_dense_state(self.h, _lib.ptr(out), ctypes.byref(nr), ctypes.byref(stt),
                                          ctypes.byref(steps)), "state")
        return {"time": out[0], "dttry": out[1], "h": out[2], "dtnext": out[3], "maxV": out[4], "err2": out[5],
                "nrjct": nr.value, "status": stt.value, "steps": steps.value}

    def reset_ctrl(self):
        """Clear a failure (status, grid barrier, reduction slots); time, step
        size and counters are kept."""
        check(_lib.lib().pq3d_dense_reset_ctrl(self.h), "reset_ctrl")

    def device_bytes(self):
        """Device memory this simulation holds."""
        b = ctypes.c_double()
        check(_lib.lib().pq3d_dense_device_bytes(self.h, ctypes.byref(b)), "device_bytes")
        return b.value

    def errors(self):
        """(Relerrormax1, Relerrormax2) of the last accepted step, or of the
        last ``dp45`` attempt if one came after it."""
        e1, e2 = ctypes.c_double(), ctypes.c_double()
        check(_lib.lib().pq3d_dense_errors(self.h, ctypes.byref(e1), ctypes.byref(e2)), "errors")
        return e1.value, e2.value

    def field_ptr(self, name):
        """Device address of field ``name`` (n doubles)."""
        p = ctypes.c_void_p()
        check(_lib.lib().pq3d_dense_field_ptr(self.h, FIELD_ID[name], ctypes.byref(p)), "field_ptr")
        return p.value

    def derivative(self, ptrs4, out_ptr):
        """RHS at the device vectors ``ptrs4`` (Tno, Tt1o, Tt2o, state) with the
        frozen traction rates into the 4 x n device buffer ``out_ptr``."""
        check(_lib.lib().pq3d_dense_derivative(self.h, *[ctypes.c_void_p(p) for p in ptrs4],
                                               ctypes.c_void_p(out_ptr)), "derivative")

    def dp45(self, h, out_ptr, gpu_semantics=True):
        """One uncommitted attempt of step h; the solution into the 4 x n device
        buffer ``out_ptr``; returns (condition1, condition2, hnew1, hnew2)."""
        out = np.zeros(4)
        check(_lib.lib().pq3d_dense_dp45(self.h, float(h), int(bool(gpu_semantics)), ctypes.c_void_p(out_ptr),
                                         _lib.ptr(out)), "dp45")
        return tuple(float(x) for x in out)

    def close(self):
        if self.h:
            _lib.lib().pq3d_dense_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceField:
    """A device-resident state vector; ``.get()`` copies it to the host."""

    def __init__(self, sim, name):
        self._sim = sim
        self._name = name

    def get(self):
        return self._sim.get(self._name)

    def set(self, arr):
        self._sim.set(self._name, arr)

    @property
    def device_ptr(self):
        return self._sim.field_ptr(self._name)

    def __array__(self, dtype=None, copy=None):
        a = self.get()
        return a if dtype is None else a.astype(dtype)

    def __len__(self):
        return self._sim.n

    @property
    def shape(self):
        return (self._sim.n,)


class DeviceArray:
    """n doubles in device memory (what the reference's CuPy arrays are to its
    callers: results of ``derivative_gpu`` and of the Dormand-Prince attempt);
    ``.get()`` copies them to the host.  Views of one buffer keep it alive."""

    def __init__(self, n, buf=None, offset=0):
        self.n = int(n)
        self._buf = buf if buf is not None else _lib.DeviceBuffer(8 * max(self.n, 1))
        self._off = int(offset)

    @classmethod
    def block(cls, n, k):
        """k arrays of n doubles, back to back in one allocation."""
        buf = _lib.DeviceBuffer(8 * max(int(n), 1) * k)
        return [cls(n, buf, 8 * int(n) * j) for j in range(k)]

    @classmethod
    def from_host(cls, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        d = cls(len(a))
        d._buf.upload(a, d._off)
        return d

    @property
    def device_ptr(self):
        return self._buf.ptr + self._off

    def get(self):
        return self._buf.download(np.empty(self.n), self._off)

    def __array__(self, dtype=None, copy=None):
        a = self.get()
        return a if dtype is None else a.astype(dtype)

    def __len__(self):
        return self.n

    @property
    def shape(self):
        return (self.n,)


def _device_operand(x, n):
    """(device address, keep-alive) of a state vector given as a device field,
    a device array or host values."""
    if isinstance(x, (DeviceField, DeviceArray)):
        if len(x) != n:
            raise ValueError(f"vector of length {len(x)}, expected {n}")
        return x.device_ptr, x
    a = np.ascontiguousarray(np.broadcast_to(np.asarray(x, dtype=np.float64), (n,)))
    d = DeviceArray.from_host(a)
    return d.device_ptr, d


_GPU_ATTRS = {"Tno_gpu": "Tno", "Tt1o_gpu": "Tt1o", "Tt2o_gpu": "Tt2o", "state_gpu": "state",
              "slipv1_gpu": "slipv1", "slipv2_gpu": "slipv2", "slipv_gpu": "slipv", "slip1_gpu": "slip1",
              "slip2_gpu": "slip2", "slip_gpu": "slip", "Tt_gpu": "Tt", "rake_gpu": "rake",
              "rake0_gpu": "rake0", "fric_gpu": "fric", "a_gpu": "a", "b_gpu": "b", "dc_gpu": "dc",
              "f0_gpu": "f0", "shear_loading_gpu": "shear_loading", "normal_loading_gpu": "normal_loading",
              "P_gpu": "P", "dPdt0_gpu": "dPdt0",
              # the frozen traction rates of the last step (set by simu_forwardGPU, QDsim_gpu.py:951-957)
              "AdotV1": "AdotV1", "AdotV2": "AdotV2", "dsigmadt": "dsigmadt"}


class QDsim:
    """Dense quasi-dynamic earthquake-cycle simulator (QDsim_gpu.py:119-186)."""

    def __init__(self, elelst, nodelst, fnamePara, device=True):
        """The reference's constructor.  ``device=False`` (an extension) stops
        after the host-side initial state (parameters, friction, stresses) and
        creates no device simulation and no Green's functions."""
        i = fnamePara.rfind("/")
        self.dirname = fnamePara[:i] if i != -1 else fnamePara
        self.Para0 = readPara0(fnamePara)
        jud_ele_order = False
        eleVec, xg = readmsh.get_eleVec(nodelst, elelst, jud_ele_order)
        self.eleVec = eleVec
        self.elelst = np.asarray(elelst)
        self.nodelst = np.asarray(nodelst, dtype=np.float64)
        self.xg = xg
        self.maxslipvque = deque(maxlen=20)
        self.val_nv = 0
        P = self.Para0
        self.mu = float(P["Shear modulus"])
        self.lambda_ = float(P["Lame constants"])
        self.density = float(P["Rock density"])
        self.htry = 1e-3
        self.Cs = sqrt(self.mu / self.density)
        self.time = 0
        self.halfspace_jud = P["Half space"] == "True"
        self.InputHetoparamter = P["InputHetoparamter"] == "True"
        self.useGPU = P.get("GPU", "True") == "True"
        self.Ifdila = P.get("If Dilatancy", "False") == "True"
        self.DilatancyC = float(P.get("Dilatancy coefficient", 0.0))
        self.Chyd = float(P.get("Hydraulic diffusivity", 0.0))
        self.hs = float(P.get("Actively shearing zone thickness", 0.0))
        self.EPermeability = float(P.get("Effective compressibility", 0.0))
        prec = P.get("GF precision", "fp64").lower()
        self.mat_type = {"fp64": _lib.MAT_F64, "fp32": _lib.MAT_F32, "bf16": _lib.MAT_BF16}[prec]
        YoungsM = self.mu * (3.0 * self.lambda_ + 2.0 * self.mu) / (self.lambda_ + self.mu)
        possonratio = self.lambda_ / 2.0 / (self.lambda_ + self.mu)
        print("Cs", self.Cs)
        print("First Lamé constants", self.lambda_)
        print("Shear Modulus", self.mu)
        print("Youngs Modulus", YoungsM)
        print("Poissons ratio", possonratio)
        self.Relerrormax1_last = 0
        self.Relerrormax2_last = 0
        if not device:
            self.Init_condition()
            return
        self.sim = DenseSim(len(elelst), self.mat_type)
        self.Init_condition()
        self.calc_corefunc()

    # ---------------------------------------------------------------- GF
    def calc_corefunc(self):
        """Load the cached Green's functions or assemble them on the GPU
        (QDsim_gpu.py:188-245)."""
        directory = self.Para0["Corefunc directory"]
        names = greenfunc.MATRIX_NAMES
        have = os.path.exists(directory) and all(os.path.exists(os.path.join(directory, n + ".npy"))
                                                 for n in names)
        if not os.path.exists(directory):
            os.mkdir(directory)
        if have:
            print("Start to load core functions...")
            for k, n in enumerate(names):
                self.sim.upload_matrix(k, np.load(os.path.join(directory, n + ".npy"), allow_pickle=False))
            print("Core functions load completed.")
        else:
            print("Start to calculate core functions...")
            P1, P2, P3 = greenfunc.source_vertices(self.nodelst, self.elelst)
            if self.mat_type == _lib.MAT_BF16:
                tmp = DenseSim(self.sim.n, _lib.MAT_F32)
                tmp.assemble(P1, P2, P3, self.xg, self.eleVec, self.mu, self.lambda_, self.halfspace_jud)
                for k in range(6):
                    self.sim.upload_matrix(k, tmp.download_matrix(k))
                tmp.close()
            else:
                self.sim.assemble(P1, P2, P3, self.xg, self.eleVec, self.mu, self.lambda_, self.halfspace_jud)
            if self.Para0.get("save Corefunc", "True") == "True":
                for k, n in enumerate(names):
                    np.save(os.path.join(directory, n), self.sim.download_matrix(k))
            print("Core functions computation completed")

    def __getattr__(self, name):
        # host copies of the matrices on demand (A1s, A2s, Bs, A1d, A2d, Bd)
        if name in greenfunc.MATRIX_NAMES and "sim" in self.__dict__:
            return self.sim.download_matrix(greenfunc.MATRIX_NAMES.index(name))
        if name in _GPU_ATTRS and self.__dict__.get("_gpu_ready"):
            return DeviceField(self.sim, _GPU_ATTRS[name])
        raise AttributeError(name)

    # ------------------------------------------------------ initial state
    def calc_nucleaszie_cohesivezone(self):
        """Element sizes, critical nucleation size and cohesive zone
        (QDsim_gpu.py:282-317)."""
        e = np.asarray(self.elelst) - 1
        P1, P2, P3 = self.nodelst[e[:, 0]], self.nodelst[e[:, 1]], self.nodelst[e[:, 2]]
        elesize = np.max(np.stack([np.linalg.norm(P1 - P2, axis=1), np.linalg.norm(P1 - P3, axis=1),
                                   np.linalg.norm(P2 - P3, axis=1)], 1), axis=1)
        maxsize = float(elesize.max())
        self.maxsize = maxsize
        self.ave_elesize = np.mean(elesize)
        b = np.max(self.b)
        a = np.min(self.a)
        sigma = np.mean(self.Tno * 1e6)
        L = np.max(self.dc)
        print("L:", L)
        self.hRA = 2.0 / np.pi * self.mu * b * L / (b - a) / (b - a) / sigma
        self.hRA = self.hRA * np.pi * np.pi / 4.0
        self.A0 = 9.0 * np.pi / 32 * self.mu * L / (b * sigma)
        print("maximum element size", maxsize)
        print("average elesize", self.ave_elesize)
        print("Critical nucleation size", self.hRA)
        print("Cohesive zone:", self.A0)
        return maxsize

    def read_parameter(self, fname):
        """10-column per-element initial state (QDsim_gpu.py:714-770)."""
        values = np.loadtxt(os.path.join(self.dirname, fname), ndmin=2)
        Ncell = self.eleVec.shape[0]
        self.rake = values[:Ncell, 0]
        self.a = values[:Ncell, 1]
        self.b = values[:Ncell, 2]
        self.dc = values[:Ncell, 3]
        self.f0 = values[:Ncell, 4]
        self.Tt = values[:Ncell, 5]
        self.Tno = values[:Ncell, 6]
        self.slipv = values[:Ncell, 7]
        self.shear_loading = values[:Ncell, 8]
        self.normal_loading = values[:Ncell, 9]
        x = np.cos(self.rake)
        y = np.sin(self.rake)
        self.slipv1 = self.slipv * x + 1e-16
        self.slipv2 = self.slipv * y + 1e-16
        self.vec_Tra = np.array([x, y]).transpose()
        self.Tt = (self.Tno - self.P * 1e-6) * self.a * np.arcsinh(
            self.slipv / (2.0 * self.V0) * np.exp((self.f0 + self.b * np.log(self.V0 / self.slipvC)) / self.a))
        self.Tt1o = self.Tt * x
        self.Tt2o = self.Tt * y
        self.fric = self.Tt / (self.Tno - self.P * 1e-6)
        self.state = np.log(np.sinh(self.Tt / (self.Tno - self.P * 1e-6) / self.a) * 2.0 * self.V0 / self.slipv) * self.a

    def Init_condition(self):
        """Initial stress, friction parameters and state (QDsim_gpu.py:319-609):
        tractions from a depth-scaled principal stress state (or the constant
        vertical stress), then either the per-element parameter file or the
        zone rules of ``_zone_init``."""
        P = self.Para0
        N = len(self.eleVec)
        flag = lambda key: P[key] == "True"  # noqa: E731
        self.fix_Tn = flag("Fix_Tn")
        scale_v = float(P["ssv_scale"])
        scale_h1 = float(P["ssh1_scale"])
        scale_h2 = float(P["ssh2_scale"])
        sigma_v0 = float(P["Vertical principal stress value"])
        self.P0 = 0
        self.P = np.zeros(N)
        self.dPdt0 = np.zeros(N)
        self.local_index = np.arange(0, N, 1)
        if self.Ifdila:
            # pore-pressure grid and profiles (QDsim_gpu.py:336-349), pressures in Pa
            nodes, stretch = 60, 1e-2
            self.P0 = float(P["Constant porepressure"]) * 1e6
            self.P = np.ones(N) * float(P["Initial porepressure"]) * 1e6
            self.yp = np.logspace(start=-3, stop=10, num=nodes, base=10)
            self.zp = np.log(1 + self.yp / stretch)
            self.Parr = np.ones([N, nodes]) * float(P["Initial porepressure"]) * 1e6
        normal_from_tensor = flag("Normal traction solved from stress tensor")
        depth_scaled = flag("Vertical principal stress value varies with depth")
        # vertical principal stress per element: constant, or growing linearly
        # with depth down to the turning depth (0.2 at the surface, 1 below)
        if depth_scaled:
            ramp = np.minimum(-self.xg[:, 2] / float(P["Turnning depth"]) + 0.2, 1.0)
            sv = sigma_v0 * ramp * scale_v
        else:
            sv = np.full(N, sigma_v0 * scale_v)
        # compressive principal stresses; the horizontal pair rotated by the
        # angle of the first one from the x axis: S = R diag(h1, h2) R^T
        phi = np.pi / 180. * float(P["Angle between ssh1 and X-axis"])
        c, s = np.cos(phi), np.sin(phi)
        h1, h2, vz = -sv * scale_h1, -sv * scale_h2, -sv
        Sxx = (c * h1) * c + (-s * h2) * -s
        Sxy = (c * h1) * s + (-s * h2) * c
        Syx = (s * h1) * c + (c * h2) * -s
        Syy = (s * h1) * s + (c * h2) * c
        nrm = self.eleVec[:, 6:9]
        t = np.stack([Sxx * nrm[:, 0] + Sxy * nrm[:, 1] + 0.0 * nrm[:, 2],
                      Syx * nrm[:, 0] + Syy * nrm[:, 1] + 0.0 * nrm[:, 2],
                      0.0 * nrm[:, 0] + 0.0 * nrm[:, 1] + vz * nrm[:, 2]], axis=1)
        ev = self.eleVec
        self.Tt1o = t[:, 0] * ev[:, 0] + t[:, 1] * ev[:, 1] + t[:, 2] * ev[:, 2]
        self.Tt2o = t[:, 0] * ev[:, 3] + t[:, 1] * ev[:, 4] + t[:, 2] * ev[:, 5]
        self.Tno = t[:, 0] * ev[:, 6] + t[:, 1] * ev[:, 7] + t[:, 2] * ev[:, 8]
        if not normal_from_tensor:
            self.Tno = vz.copy()
        self.Tno = np.abs(self.Tno)
        self.Tt = np.sqrt(self.Tt1o * self.Tt1o + self.Tt2o * self.Tt2o)
        with np.errstate(invalid="ignore", divide="ignore"):
            cos_r = self.Tt1o / self.Tt
            sin_r = self.Tt2o / self.Tt
        shear_from_tensor = flag("Shear traction solved from stress tensor")
        if flag("Rake solved from stress tensor"):
            self.rake = np.arctan2(sin_r, cos_r)
        else:
            self.rake = np.ones(N) * float(P["Fix_rake"]) / 180.0 * np.pi
        if not normal_from_tensor and not depth_scaled:
            self.Tno = np.ones(N) * sigma_v0
        self.vec_Tra = np.array([cos_r, sin_r]).transpose()
        self.Vpl_con = float(P["Plate loading rate"])
        self.slipvC = np.ones(len(self.xg)) * self.Vpl_con
        self.shear_loading = np.zeros(N)
        self.normal_loading = np.zeros(N)
        self.V0 = float(P["Reference slip rate"])
        self.dc = np.ones(N) * 0.02
        self.f0 = np.ones(N) * float(P["Reference friction coefficient"])
        self.a = np.zeros(N)
        self.b = np.ones(N) * 0.03
        self.slipv1 = np.zeros(N)
        self.slipv2 = np.zeros(N)
        # the reference aliases slipv and slipvC here (QDsim_gpu.py:446): writes
        # to slipv in the nucleation patch also change slipvC
        self.slipv = self.slipvC
        self.slip1 = np.zeros(N)
        self.slip2 = np.zeros(N)
        self.slip = np.zeros(N)
        self.arriT = np.ones(N) * 1e9
        if self.InputHetoparamter:
            self.read_parameter(P["Inputparamter file"])
            self.maxslipv0 = np.max(self.slipv)
            self.rake0 = np.copy(self.rake)
            self.calc_nucleaszie_cohesivezone()
            return
        self._zone_init(shear_from_tensor)

    @staticmethod
    def _nearest(points, targets, chunk=4096):
        """Distance from every point to the nearest target (sqrt of the least
        squared distance: the same value as the least distance)."""
        out = np.empty(len(points))
        for s in range(0, len(points), chunk):
            d = points[s:s + chunk, None, :] - targets[None, :, :]
            out[s:s + chunk] = np.sqrt(np.min(np.sum(d * d, axis=2), axis=1))
        return out

    def _zone_init(self, shear_from_tensor):
        """Friction zones from the distance to the mesh boundary
        (QDsim_gpu.py:473-609): velocity-strengthening within a band along the
        buried edges (and a narrower one along the free surface), a linear
        transition band, velocity-weakening inside; an optional nucleation
        disc; then steady-state traction and state at the initial slip rate."""
        P = self.Para0
        num = lambda key: float(P[key])  # noqa: E731
        _, edge_nodes = find_boundary_edges_and_nodes(self.elelst)
        edge_xyz = self.nodelst[edge_nodes - 1]
        at_surface = np.abs(edge_xyz[:, 2] - 0.0) < 1e-5
        buried = edge_xyz[~at_surface]
        surface = edge_xyz[at_surface]
        band = num("Widths of VS region")
        band_surface = num("Widths of surface VS region")
        ramp = num("Transition region from VS to VW region")
        vs = [num(f"Rate-and-state parameters {k} in VS region") for k in ("a", "b")] + \
            [num("Characteristic slip distance in VS region")]
        vw = [num(f"Rate-and-state parameters {k} in VW region") for k in ("a", "b")] + \
            [num("Characteristic slip distance in VW region")]
        nuc = [num(f"Rate-and-state parameters {k} in nucleation region") for k in ("a", "b")] + \
            [num("Characteristic slip distance in nucleation region")]
        d_edge = self._nearest(self.xg, buried)
        u_edge = d_edge / band                    # < 1 inside the buried-edge band
        u_surf = u_edge
        u_ramp = (d_edge - band) / ramp
        if len(surface) > 10:
            d_surf = self._nearest(self.xg, surface)
            u_surf = d_surf / band_surface
            u_ramp = np.minimum(d_edge - band, d_surf - band_surface) / ramp
        in_vs = (u_edge < 1.0) | (u_surf < 1.0)
        in_ramp = ~in_vs & (u_ramp < 1.0)
        in_vw = ~in_vs & ~in_ramp
        a_ramp_mode = P["ChangefriA"] == "True"   # the transition tapers a (else b)
        self.a = np.where(in_vs, vs[0], np.where(in_vw, vw[0], vs[0] - (vs[0] - vw[0]) * u_ramp if a_ramp_mode
                                                  else vs[0]))
        self.b = np.where(in_vs, vs[1], np.where(in_vw, vw[1], vs[1] if a_ramp_mode
                                                  else vs[1] - (vs[1] - vw[1]) * u_ramp))
        self.dc = np.where(in_vw, vw[2], vs[2])
        if P["Set_nucleation"] == "True":
            centre = np.array([num("Nuclea_posx"), num("Nuclea_posy"), num("Nuclea_posz")])
            disc = np.linalg.norm(self.xg - centre, axis=1) < num("Radius of nucleation")
            self.slipv[disc] = num("Initial slip rate in nucleation region")
            self.dc = np.where(disc, nuc[2], self.dc)
            self.a = np.where(disc, nuc[0], self.a)
            self.b = np.where(disc, nuc[1], self.b)
        # slip direction in global coordinates: frame^T (cos rake, sin rake, 0)
        frames = self.eleVec.reshape(-1, 3, 3)
        slip_dir = np.stack([np.cos(self.rake), np.sin(self.rake), np.zeros(len(self.rake))], axis=1)
        self.T_globalarr = np.einsum("nki,nk->ni", frames, slip_dir)
        self.slipv1 = self.slipv * np.cos(self.rake) + 1e-16
        self.slipv2 = self.slipv * np.sin(self.rake) + 1e-16
        if not shear_from_tensor:
            self.Tt = self.Tno * self.a * np.arcsinh(
                self.slipv / (2.0 * self.V0) * np.exp((self.f0 + self.b * np.log(self.V0 / self.slipvC)) / self.a))
        self.vec_Tra = np.array([np.cos(self.rake), np.sin(self.rake)]).transpose()
        self.fric = self.Tt / self.Tno
        self.state = np.log(np.sinh(self.Tt / self.Tno / self.a) * 2.0 * self.V0 / self.slipv) * self.a
        self.Tt1o = self.Tt * np.cos(self.rake)
        self.Tt2o = self.Tt * np.sin(self.rake)
        self.maxslipv0 = np.max(self.slipv)
        self.rake0 = np.copy(self.rake)
        self.calc_nucleaszie_cohesivezone()

    # ----------------------------------------------------------- stepping
    def initGPUvariable(self):
        """Upload the state and parameters (QDsim_gpu.py:893-939)."""
        s = self.sim
        for name in ("Tno", "Tt1o", "Tt2o", "state", "slipv1", "slipv2", "slipv", "slip1", "slip2", "slip", "Tt",
                     "rake", "fric", "a", "b", "dc", "f0", "shear_loading", "normal_loading", "P", "dPdt0"):
            s.set(name, getattr(self, name))
        s.set("rake0", self.rake)
        s.set_params(self.V0, self.Vpl_con, self.mu / (2.0 * self.Cs), self.fix_Tn, True)
        s.set_time(self.time)
        if s.has_dense and autotune.enabled():
            autotune.ensure_dense(s)   # split-K tile width for this mesh and device
        if self.Ifdila:
            s.set_dilatancy(self.zp, self.Parr, self.Chyd, self.DilatancyC, self.hs, self.EPermeability, self.P0,
                            self.V0)
        self._gpu_ready = True

    def _step(self, dttry, gpu_semantics):
        if not self.__dict__.get("_gpu_ready"):
            self.initGPUvariable()
        self.sim.set_params(self.V0, self.Vpl_con, self.mu / (2.0 * self.Cs), self.fix_Tn, gpu_semantics)
        self.sim.set_time(self.time)
        h, dtnext, nrjct = self.sim.step(dttry)
        if h is None:
            print("error: dt is too small")
            sys.exit()
        st = self.sim.state()
        self.Relerrormax1, self.Relerrormax2 = self.sim.errors()
        self.Relerrormax1_last, self.Relerrormax2_last = self.Relerrormax1, self.Relerrormax2
        self.time = self.time + h
        self.maxslipv0 = st["maxV"]
        return h, dtnext

    # (field name in the file, attribute, scale, number format): the cell data
    # of the reference's legacy VTK output, in its order (QDsim_gpu.py:1927-2073)
    _VTK_FIELDS = (("Normal_[MPa]", "Tno", 1.0, "%f"), ("Pore_pressure[MPa]", "P", 1e-6, "%f"),
                   ("Shear_[MPa]", "Tt", 1.0, "%f"), ("Shear_1[MPa]", "Tt1o", 1.0, "%f"),
                   ("Shear_2[MPa]", "Tt2o", 1.0, "%f"), ("rake[Degree]", "rake", 180. / np.pi, "%f"),
                   ("state", "state", 1.0, "%f"), ("Slipv[m/s]", "slipv", 1.0, "%.20f"),
                   ("Slipv1[m/s]", "slipv1", 1.0, "%.20f"), ("Slipv2[m/s]", "slipv2", 1.0, "%.20f"),
                   ("a", "a", 1.0, "%f"), ("b", "b", 1.0, "%f"), ("a-b", None, 1.0, "%f"), ("dc", "dc", 1.0, "%f"),
                   ("fric", "fric", 1.0, "%f"), ("slip", "slip", 1.0, "%.20f"), ("slip1", "slip1", 1.0, "%.20f"),
                   ("slip2", "slip2", 1.0, "%.20f"))

    def ouputVTK(self, fname):
        """Legacy-VTK snapshot of the fault in the reference's layout
        (QDsim_gpu.py:1927-2073): nodes, triangles, and per-cell fields.  The
        values are the current device state (the reference's GPU driver
        writes some fields from host copies it does not refresh)."""
        if self.__dict__.get("_gpu_ready"):
            for name in ("Tno", "P", "Tt", "Tt1o", "Tt2o", "rake", "state", "slipv", "slipv1", "slipv2", "fric",
                         "slip", "slip1", "slip2"):
                setattr(self, name, self.sim.get(name))
        fields = []
        for label, attr, scale, fmt in self._VTK_FIELDS:
            vals = (np.asarray(self.a) - np.asarray(self.b)) if attr is None else \
                np.asarray(getattr(self, attr), dtype=np.float64) * scale
            fields.append((label, vals, fmt))
        outputs.write_legacy_vtk(fname, self.nodelst, np.asarray(self.elelst, dtype=np.int64) - 1, fields)

    def outputtxt(self, fname):
        """Surface grids of slip rate and traction (QDsim_gpu.py:2085-2151):
        slipv, slipv1, slipv2 and Tt interpolated with SciPy's cubic
        ``griddata`` from the element centroids' (x, z) onto a 500 x 300 grid
        inset by the largest element size, written under ``out_txt/`` (in the
        working directory) as ``X_grid.txt``, ``Y_grid.txt`` (rows of '%f ')
        and ``<fname>slipv.txt``, ``slipv1``, ``slipv2``, ``Traction`` (rows of
        '%.5f '; points outside the centroids' hull are nan).  The values
        are the current device state."""
        from scipy.interpolate import griddata  # host post-processing only
        if self.__dict__.get("_gpu_ready"):
            for name in ("slipv", "slipv1", "slipv2", "Tt"):
                setattr(self, name, self.sim.get(name))
        directory = "out_txt"
        os.makedirs(directory, exist_ok=True)
        xg = np.asarray(self.xg, dtype=np.float64)
        xmin, xmax = np.min(xg[:, 0]), np.max(xg[:, 0])
        zmin, zmax = np.min(xg[:, 2]), np.max(xg[:, 2])
        X1 = np.linspace(xmin + self.maxsize, xmax - self.maxsize, 500)
        Y1 = np.linspace(zmin + self.maxsize, zmax - self.maxsize, 300)
        X_grid, Y_grid = np.meshgrid(X1, Y1)
        pts = np.column_stack((X_grid.ravel(), Y_grid.ravel()))

        def write(path, grid, fmt):
            with open(path, "w") as f:
                for row in grid:
                    f.write("".join(fmt % v for v in row) + "\n")

        write(os.path.join(directory, "X_grid.txt"), X_grid, "%f ")
        write(os.path.join(directory, "Y_grid.txt"), Y_grid, "%f ")
        for attr, tag in (("slipv", "slipv"), ("slipv1", "slipv1"), ("slipv2", "slipv2"), ("Tt", "Traction")):
            vals = np.asarray(getattr(self, attr), dtype=np.float64)
            grid = griddata(xg[:, [0, 2]], vals, pts, method="cubic").reshape(len(Y1), len(X1))
            write(os.path.join(directory, fname + tag + ".txt"), grid, "%.5f ")

    def Calc_P_implicit(self, dt):
        """One pore-pressure update on the device (QDsim_gpu.py:628-671); the
        step drivers already run it after every accepted step."""
        if not self.__dict__.get("_gpu_ready"):
            self.initGPUvariable()
        self.sim.dilatancy_update(dt)
        self.P = self.sim.get("P")
        self.dPdt0 = self.sim.get("dPdt0")
        self.Parr = self.sim.get_dilatancy()

    # ---- the stepping pieces, one call at a time
    def derivative_gpu(self, Tno, Tt1o, Tt2o, state):
        """The rate-and-state RHS (QDsim_gpu.py:830-891) at the given state
        vectors (device fields such as ``Tno_gpu``, ``DeviceArray`` results or
        host arrays) with the frozen traction rates of the last step (the
        ``AdotV1``, ``AdotV2``, ``dsigmadt`` fields).  Returns the reference's
        4-tuple (dstate/dt, dsigma/dt, dtau1/dt, dtau2/dt) as device arrays."""
        if not self.__dict__.get("_gpu_ready"):
            self.initGPUvariable()
        n = self.sim.n
        ops = [_device_operand(x, n) for x in (Tno, Tt1o, Tt2o, state)]
        out = DeviceArray.block(n, 4)
        self.sim.derivative([p for p, _ in ops], out[0].device_ptr)
        return tuple(out)

    def RungeKutte_solve_Dormand_Prince_gpu(self, h):
        """One Dormand-Prince attempt of step h from the current device state
        with the frozen traction rates, nothing committed (QDsim_gpu.py:1128-1262):
        returns (Tno_yhk, Tt1o_yhk, Tt2o_yhk, state_yhk, condition1, condition2,
        hnew1, hnew2), the vectors as device arrays; sets Relerrormax1/2.  The
        normal-traction error is dropped as in the reference's GPU method."""
        return self._attempt(h, True)

    def _attempt(self, h, gpu_semantics):
        if not self.__dict__.get("_gpu_ready"):
            self.initGPUvariable()
        out = DeviceArray.block(self.sim.n, 4)
        c1, c2, hn1, hn2 = self.sim.dp45(h, out[0].device_ptr, gpu_semantics)
        self.Relerrormax1, self.Relerrormax2 = self.sim.errors()
        return out[0], out[1], out[2], out[3], c1, c2, hn1, hn2

    def derivative(self, Tno, Tt1o, Tt2o, state):
        """The CPU branch's RHS (QDsim_gpu.py:772-828; the same formulas),
        computed on the device: host arrays in, host arrays out."""
        return tuple(x.get() for x in self.derivative_gpu(Tno, Tt1o, Tt2o, state))

    def RungeKutte_solve_Dormand_Prince(self, h):
        """The CPU branch's attempt (QDsim_gpu.py:1266-1397): as the GPU method
        but with the normal-traction error kept; host arrays out."""
        r = self._attempt(h, False)
        return tuple(x.get() for x in r[:4]) + r[4:]

    def simu_forwardGPU(self, dttry):
        """One accepted step on the device; returns (h, dtnext)."""
        return self._step(dttry, True)

    def simu_forward(self, dttry):
        """The CPU branch's step semantics, computed on the device; refreshes the
        host attributes the CPU branch maintains."""
        h, dtnext = self._step(dttry, False)
        for name in ("Tno", "Tt1o", "Tt2o", "state", "slipv1", "slipv2", "slipv", "slip1", "slip2", "slip", "Tt",
                     "rake", "fric", "AdotV1", "AdotV2", "dsigmadt"):
            setattr(self, name, self.sim.get(name))
        if self.Ifdila:
            self.P = self.sim.get("P")
            self.dPdt0 = self.sim.get("dPdt0")
            self.Parr = self.sim.get_dilatancy()
        return h, dtnext

    def run(self, nsteps, dttry=None, graph_steps=8, records=True):
        """``nsteps`` device-resident steps with no host synchronisation in
        between (an extension for throughput; the reference steps one at a
        time).  Returns the per-step log (h, dtnext, time, max slip rate)."""
        if not self.__dict__.get("_gpu_ready"):
            self.initGPUvariable()
        self.sim.set_time(self.time)
        first = None
        if dttry is not None:  # the first step takes dttry; later ones take dtnext on the device
            h, dtnext, _ = self.sim.step(dttry)
            if h is None:
                raise RuntimeError("error: dt is too small")
            first = np.array([[h, dtnext, self.time + h, self.sim.state()["maxV"]]])
            nsteps -= 1
        status, rec = self.sim.run(nsteps, graph_steps, records)
        if status != 0:
            raise RuntimeError("error: dt is too small" if status == 1 else f"device status {status}")
        st = self.sim.state()
        self.time = st["time"]
        if records and first is not None:
            rec = np.concatenate([first, rec])
        return rec
