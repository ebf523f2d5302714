"""Continuous-batching server for earthquake-cycle simulation requests.

There is no reference counterpart (the reference runs one simulation per
process, main_gpu.py); this layer turns the ensemble kernels into a service:

* every mesh that is resident keeps its FP32 traction Green's functions and a
  64-slot ensemble (state of up to 64 simulations) in HBM;
* meshes are made resident on demand when the HBM budget allows, and an idle
  mesh (no queued or running request) is evicted to make room;
* a request (mesh, friction parameters in the reference's init-file layout,
  simulated duration) is admitted into a free slot of its mesh as soon as one
  frees up -- continuous batching: new requests join while others are
  mid-simulation, and finished ones leave;
* each serving iteration enqueues a burst of global steps on every active
  mesh's own stream, so the meshes' work overlaps on the device, then waits,
  retires finished requests and admits queued ones.

A request's result is independent of what else runs: slots are independent
columns of the GEMM and independent RK members, so a simulation gives the same
state alone and in the mix (tests/test_serving_gpu.py).
"""
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib, autotune, greenfunc, readmsh
from .QDsim_gpu import DenseSim
from .ensemble import Ensemble, load_member

YEAR = 365.0 * 24 * 3600


@dataclass
class Request:
    rid: int
    mesh: str
    cols: np.ndarray          # (n, 10) init-file columns
    duration_s: float         # simulated time to reach
    arrival: float = 0.0      # wall-clock offset at which it is submitted
    submitted: float = None
    admitted: float = None
    finished: float = None
    steps: int = 0
    status: int = 0
    result: dict = field(default_factory=dict)


class _Mesh:
    def __init__(self, name, nodes, tri, mu, lam, halfspace):
        self.name = name
        self.nodes = np.asarray(nodes, dtype=np.float64)
        self.tri = np.asarray(tri)
        self.n = len(tri)
        self.mu, self.lam, self.halfspace = mu, lam, halfspace
        self.sim = None
        self.ens = None
        self.slots = [None] * 64
        self.busy_since = None

    def hbm_bytes(self):
        """Device memory the mesh needs when resident, an upper bound of what
        the library then reports (``DenseSim.device_bytes`` +
        ``Ensemble.device_bytes``): the FP32 Green's functions, the dense
        simulation's fields and split-K partials (kc 256), and the ensemble's
        fields, operands, partials (at most 32 chunks), the split GEMM's BF16
        operand planes and panel copies of all six matrices."""
        n = self.n
        ld = (n + 3) // 4 * 4
        npad = (n + 63) // 64 * 64
        kpad = (n + 31) // 32 * 32
        gf = 6 * ld * n * 4
        dense = 27 * n * 8 + 3 * n * ((n + 255) // 256) * 8 + 4096
        ens = (31 * 64 * n * 8 + 2 * 64 * n * 4 + 32 * 3 * 64 * npad * 4 + 2 * 3 * (npad // 8) * 64 * 16
               + 6 * npad * kpad * 4 + 8192)
        return gf + dense + ens

    @property
    def resident(self):
        return self.ens is not None

    def load(self):
        ev, xg = readmsh.get_eleVec(self.nodes, self.tri, False)
        self.sim = DenseSim(self.n, _lib.MAT_F32)
        self.sim.assemble(*greenfunc.source_vertices(self.nodes, self.tri), xg, ev, self.mu, self.lam,
                          self.halfspace)
        self.ens = Ensemble(self.n, 64, matrices_from=self.sim)
        if autotune.enabled():
            autotune.ensure_ensemble(self.ens)   # split-K tile width for this mesh and device
        for m in range(64):
            self.ens.reset_member(m, active=False)
        self.slots = [None] * 64

    def unload(self):
        if self.ens is not None:
            self.ens.close()
            self.ens = None
        if self.sim is not None:
            self.sim.close()
            self.sim = None

    # slots in member-subtile order (members 0, 4, ..., 60, then 1, 5, ...):
    # the ensemble GEMM skips the MFMAs of the subtiles above the highest one
    # holding an active member, so a lightly loaded mesh uses few subtiles
    SLOT_ORDER = tuple((s % 16) * 4 + s // 16 for s in range(64))

    def free_slot(self):
        for m in self.SLOT_ORDER:
            if self.slots[m] is None:
                return m
        return None

    def active(self):
        return sum(r is not None for r in self.slots)


class Server:
    """Continuous-batching simulation server on one GPU."""

    def __init__(self, meshes, hbm_budget_bytes, V0=1e-6, Vpl=1e-9, mu=32038120320.0, lam=32038120320.0,
                 Cs=3464.0, halfspace=True, steps_per_iteration=64, graph_steps=16):
        self.meshes = {k: _Mesh(k, nodes, tri, mu, lam, halfspace) for k, (nodes, tri) in meshes.items()}
        self.budget = hbm_budget_bytes
        self.V0, self.Vpl, self.eta = V0, Vpl, mu / (2.0 * Cs)
        self.K = steps_per_iteration
        self.graph_steps = graph_steps
        self.queue = []
        self.done = []
        self.occupancy = []
        self.member_steps = 0
        self.log = []

    # ------------------------------------------------------------- residency
    def _resident_bytes(self):
        return sum(m.hbm_bytes() for m in self.meshes.values() if m.resident)

    def _make_resident(self, mesh):
        need = mesh.hbm_bytes()
        if need > self.budget:
            raise ValueError(f"mesh {mesh.name} needs {need / 1e9:.1f} GB, over the HBM budget")
        while self._resident_bytes() + need > self.budget:
            idle = [m for m in self.meshes.values() if m.resident and m.active() == 0
                    and not any(r.mesh == m.name for r in self.queue if r.submitted is not None)]
            if not idle:
                return False
            idle[0].unload()
        mesh.load()
        mesh.ens.set_params(self.V0, self.Vpl, self.eta)
        return True

    # ------------------------------------------------------------- serving
    def submit(self, req, now):
        req.submitted = now
        self.queue.append(req)

    def _admit(self, now):
        still = []
        for req in self.queue:
            mesh = self.meshes[req.mesh]
            if not mesh.resident and not self._make_resident(mesh):
                still.append(req)
                continue
            m = mesh.free_slot()
            if m is None:
                still.append(req)
                continue
            load_member(mesh.ens, m, req.cols, self.V0, self.Vpl, t_end=req.duration_s)
            mesh.slots[m] = req
            req.admitted = now
        self.queue = still

    def _retire(self, mesh, now):
        mem = mesh.ens.members()
        for m, req in enumerate(mesh.slots):
            if req is None:
                continue
            if mem["finished"][m] or mem["status"][m] != 0:
                req.finished = now
                req.steps = int(mem["steps"][m])
                req.status = int(mem["status"][m])
                req.result = {"time": float(mem["time"][m]),
                              "state": mesh.ens.get_member(m, "state"),
                              "slip": mesh.ens.get_member(m, "slip"),
                              "slipv": mesh.ens.get_member(m, "slipv")}
                mesh.ens.reset_member(m, active=False)
                mesh.slots[m] = None
                self.done.append(req)

    def _count_steps(self, mesh, before):
        mem = mesh.ens.members()
        after = mem["steps"]
        self.member_steps += int(np.sum(np.maximum(after - before, 0)))

    def run(self, requests, max_wall_s=600.0):
        """Serve `requests` (each with an `arrival` offset in seconds) until all
        have finished.  Returns the metrics dict."""
        pending = sorted(requests, key=lambda r: r.arrival)
        t0 = time.perf_counter()
        i = 0
        while True:
            now = time.perf_counter() - t0
            while i < len(pending) and pending[i].arrival <= now:
                self.submit(pending[i], pending[i].arrival)
                i += 1
            self._admit(now)
            running = [m for m in self.meshes.values() if m.resident and m.active() > 0]
            if not running:
                if i >= len(pending) and not self.queue:
                    break
                if i < len(pending):
                    time.sleep(max(0.0, min(0.01, pending[i].arrival - now)))
                continue
            before = {m.name: m.ens.members()["steps"] for m in running}
            self.occupancy.append(sum(m.active() for m in running) / (64.0 * len(running)))
            for m in running:           # enqueue on every mesh's stream, then wait
                m.ens.run(self.K, self.graph_steps, wait=False)
            for m in running:
                m.ens.sync()
            now = time.perf_counter() - t0
            for m in running:
                self._count_steps(m, before[m.name])
                self._retire(m, now)
            if now > max_wall_s:
                raise RuntimeError("serving exceeded its wall-clock limit")
        wall = time.perf_counter() - t0
        ttr = np.array([r.finished - r.arrival for r in self.done])
        return {"requests": len(self.done), "wall_s": wall,
                "p50_time_to_result_s": float(np.percentile(ttr, 50)) if len(ttr) else None,
                "p99_time_to_result_s": float(np.percentile(ttr, 99)) if len(ttr) else None,
                "member_steps": self.member_steps, "member_steps_per_s": self.member_steps / wall,
                "mean_occupancy": float(np.mean(self.occupancy)) if self.occupancy else 0.0,
                "failed": int(sum(r.status != 0 for r in self.done))}

    def close(self):
        for m in self.meshes.values():
            m.unload()


def poisson_requests(meshes_xg, rate_per_s, count, seed=0, years=(50.0, 500.0), member_params=None):
    """`count` requests with Poisson arrivals at `rate_per_s`, uniformly over the
    meshes, each with seeded friction parameters and a duration uniform in
    `years`."""
    from . import synth
    rng = np.random.default_rng(seed)
    names = sorted(meshes_xg)
    t = 0.0
    out = []
    for k in range(count):
        t += rng.exponential(1.0 / rate_per_s)
        name = names[rng.integers(len(names))]
        cols = (member_params or synth.member_params)(meshes_xg[name], int(rng.integers(1 << 30)))
        out.append(Request(rid=k, mesh=name, cols=cols, duration_s=float(rng.uniform(*years)) * YEAR, arrival=t))
    return out
