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
* every resident mesh runs bursts of global steps on its own stream; the
  loop polls the streams, and a mesh whose burst has finished retires its
  finished requests, admits queued ones and starts its next burst at once,
  whatever the other meshes are doing, so the meshes' work overlaps on the
  device and a small mesh never waits for a large one;
* a mesh is made resident by a loader thread (Green's-function assembly,
  structure scan, ensemble set-up) while the other meshes keep stepping: the
  library's calls on that path wait for the null stream only, never for the
  whole device.

A request's result is independent of what else runs: slots are independent
columns of the GEMM and independent RK members, so a simulation gives the same
state alone and in the mix (tests/test_serving_gpu.py).
"""
import threading
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
        self.loader = None        # thread making the mesh resident
        self.load_error = None
        self.busy = False         # a burst is in flight on the ensemble's stream
        self.burst_t0 = 0.0
        self.burst_active = 0
        self.steps_before = None
        self.slot_seconds = 0.0   # sum over bursts of active slots x burst wall time

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

    def load(self, tune=True):
        ev, xg = readmsh.get_eleVec(self.nodes, self.tri, False)
        sim = DenseSim(self.n, _lib.MAT_F32)
        sim.assemble(*greenfunc.source_vertices(self.nodes, self.tri), xg, ev, self.mu, self.lam, self.halfspace)
        ens = Ensemble(self.n, 64, matrices_from=sim)
        for m in range(64):
            ens.reset_member(m, active=False)
        ens.structure  # zero-matrix scan and the split GEMM's panel copies, before the first burst
        if autotune.enabled():
            autotune.ensure_ensemble(ens, tune=tune)   # split-K tile width for this mesh and device
        self.slots = [None] * 64
        self.sim, self.ens = sim, ens

    def start_load(self, tune):
        """Make the mesh resident on a thread of its own."""
        def work():
            try:
                self.load(tune)
            except Exception as e:  # reported by the serving loop
                self.load_error = e
        self.loader = threading.Thread(target=work, name=f"load-{self.name}", daemon=True)
        self.loader.start()

    @property
    def loading(self):
        return self.loader is not None and self.loader.is_alive()

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
        self.member_steps = 0
        self.log = []

    # ------------------------------------------------------------- residency
    def _resident_bytes(self):
        return sum(m.hbm_bytes() for m in self.meshes.values() if m.resident or m.loading)

    def _wanted(self, mesh):
        return any(r.mesh == mesh.name for r in self.queue)

    def _start_loads(self):
        """Start loading every non-resident mesh with queued requests that fits
        the budget, evicting idle meshes (no request running or queued) to
        make room."""
        for mesh in self.meshes.values():
            if mesh.resident or mesh.loading or not self._wanted(mesh):
                continue
            need = mesh.hbm_bytes()
            if need > self.budget:
                raise ValueError(f"mesh {mesh.name} needs {need / 1e9:.1f} GB, over the HBM budget")
            while self._resident_bytes() + need > self.budget:
                idle = [m for m in self.meshes.values() if m.resident and not m.busy and m.active() == 0
                        and not self._wanted(m)]
                if not idle:
                    break
                idle[0].unload()
            if self._resident_bytes() + need > self.budget:
                continue
            # tune the tile width only when nothing else shares the device
            quiet = not any(m.busy or m.loading for m in self.meshes.values())
            mesh.start_load(tune=quiet)

    def _finish_loads(self):
        for mesh in self.meshes.values():
            if mesh.loader is not None and not mesh.loader.is_alive():
                mesh.loader = None
                if mesh.load_error is not None:
                    raise RuntimeError(f"loading mesh {mesh.name} failed") from mesh.load_error
                mesh.ens.set_params(self.V0, self.Vpl, self.eta)

    # ------------------------------------------------------------- serving
    def submit(self, req, now):
        req.submitted = now
        self.queue.append(req)

    def _admit(self, mesh, now):
        """Queued requests of this (idle, resident) mesh into its free slots."""
        still = []
        for req in self.queue:
            m = mesh.free_slot() if req.mesh == mesh.name else None
            if m is None:
                still.append(req)
                continue
            load_member(mesh.ens, m, req.cols, self.V0, self.Vpl, t_end=req.duration_s)
            mesh.slots[m] = req
            req.admitted = now
        self.queue = still

    def _retire(self, mesh, mem, now):
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

    def _burst_done(self, mesh, now):
        mesh.busy = False
        mem = mesh.ens.members()
        self.member_steps += int(np.sum(np.maximum(mem["steps"] - mesh.steps_before, 0)))
        mesh.slot_seconds += mesh.burst_active * (now - mesh.burst_t0)
        self._retire(mesh, mem, now)

    def _start_burst(self, mesh, now):
        mesh.steps_before = mesh.ens.members()["steps"]
        mesh.burst_active = mesh.active()
        mesh.ens.run(self.K, self.graph_steps, wait=False)
        mesh.busy = True
        mesh.burst_t0 = now

    def run(self, requests, max_wall_s=600.0):
        """Serve `requests` (each with an `arrival` offset in seconds) until all
        have finished.  Returns the metrics dict: p50/p99 time-to-result
        (arrival to result), accepted member-steps/s over the run, and the
        occupancy (time-averaged share of a mesh's 64 slots holding a
        request) per mesh over the run and over the arrival window."""
        pending = sorted(requests, key=lambda r: r.arrival)
        t0 = time.perf_counter()
        i = 0
        first_arrival = pending[0].arrival if pending else 0.0
        last_arrival = pending[-1].arrival if pending else 0.0
        window = {name: 0.0 for name in self.meshes}   # slot-seconds inside the arrival window
        while True:
            now = time.perf_counter() - t0
            while i < len(pending) and pending[i].arrival <= now:
                self.submit(pending[i], pending[i].arrival)
                i += 1
            self._start_loads()
            self._finish_loads()
            progressed = False
            for mesh in self.meshes.values():
                if not mesh.resident or mesh.loading:
                    continue
                if mesh.busy:
                    if not mesh.ens.done():
                        continue
                    t = time.perf_counter() - t0
                    lo, hi = max(mesh.burst_t0, first_arrival), min(t, last_arrival)
                    if hi > lo:
                        window[mesh.name] += mesh.burst_active * (hi - lo)
                    self._burst_done(mesh, t)
                    progressed = True
                self._admit(mesh, now)
                if mesh.active() > 0:
                    self._start_burst(mesh, time.perf_counter() - t0)
                    progressed = True
            if i >= len(pending) and not self.queue and not any(m.busy or m.loading
                                                                 for m in self.meshes.values()):
                break
            if now > max_wall_s:
                raise RuntimeError("serving exceeded its wall-clock limit")
            if not progressed:
                time.sleep(2e-4)
        wall = time.perf_counter() - t0
        ttr = np.array([r.finished - r.arrival for r in self.done])
        span = max(last_arrival - first_arrival, 1e-9)
        occ_run = {k: m.slot_seconds / (64.0 * wall) for k, m in self.meshes.items()}
        occ_win = {k: window[k] / (64.0 * span) for k in self.meshes}
        used = [k for k in self.meshes if self.meshes[k].slot_seconds > 0]
        return {"requests": len(self.done), "wall_s": wall,
                "p50_time_to_result_s": float(np.percentile(ttr, 50)) if len(ttr) else None,
                "p99_time_to_result_s": float(np.percentile(ttr, 99)) if len(ttr) else None,
                "member_steps": self.member_steps, "member_steps_per_s": self.member_steps / wall,
                "mean_occupancy": float(np.mean([occ_run[k] for k in used])) if used else 0.0,
                "occupancy_per_mesh": occ_run, "occupancy_arrival_window": occ_win,
                "arrival_window_s": span,
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
