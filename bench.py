"""Benchmark of the PyQuake3D hot path on one MI355X.

Mesh (all legs): planar vertical strike-slip fault, 40 km x 40 km, 100 x 100
squares split into 20,000 triangular dislocations, half space.

Legs
----
* Green's functions: FP64 triangular-dislocation assembly on the GPU (six dense
  20k x 20k traction matrices), stored in FP32; pairs/s and FP64 FLOP/s.
* value -- batched ensemble: 64 members with seeded friction parameters
  (a-b ~ U[-0.006,-0.002] on the patch, Dc = 0.14 exp(0.3 N), sigma_n ~
  U[20,30] MPa) share the FP32 matrices; the traction rates of all members
  come from one tall-skinny GEMM on MFMA per global step, and every member
  takes one accepted adaptive RK45 step with its own step size.  The GEMM
  splits every FP32 operand exactly into three BF16 parts and accumulates the
  six leading products in FP32 on the BF16 MFMA units (the same ensemble with
  the FP32-MFMA GEMM is reported beside it).
  value = accepted member-steps per second (64 x global steps / time).
* ensemble.rupture: the same mesh and matrices with 64 members that rupture
  (a-b ~ U[-0.026,-0.022], rake 0.5: nucleation size near 7 km, while the
  value's members creep), timed over 200 global steps once eight members
  are coseismic (slip rate above 1e-2 m/s).
* serving: Poisson arrivals over 2k/10k/30k planar meshes; p50/p99
  time-to-result, member-steps/s and slot occupancy.
* config1: the 1,600-element case, 2,000 adaptive steps on the plain FP64
  path, beside the NumPy restatement on the host.
* plain: one simulation (aging law, Dc lognormal sigma 0.2, V = Vpl = 1e-9
  m/s, steady state) with FP32 and with FP64 matrix storage; member-steps/s.
* roofline: the ensemble GEMM (dominant kernel of `value`): with the BF16
  split it executes 6 x 2 x 64 = 768 flops per 4-byte matrix entry, 192
  flop/B, below the BF16 MFMA ridge (2.5 PFLOP/s / 8 TB/s = 312 flop/B), so it
  is HBM-bound: algorithmic bytes (matrices, slip rates, outputs) per launch over
  the mean launch time from HIP events, against 8 TB/s; the FP32-MFMA GEMM's
  roofline (2 nmat N^2 64 flops against 157.3 TFLOP/s) and the plain FP32
  matvec against 8 TB/s are reported beside it.
* cpu_baseline: the NumPy FP64 restatement of the reference's dense step
  (oracle/qdsim_oracle.py: six np.dot per step, BLAS on all host cores) on the
  same 20k FP64 matrices, over a bounded number of steps.
* hmatrix (separate mesh): 60,000-element rough vertical fault, H-matrix with
  leaf 64, eta 2, ACA 1e-4, BF16 low-rank factors and FP32 dense leaves: build
  time, product time against 8 TB/s, error against the exact FP64 dense
  operator (assembled in HBM, 173 GB) overall and per block class, stepping
  rate with the H-matrix and with the dense FP64 matrices.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-hmat | --hmat-only]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "member-steps/s at 20k tri (batched vs plain) + p99 solve latency; GF pairs/s"
MU = LAM = 32038120320.0
CS = 3464.0
V0, VPL = 1e-6, 1e-9
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_MFMA_PEAK_TFLOPS = 157.3
MEMBERS = 64


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def build_case(nx=100, nz=100, length=40.0, depth=40.0, dc_sigma=0.2, seed=1):
    from pyquake3d_amd import greenfunc, readmsh, synth
    nodes_km, tri = synth.planar_fault_mesh(nx, nz, length, depth)
    nodes = nodes_km * 1e3
    ev, xg = readmsh.get_eleVec(nodes, tri, False)
    cols = synth.friction_params(xg, dc_sigma=dc_sigma, seed=seed)
    return nodes, tri, ev, xg, cols, greenfunc.source_vertices(nodes, tri)


def init_state(sim, cols):
    """Steady-state initial condition from the init-file columns
    (QDsim_gpu.py:714-770)."""
    rake, a, b, dc, f0, _, Tn, V, sl, nl = np.asarray(cols, np.float64).T
    Tt = Tn * a * np.arcsinh(V / (2.0 * V0) * np.exp((f0 + b * np.log(V0 / VPL)) / a))
    state = np.log(np.sinh(Tt / Tn / a) * 2.0 * V0 / V) * a
    x, y = np.cos(rake), np.sin(rake)
    vals = {"Tno": Tn, "Tt1o": Tt * x, "Tt2o": Tt * y, "state": state, "slipv1": V * x + 1e-16,
            "slipv2": V * y + 1e-16, "a": a, "b": b, "dc": dc, "f0": f0, "shear_loading": sl,
            "normal_loading": nl}
    vals["slipv"] = np.sqrt(vals["slipv1"] ** 2 + vals["slipv2"] ** 2)
    for k, v in vals.items():
        sim.set(k, v)
    sim.set("rake0", rake)
    sim.set_params(V0, VPL, MU / (2 * CS), False, True)


def time_plain(sim, steps, warmup, graph_steps):
    from pyquake3d_amd import _lib
    sim.step(1e-3)
    if warmup > 1:
        st, _ = sim.run(warmup - 1, graph_steps)
        assert st == 0, f"warm-up failed with status {st}"
    t = _lib.Timer()
    _lib.lib().pq3d_synchronize()
    t.start()
    st, _ = sim.run(steps, graph_steps)
    t.stop()
    ms = t.ms()
    assert st == 0, f"timed steps failed with status {st}"
    return ms


def time_ensemble(ens, steps, warmup, graph_steps):
    """Wall time of `steps` global steps (host clock around an enqueue plus
    stream sync; the ensemble runs on its own stream).  Returns (ms, accepted
    member-steps in the timed region)."""
    if warmup > 0:
        ens.run(warmup, graph_steps)
    before = ens.members()["steps"].copy()
    t0 = time.perf_counter()
    ens.run(steps, graph_steps)
    ms = (time.perf_counter() - t0) * 1e3
    mem = ens.members()
    assert np.all(mem["status"] == 0), "an ensemble member failed"
    return ms, int(np.sum(mem["steps"] - before))


def rupture_ensemble_leg(sim, xg, timed=200, graph_steps=8, vdyn=1e-2, min_members=8, chunk=200,
                         max_steps=12000):
    """The batched rate while members rupture.  The metric's members creep
    (their nucleation size exceeds the patch), so this leg runs 64 members
    with a - b in [-0.026, -0.022] (nucleation size near 7 km), rake 0.5, on the same
    mesh and matrices, steps them until at least `min_members` members have
    slip rates above `vdyn` (coseismic: steps shrink, members desynchronise,
    rejections make the second attempt round live), then times `timed`
    global steps there."""
    from pyquake3d_amd import synth
    from pyquake3d_amd.ensemble import Ensemble, init_members
    mcols = np.stack([synth.member_params(xg, m, base_seed=5000, amb_range=(-0.026, -0.022), rake=0.5)
                      for m in range(MEMBERS)])
    ens = Ensemble(len(xg), MEMBERS, matrices_from=sim)
    init_members(ens, mcols, V0, VPL, MU, CS)
    try:
        return time_ensemble_rupture(ens, vdyn, min_members, chunk, max_steps, timed, graph_steps)
    finally:
        ens.close()


def time_ensemble_rupture(ens, vdyn=1e-2, min_members=8, chunk=200, max_steps=12000, timed=200, graph_steps=8):
    """Step the ensemble on until at least `min_members` members have slip
    rates above `vdyn`, then time `timed` global steps there."""
    searched = 0
    while searched < max_steps:
        vmax = ens.get("slipv").max(axis=1)
        if int(np.sum(vmax > vdyn)) >= min_members:
            break
        ens.run(chunk, graph_steps)
        searched += chunk
    vmax0 = ens.get("slipv").max(axis=1)
    before = ens.members()
    t0 = time.perf_counter()
    ens.run(timed, graph_steps)
    ms = (time.perf_counter() - t0) * 1e3
    after = ens.members()
    vmax1 = ens.get("slipv").max(axis=1)
    acc = int(np.sum(after["steps"] - before["steps"]))
    return {"value": acc / (ms * 1e-3), "unit": "member-steps/s", "global_steps": timed,
            "ms_per_global_step": ms / timed, "steps_before_window": searched,
            "members_above_vdyn_at_start": int(np.sum(vmax0 > vdyn)),
            "members_above_vdyn_at_end": int(np.sum(vmax1 > vdyn)), "vdyn_m_per_s": vdyn,
            "peak_slip_rate_m_per_s": float(max(vmax0.max(), vmax1.max())),
            "min_step_s": float(min(before["h"].min(), after["h"].min())),
            "failed_members": int(np.sum(after["status"] != 0)),
            "members": "64 members, a - b uniform in [-0.026, -0.022] on the patch (seed 5000 + member), Dc "
                       "lognormal (median 0.14 m, sigma 0.3), normal stress uniform 20-30 MPa, rake 0.5; same "
                       "mesh and FP32 matrices, BF16-split GEMM",
            "note": "a member-step is one accepted step of one member; the window starts once members rupture"}


def serve(n_requests, rate, seed=7):
    """Continuous-batching serving over 2k, 10k and 30k planar meshes."""
    from pyquake3d_amd import readmsh, synth
    from pyquake3d_amd.serving import Server, poisson_requests
    meshes, xg = {}, {}
    for name, (nx, nz, L, D) in {"2k": (50, 20, 50.0, 20.0), "10k": (100, 50, 50.0, 25.0),
                                 "30k": (150, 100, 60.0, 40.0)}.items():
        nodes_km, tri = synth.planar_fault_mesh(nx, nz, L, D)
        nodes = nodes_km * 1e3
        meshes[name] = (nodes, tri)
        xg[name] = readmsh.get_eleVec(nodes, tri, False)[1]
    reqs = poisson_requests(xg, rate, n_requests, seed=seed)
    srv = Server(meshes, hbm_budget_bytes=200e9, steps_per_iteration=64, graph_steps=16)
    try:
        m = srv.run(reqs, max_wall_s=900.0)
    finally:
        srv.close()
    m.update({"arrival_rate_per_s": rate, "meshes": {k: len(v[1]) for k, v in meshes.items()},
              "durations_years": [50.0, 500.0], "hbm_budget_gb": 200.0})
    return m


def assembly_flops_per_pair(every_side=False):
    """FP64 operations per source-receiver pair (both slip directions) from the
    counting build of the kernel's own source (pyquake3d_amd/csrc/flopcount.cpp):
    as the kernel runs them (a side shared by consecutive sources evaluated
    once), or with every side evaluated for every pair (every_side)."""
    p = os.path.join(ROOT, "pyquake3d_amd", "csrc", "flops_per_pair.txt")
    try:
        with open(p) as f:
            return float(f.read().splitlines()[1 if every_side else 0].split()[0])
    except (OSError, ValueError, IndexError):
        return None


def host_threads():
    """Host threads the CPU baselines use: the cores this process may run on,
    capped by OMP_NUM_THREADS (the BLAS thread count) when it is set."""
    n = len(os.sched_getaffinity(0))
    try:
        return max(1, min(n, int(os.environ.get("OMP_NUM_THREADS", n))))
    except ValueError:
        return n


def cpu_baseline_step(cols, mats64, budget_s=15.0, max_steps=40):
    """NumPy FP64 restatement of the dense step on the host (the oracle)."""
    from oracle import qdsim_oracle as Q
    threads = host_threads()
    orc = Q.from_init_columns(mats64, cols, V0, VPL, MU, CS)
    orc.step(1e-3)  # untimed first step
    dt = 1.5e-3
    t0 = time.perf_counter()
    n = 0
    while n < max_steps and time.perf_counter() - t0 < budget_s:
        _, dt = orc.step(dt)
        n += 1
    el = time.perf_counter() - t0
    return {"value": n / el, "unit": "member-steps/s", "cores": threads, "kind": "port",
            "sample": f"{n} steps of the NumPy FP64 restatement of simu_forward (six np.dot per step) on the "
                      f"same 20,000-element FP64 matrices, BLAS threads = {threads}, CPU {_cpu_model()}"}


def config1_leg(steps=2000, cpu_budget_s=8.0, graph_steps=8):
    """The reference's CPU-path case at its own size: 80 x 40 km planar fault,
    40 x 20 squares split into 1,600 TDs, half space, 2,000 adaptive steps.
    GPU: the plain FP64 dense path over all 2,000 steps (wall clock of the
    run, host sync at the end).  CPU: the NumPy FP64 restatement of
    simu_forward on the same matrices for a bounded sample of steps."""
    from oracle import qdsim_oracle as Q
    from pyquake3d_amd import _lib, greenfunc, readmsh, synth
    from pyquake3d_amd.QDsim_gpu import DenseSim
    nodes_km, tri = synth.planar_fault_mesh(40, 20, 80.0, 40.0)
    nodes = nodes_km * 1e3
    ev, xg = readmsh.get_eleVec(nodes, tri, False)
    cols = synth.friction_params(xg)
    n = len(tri)
    s = DenseSim(n, _lib.MAT_F64)
    s.assemble(*greenfunc.source_vertices(nodes, tri), xg, ev, MU, LAM, True)
    init_state(s, cols)
    mats = [s.download_matrix(k) for k in range(6)]
    s.step(1e-3)
    _lib.lib().pq3d_synchronize()
    t0 = time.perf_counter()
    st, rec = s.run(steps - 1, graph_steps, records=True)
    wall = time.perf_counter() - t0
    assert st == 0, f"config-1 run failed with status {st}"
    sim_years = float(rec[-1, 2]) / (365.25 * 86400)
    max_v = float(rec[:, 3].max())
    s.close()
    out = {"n_elements": n, "steps": steps, "gpu_wall_s": wall, "gpu_steps_per_s": (steps - 1) / wall,
           "simulated_years": sim_years, "max_slip_rate": max_v, "cpu_baseline": None}
    if cpu_budget_s > 0:
        orc = Q.from_init_columns(mats, cols, V0, VPL, MU, CS, False)
        orc.step(1e-3)
        dt = 1.5e-3
        t0 = time.perf_counter()
        k = 0
        while k < steps and time.perf_counter() - t0 < cpu_budget_s:
            _, dt = orc.step(dt)
            k += 1
        threads = host_threads()
        out["cpu_baseline"] = {"value": k / (time.perf_counter() - t0), "unit": "steps/s", "cores": threads,
                               "kind": "port",
                               "sample": f"{k} steps of the NumPy FP64 restatement of simu_forward on the same "
                                         f"matrices, BLAS threads = {threads}, CPU {_cpu_model()}"}
    cpu = out["cpu_baseline"]
    log(f"[bench] config 1 (1,600 TDs): {steps} steps in {wall:.2f} s on the GPU ({out['gpu_steps_per_s']:.0f} "
        f"steps/s, {sim_years:.1f} years, max V {max_v:.3g} m/s)"
        + (f"; CPU port {cpu['value']:.1f} steps/s on {cpu['cores']} threads" if cpu else ""))
    return out


def _ref_column(args):
    from oracle import refbind
    P1, P2, P3, xg = args
    for Ss, Ds in ((1.0, 0.0), (0.0, 1.0)):
        refbind.td_stress_fs_c(xg[:, 0], xg[:, 1], xg[:, 2], P1, P2, P3, Ss, Ds, 0.0, MU, LAM, True)
    return 0


def _ref_warm(_):
    from oracle import refbind
    refbind.lib()
    return 0


def cpu_baseline_assembly(P, xg):
    """The reference's own TDstressFS_C (built from source into oracle/_ref)
    over sampled source columns, one process per host thread.  The workers
    are spawned (this process has opened the GPU, which a forked child would
    inherit) and loaded before the timed map."""
    from oracle import refbind
    if not refbind.available():
        return None
    import multiprocessing as mp
    threads = host_threads()
    ncol = threads
    with mp.get_context("spawn").Pool(threads) as pool:
        pool.map(_ref_warm, range(threads))
        t0 = time.perf_counter()
        pool.map(_ref_column, [(P[0][j], P[1][j], P[2][j], xg) for j in range(ncol)])
        el = time.perf_counter() - t0
    return {"value": ncol * len(xg) / el, "unit": "pairs/s", "cores": threads, "kind": "reference",
            "sample": f"{ncol} source columns x {len(xg)} receivers through the reference TDstressFS_C "
                      f"(strike and dip calls), {threads} processes"}


def _bf16(a):
    """Round to BF16 (nearest even) and back, as the device stores factors."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    b = f.view(np.uint32)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16).astype(np.uint32) << 16
    return r.view(np.float32).astype(np.float64)


def hmatrix_class_errors(tree, factor_round, rng, per_class=40, sub=128):
    """Relative Frobenius error of the stored blocks against exact FP64
    entries on random sub-blocks (<= sub x sub), per block class."""
    classes = {"dense_leaf": [], "lowrank_m<=256": [], "lowrank_256<m<=2048": [], "lowrank_m>2048": []}
    for b in tree.blocks_to_process:
        m = len(b.row_cluster)
        key = "dense_leaf" if not b.judaca else (
            "lowrank_m<=256" if m <= 256 else ("lowrank_256<m<=2048" if m <= 2048 else "lowrank_m>2048"))
        classes[key].append(b)
    out = {}
    for key, blocks in classes.items():
        if not blocks:
            continue
        pick = rng.choice(len(blocks), size=min(per_class, len(blocks)), replace=False)
        reqs, sel = [], []
        for i in pick:
            b = blocks[i]
            ri = np.sort(rng.choice(len(b.row_cluster), size=min(sub, len(b.row_cluster)), replace=False))
            ci = np.sort(rng.choice(len(b.col_cluster), size=min(sub, len(b.col_cluster)), replace=False))
            reqs.append((b.row_cluster[ri], b.col_cluster[ci]))
            sel.append((b, ri, ci))
        exact = tree.src.segments(reqs)
        shear, normal = [], []
        for (b, ri, ci), A in zip(sel, exact):
            strongest = max(np.linalg.norm(A[c]) for c in range(6))
            for c, name in enumerate(("A1s", "A2s", "Bs", "A1d", "A2d", "Bd")):
                if b.judaca:
                    U, V = b.factors(name)
                    approx = factor_round(U[ri]) @ factor_round(V[:, ci]) if U.shape[1] else np.zeros_like(A[c])
                else:
                    approx = np.asarray(getattr(b, "Mf_" + name))[np.ix_(ri, ci)].astype(np.float32)
                e = np.linalg.norm(A[c] - approx) / max(np.linalg.norm(A[c]), 1e-12 * strongest, 1e-300)
                (normal if c in (2, 5) else shear).append(e)
        out[key] = {"blocks": len(blocks), "sampled": len(sel), "shear_max": float(np.max(shear)),
                    "shear_median": float(np.median(shear)), "normal_max": float(np.max(normal)),
                    "normal_median": float(np.median(normal))}
    return out


def hmatrix_leg(steps=200, warmup=10, graph_steps=8, cpu_budget_s=20.0):
    """Config 3: 60,000-element rough fault, H-matrix (leaf 64, eta 2, ACA 1e-4,
    BF16 low-rank factors, FP32 dense leaves).  Build cost, product time
    against the HBM roofline, error against the exact FP64 dense operator,
    stepping throughput with the H-matrix and with the dense FP64 matrices,
    and the reference block-loop product on the host."""
    from oracle import hmat_oracle as HO
    from pyquake3d_amd import Hmatrix as H, _lib, synth
    from pyquake3d_amd.QDsim_gpu import DenseSim
    nodes_km, tri = synth.rough_fault_mesh(300, 100, 120.0, 30.0, amp_km=1.0, wavelength_km=40.0, hurst=0.8,
                                           rough_ratio=1e-3, seed=7)
    nodes = nodes_km * 1e3
    N = len(tri)
    log(f"[hmat] N={N}: building tree + ACA")
    t0 = time.perf_counter()
    tree = H.build(nodes, tri, MU, LAM, True, mini_leaf=64, min_size=16, eta=2.0, eps=1e-4)
    build_s = time.perf_counter() - t0
    nlow = sum(1 for b in tree.blocks_to_process if b.judaca)
    ranks = [tree.blocks_to_process[i].factors(nm)[0].shape[1] for i in range(len(tree.blocks_to_process))
             if tree.blocks_to_process[i].judaca for nm in ("A1s", "A2s", "Bs", "A1d", "A2d", "Bd")]
    log(f"[hmat] built in {build_s:.1f} s: {len(tree.blocks_to_process)} leaves, {nlow} low-rank; "
        f"{tree.build_stats}")
    res = {"n_elements": N, "leaves": len(tree.blocks_to_process), "lowrank_leaves": nlow,
           "rank_mean": float(np.mean(ranks)), "rank_max": int(np.max(ranks)), "build_s": build_s,
           "build_stats": tree.build_stats,
           "mesh": "rough vertical fault 120 x 30 km, near-equilateral lattice 300 x 100 x 2; y = 1 km sin(2 pi x / "
                   "40 km) + self-affine roughness (H = 0.8, RMS 1e-3 x 120 km, seed 7)",
           "tree": {"leaf": 64, "cluster_min_size": 16, "eta": 2.0, "aca_eps": 1e-4, "recompressed": True}}
    x1 = np.random.default_rng(2).standard_normal(N)
    x2 = np.random.default_rng(3).standard_normal(N)
    prods = {}
    for label, ft in (("bf16", _lib.MAT_BF16), ("fp32", _lib.MAT_F32)):
        dev = tree.device(factor_type=ft)
        info = dev.info()
        y = dev.mvm(x1, x2)
        ms = dev.time_mvm(20)
        ms_fix = dev.time_mvm(20, fix_tn=True)
        gbs = info["bytes_per_mvm"] / (ms * 1e-3) / 1e9
        prods[label] = {"y": y, "ms": ms, "ms_fix_tn": ms_fix, "bytes": info["bytes_per_mvm"], "gbs": gbs,
                        "info": info}
        log(f"[hmat] product {label}: {ms:.3f} ms, {info['bytes_per_mvm'] / 1e9:.2f} GB -> {gbs:.0f} GB/s")
    # stepping with the H-matrix (BF16 factors)
    cols = synth.friction_params(tree.xg, dc_sigma=0.2, seed=5)
    sh = DenseSim(N, _lib.MAT_NONE)
    sh.attach_hmatrix(tree.device(factor_type=_lib.MAT_BF16))
    init_state(sh, cols)
    h_ms = time_plain(sh, steps, warmup, graph_steps)
    h_rate = steps / (h_ms * 1e-3)
    sh.close()
    log(f"[hmat] stepping with the H-matrix: {h_rate:.1f} steps/s")
    # exact FP64 operator (173 GB of HBM): product truth and the dense step rate
    s64 = DenseSim(N, _lib.MAT_F64)
    P = (tree.src.P1, tree.src.P2, tree.src.P3)
    t = _lib.Timer()
    t.start()
    s64.assemble(*P, tree.xg, tree.eleVec, MU, LAM, True)
    t.stop()
    asm64_ms = t.ms()
    s64.set("rake0", np.zeros(N))
    s64.set_params(V0, 0.0, MU / (2 * CS), False, True)
    s64.set("slipv1", x1)
    s64.set("slipv2", x2)
    s64.set("normal_loading", np.zeros(N))
    s64.matvec()
    exact = [s64.get("AdotV1"), s64.get("AdotV2"), s64.get("dsigmadt")]
    mv64_ms = s64.time_matvec(5)
    init_state(s64, cols)
    k64 = max(20, steps // 10)
    d_ms = time_plain(s64, k64, 2, graph_steps)
    d_rate = k64 / (d_ms * 1e-3)
    s64.close()
    log(f"[hmat] dense FP64: assembly {asm64_ms:.0f} ms, matvec {mv64_ms:.2f} ms, {d_rate:.1f} steps/s")
    for label, p in prods.items():
        p["rel_err_vs_fp64_dense"] = [float(np.linalg.norm(p["y"][q] - exact[q]) / np.linalg.norm(exact[q]))
                                      for q in range(3)]
        del p["y"]
    rng = np.random.default_rng(11)
    res["block_class_error_bf16"] = hmatrix_class_errors(tree, _bf16, rng)
    res["block_class_error_fp32"] = hmatrix_class_errors(tree, lambda a: a.astype(np.float32).astype(np.float64),
                                                         np.random.default_rng(11))
    # the reference's block-loop product on the host (Hmatrix.py:1773-1866, six calls per step)
    leaves = []
    for b in tree.blocks_to_process:
        if b.judaca:
            leaves.append((b.row_cluster, b.col_cluster, None, {nm: b.factors(nm) for nm in HO.NAMES}))
        else:
            leaves.append((b.row_cluster, b.col_cluster, {nm: getattr(b, "Mf_" + nm) for nm in HO.NAMES}, None))
    tc = time.perf_counter()
    nprod = 0
    while cpu_budget_s > 0 and nprod < 3 and (nprod == 0 or time.perf_counter() - tc < cpu_budget_s):
        for q in range(3):
            HO.blocks_mvm(x1, leaves, HO.NAMES[q])
            HO.blocks_mvm(x2, leaves, HO.NAMES[q + 3])
        nprod += 1
    cpu_s = (time.perf_counter() - tc) / nprod if nprod else None
    tree.close()
    bf = prods["bf16"]
    res.update({
        "product": prods,
        "roofline": {"bound": "hbm", "achieved": bf["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": bf["gbs"] / HBM_PEAK_GBS, "traffic": _pmc_traffic("hmat_bf16_60k_traffic.json"),
                     "kernel": "H-matrix product, BF16 factors (gather + phase A + phase B + reduce)",
                     "alg_bytes_per_launch": bf["bytes"], "avg_launch_ms": bf["ms"]},
        "stepping_hmatrix_bf16_steps_per_s": h_rate,
        "stepping_dense_fp64_steps_per_s": d_rate,
        "dense_fp64_matvec_ms": mv64_ms,
        "dense_fp64_assembly_ms": asm64_ms,
        "dense_fp64_assembly_pairs_per_s": float(N) * N / (asm64_ms * 1e-3),
        "product_speedup_vs_dense_fp64": mv64_ms / bf["ms"],
        "cpu_baseline": None if cpu_s is None else {
            "value": cpu_s, "unit": "s per step (6 block products)", "cores": 1, "kind": "port",
            "sample": f"{nprod} step(s) of the reference's block-loop product (NumPy FP64, "
                      "oracle/hmat_oracle.blocks_mvm) over the same leaves"},
    })
    return res


def traction_error(s64, gemm_out, member_u, members=(0, 21, 42, 63)):
    """Traction rates of the batched path (FP32 Green's functions, MFMA GEMM,
    split-K partials summed in FP64) against the FP64 dense product of the
    same FP64-assembled matrices, for the members' slip rates after the timed
    steps.  Per member and component: ||t - t64||_2 / ||t64||_2 and
    max|t - t64| / max|t64| (a component that is identically zero, as the
    normal traction rate of a planar fault, is skipped)."""
    nl = s64.get("normal_loading")
    rows = []
    for m in members:
        for k in ("slipv1", "slipv2", "rake0"):
            s64.set(k, member_u[k][m])
        s64.matvec()
        ref = [s64.get("AdotV1"), s64.get("AdotV2"), s64.get("dsigmadt") - nl]
        for q, name in enumerate(("strike", "dip", "normal")):
            r, g = ref[q], gemm_out[q, m]
            nr = float(np.linalg.norm(r))
            if nr == 0.0:
                continue
            rows.append({"member": m, "component": name, "rel_l2": float(np.linalg.norm(g - r)) / nr,
                         "rel_max": float(np.abs(g - r).max() / np.abs(r).max())})
    return {"members": list(members), "per_result": rows,
            "max_rel_l2": max(r["rel_l2"] for r in rows), "max_rel_max": max(r["rel_max"] for r in rows),
            "target_rel": 1e-5,
            "reference": "FP64 dense product (FP64 matrices, FP64 sums) of the same slip rates"}


MATRIX_NAMES = ("A1s", "A2s", "Bs", "A1d", "A2d", "Bd")


def _tuned(obj, res, builtin_kc):
    """Autotuner result for the JSON line: the chosen split-K width and its
    time beside the built-in width's (both from the same sweep)."""
    table = dict(res["table"])
    kc_b = min(table, key=lambda k: abs(k - builtin_kc))
    log(f"[bench] autotune {res['key'].split('|')[2]}: kc {res['kc']} ({table[res['kc']]:.3f} ms per product in "
        f"the sweep), built-in kc {builtin_kc}: {table[kc_b]:.3f} ms, {len(table)} candidates")
    return {"kc": res["kc"], "ms_per_product": table[res["kc"]], "final_median_ms": res["ms"],
            "builtin_kc": builtin_kc, "builtin_ms_per_product": table[kc_b], "candidates": len(table),
            "table": [[k, round(v, 5)] for k, v in res["table"]]}


def measured_peaks():
    """What this device sustains, from the library's own probes, beside the
    published peaks: a read-only stream of 16-byte vectors over an 8 GiB
    buffer (the fastest of five timed passes), and independent FP64 FMA
    chains at full occupancy (five timed launches)."""
    import ctypes
    from pyquake3d_amd import _lib
    L = _lib.lib()
    fp, dp = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double)
    _lib._sig(L, "pq3d_probe_stream", [ctypes.c_longlong, ctypes.c_int, fp, dp])
    _lib._sig(L, "pq3d_probe_fma64", [ctypes.c_int, ctypes.c_int, fp, dp])
    ms, gbs = ctypes.c_float(), ctypes.c_double()
    nbytes = 8 << 30
    _lib.check(L.pq3d_probe_stream(nbytes, 5, ctypes.byref(ms), ctypes.byref(gbs)), "probe_stream")
    ms2, tf = ctypes.c_float(), ctypes.c_double()
    _lib.check(L.pq3d_probe_fma64(4096, 5, ctypes.byref(ms2), ctypes.byref(tf)), "probe_fma64")
    out = {"hbm_read_gbs": gbs.value, "hbm_read_frac_of_published": gbs.value / HBM_PEAK_GBS,
           "stream_bytes": nbytes, "stream_ms": ms.value,
           "fp64_fma_tflops": tf.value, "fp64_frac_of_published": tf.value / FP64_PEAK_TFLOPS, "fma_ms": ms2.value,
           "note": "the library's probes (pq3d_probe_stream: a contiguous share per workgroup, 16-byte loads, "
                   "eight in flight per lane, 32 workgroups per CU, fastest pass; pq3d_probe_fma64: eight independent FP64 FMA chains per lane, 64 "
                   "workgroups per CU); the rooflines' frac stays against the published peaks, frac_of_measured "
                   "is against these"}
    log(f"[bench] measured: read stream {gbs.value:.0f} GB/s ({gbs.value / HBM_PEAK_GBS:.2f} of 8 TB/s), "
        f"FP64 FMA {tf.value:.1f} TFLOP/s ({tf.value / FP64_PEAK_TFLOPS:.2f} of 78.6)")
    return out


def _pmc_json(name):
    """A committed rocprofv3 --pmc summary (profiles/<name>), or None."""
    p = os.path.join(ROOT, "profiles", name)
    try:
        with open(p) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def _pmc_traffic(name):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary, or None."""
    d = _pmc_json(name)
    try:
        return float(d["hbm_bytes_per_launch"]) if d else None
    except (KeyError, TypeError, ValueError):
        return None


def _pmc_mfma(name):
    """MFMA busy fraction and counted MFMA flops of the ensemble GEMM from the
    committed counter summary (SQ_VALU_MFMA_BUSY_CYCLES, SQ_INSTS_VALU_MFMA_MOPS_F32)."""
    d = _pmc_json(name)
    if not d:
        return None
    return {"mfma_util": d.get("mfma_util"), "mfma_flops_per_launch": d.get("mfma_flops_per_launch"),
            "mfma_flops_over_algorithmic": d.get("mfma_flops_over_algorithmic"), "source": "profiles/" + name}


def _pmc_fp64(name, pairs, ms):
    """Hardware-counted FP64 flops of the assembly kernel ((2 FMA + ADD + MUL) x 64
    per wave instruction, committed counter summary) at this run's timing."""
    d = _pmc_json(name)
    try:
        k = d["per_output_type"]["assemble_f32out"]
        fpp = float(k["fp64_flops_per_pair"])
    except (KeyError, TypeError, ValueError):
        return None
    tf = pairs * fpp / (ms * 1e-3) / 1e12
    return {"fp64_flops_per_pair": fpp, "fp64_tflops": tf, "frac_fp64_peak": tf / FP64_PEAK_TFLOPS,
            "valu_issue_util": k.get("valu_issue_util"), "source": "profiles/" + name}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500, help="timed ensemble global steps")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--plain-steps", type=int, default=2000)
    ap.add_argument("--graph-steps", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=400, help="0 skips the serving leg")
    ap.add_argument("--serve-rate", type=float, default=40.0,
                    help="Poisson arrival rate, requests/s over the three meshes (40/s keeps the 30k mesh's "
                         "64 slots more than half full)")
    ap.add_argument("--serve-only", action="store_true", help="run only the serving leg (JSON of that leg)")
    ap.add_argument("--rupture-steps", type=int, default=200,
                    help="global steps timed once members rupture (0 skips that window)")
    ap.add_argument("--no-hmat", action="store_true", help="skip the 60k H-matrix leg")
    ap.add_argument("--no-config1", action="store_true", help="skip the 1,600-element 2,000-step leg")
    ap.add_argument("--hmat-only", action="store_true", help="run only the 60k H-matrix leg (JSON of that leg)")
    ap.add_argument("--hmat-steps", type=int, default=200)
    ap.add_argument("--copy-probe-gib", type=float, default=0.0,
                    help="only copy a device buffer of this size 5 times (calibrates PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: one GPU only; runs on more than one GPU are not supported (members, row panels or "
              "H-matrix row clusters would be sharded across GPUs, which this code does not do)", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe_gib > 0:
        from pyquake3d_amd import _lib
        nbytes = int(args.copy_probe_gib * 2 ** 30) // 256 * 256
        src, dst = _lib.DeviceBuffer(nbytes), _lib.DeviceBuffer(nbytes)
        _lib.check(_lib.lib().pq3d_memset(src.ptr, 1, nbytes), "memset")
        for _ in range(5):
            _lib.check(_lib.lib().pq3d_memcpy_d2d(dst.ptr, src.ptr, nbytes), "copy")
        _lib.lib().pq3d_synchronize()
        print(json.dumps({"copy_bytes": nbytes, "copies": 5}))
        return

    if args.serve_only:
        print(json.dumps(serve(args.serve_requests, args.serve_rate)))
        return

    if args.hmat_only:
        print(json.dumps(hmatrix_leg(steps=args.hmat_steps, warmup=args.warmup, graph_steps=args.graph_steps,
                                     cpu_budget_s=0.0 if args.no_cpu_baseline else 20.0)))
        return

    from pyquake3d_amd import _lib, autotune, synth
    from pyquake3d_amd.QDsim_gpu import DenseSim
    from pyquake3d_amd.ensemble import Ensemble, init_members

    peaks = measured_peaks()
    nodes, tri, ev, xg, cols, P = build_case()
    N = len(tri)
    log(f"[bench] N={N} elements")

    # --- Green's functions: FP64 arithmetic, FP32 storage
    sim = DenseSim(N, _lib.MAT_F32)
    t = _lib.Timer()
    sim.assemble(*P, xg, ev, MU, LAM, True)   # first call loads the code object
    t.start()
    sim.assemble(*P, xg, ev, MU, LAM, True)
    t.stop()
    asm_ms = t.ms()
    pairs = float(N) * N
    gf = {"pairs_per_s": pairs / (asm_ms * 1e-3), "ms": asm_ms, "pairs": pairs,
          "note": "pair = receiver centroid x source TD, strike and dip slip, half space, traction projection"}
    fpp = assembly_flops_per_pair()
    if fpp:
        tf = pairs * fpp / (asm_ms * 1e-3) / 1e12
        gf.update({"fp64_flops_per_pair": fpp, "fp64_tflops": tf, "fp64_peak_tflops": FP64_PEAK_TFLOPS,
                   "frac_fp64_peak": tf / FP64_PEAK_TFLOPS,
                   "flop_convention": "source-level count of the operations the kernel runs (a free-surface "
                                      "side shared by consecutive sources evaluated once), every "
                                      "add/sub/mul/div/sqrt/log/atan = 1"})
    fpp_all = assembly_flops_per_pair(every_side=True)
    if fpp_all:
        tf = pairs * fpp_all / (asm_ms * 1e-3) / 1e12
        gf["every_side_equivalent"] = {
            "fp64_flops_per_pair": fpp_all, "fp64_tflops": tf, "frac_fp64_peak": tf / FP64_PEAK_TFLOPS,
            "note": "the same timing priced at the count with every side evaluated for every pair (the "
                    "reference's operation sequence): an effective rate, not executed flops"}
    gf["pmc_counted"] = _pmc_fp64("gf_assemble_20k_fp64_pmc.json", pairs, asm_ms)
    log(f"[bench] assembly {asm_ms:.1f} ms, {gf['pairs_per_s']:.3e} pairs/s")

    # --- plain single simulation, FP32 storage
    init_state(sim, cols)
    tuning = {"matvec_f32": _tuned(sim, autotune.tune_dense(sim, use_cache=False), 256)}
    zmask, skip = sim.structure
    nmat = 4 if skip else 6   # Bs = Bd = 0 on a planar fault: the products skip them
    log(f"[bench] zero matrices: {[n for k, n in enumerate(MATRIX_NAMES) if zmask >> k & 1]}; "
        f"products stream {nmat} of 6")
    mv_ms = sim.time_matvec(20)
    p32_ms = time_plain(sim, args.plain_steps, args.warmup, args.graph_steps)
    plain32 = args.plain_steps / (p32_ms * 1e-3)
    mv_alg = nmat * N * N * 4.0 + 4.0 * N * 8 + 3.0 * ((N + 255) // 256) * N * 8
    mv_gbs = mv_alg / (mv_ms * 1e-3) / 1e9
    log(f"[bench] plain FP32: {plain32:.1f} member-steps/s; matvec {mv_ms:.3f} ms = {mv_gbs:.0f} GB/s")

    # --- batched ensemble on the same FP32 matrices (the measured value)
    mcols = np.stack([synth.member_params(xg, m) for m in range(MEMBERS)])
    ens = Ensemble(N, MEMBERS, matrices_from=sim)
    init_members(ens, mcols, V0, VPL, MU, CS)
    kc0 = ens.kc
    tuning["ens_gemm"] = _tuned(ens, autotune.tune_ensemble(ens, use_cache=False), kc0)
    precision, split = ens.precision
    ens_ms, accepted = time_ensemble(ens, args.steps, args.warmup, args.graph_steps)
    value = accepted / (ens_ms * 1e-3)
    gemm_ms = ens.gemm(reps=10)
    mem = ens.members()
    gemm_flops = 2.0 * nmat * N * N * MEMBERS
    gemm_tf = gemm_flops / (gemm_ms * 1e-3) / 1e12
    gemm_bytes = nmat * N * N * 4.0 + 2.0 * N * MEMBERS * 4 + 3.0 * N * MEMBERS * 4
    gemm_gbs = gemm_bytes / (gemm_ms * 1e-3) / 1e9
    log(f"[bench] ensemble ({precision}): {value:.1f} member-steps/s ({args.steps} global steps in {ens_ms:.1f} ms); "
        f"GEMM {gemm_ms:.3f} ms = {gemm_tf:.1f} TFLOP/s of the product, {gemm_gbs:.0f} GB/s")
    ens_info = {"members": MEMBERS, "global_steps": args.steps, "ms_per_global_step": ens_ms / args.steps,
                "gemm_ms": gemm_ms, "gemm_precision": precision, "gemm_bf16_split": split,
                "member_sim_time_s_min": float(mem["time"].min()),
                "member_sim_time_s_max": float(mem["time"].max()), "failed_members": int((mem["status"] != 0).sum())}
    if args.rupture_steps > 0:
        ens_info["rupture"] = rupture_ensemble_leg(sim, xg, timed=args.rupture_steps, graph_steps=args.graph_steps)
        r = ens_info["rupture"]
        log(f"[bench] ensemble in rupture: {r['value']:.1f} member-steps/s over {r['global_steps']} global steps "
            f"after {r['steps_before_window']} more; members above Vdyn {r['members_above_vdyn_at_start']} -> "
            f"{r['members_above_vdyn_at_end']}, peak {r['peak_slip_rate_m_per_s']:.3g} m/s")
    # traction rates of the members' current slip rates through the batched
    # path (FP32 storage, MFMA GEMM), checked below against the FP64 product
    ens.gemm()
    gemm_out = ens.gemm_result()
    member_u = {k: ens.get(k) for k in ("slipv1", "slipv2", "rake0")}
    ens.close()
    # the same ensemble with the FP32-MFMA GEMM
    e32 = Ensemble(N, MEMBERS, matrices_from=sim, precision="fp32")
    init_members(e32, mcols, V0, VPL, MU, CS)
    tuning["ens_gemm_fp32"] = _tuned(e32, autotune.tune_ensemble(e32, use_cache=False), kc0)
    f32_ms, f32_acc = time_ensemble(e32, args.steps, args.warmup, args.graph_steps)
    f32_gemm_ms = e32.gemm(reps=10)
    f32_tf = gemm_flops / (f32_gemm_ms * 1e-3) / 1e12
    e32.gemm()
    gemm_out32 = e32.gemm_result()
    member_u32 = {k: e32.get(k) for k in ("slipv1", "slipv2", "rake0")}
    e32.close()
    fp32_gemm = {"value": f32_acc / (f32_ms * 1e-3), "gemm_ms": f32_gemm_ms,
                 "roofline": {"bound": "mfma", "achieved": f32_tf, "peak": FP32_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                              "frac": f32_tf / FP32_MFMA_PEAK_TFLOPS,
                              "kernel": "ens_gemm_kernel (v_mfma_f32_16x16x4f32)",
                              "pmc": _pmc_mfma("ens_gemm_20k_mfma_pmc.json"),
                              "traffic": _pmc_traffic("ens_gemm_20k_traffic.json")}}
    log(f"[bench] same ensemble, FP32-MFMA GEMM: {fp32_gemm['value']:.1f} member-steps/s, GEMM {f32_gemm_ms:.3f} ms "
        f"= {f32_tf:.1f} TFLOP/s")
    elision = {"zero_matrices": [n for k, n in enumerate(MATRIX_NAMES) if zmask >> k & 1],
               "matrices_streamed": nmat}
    if skip:
        # the same ensemble streaming all six matrices (results bit-identical)
        full = Ensemble(N, MEMBERS, matrices_from=sim)
        full.exploit_structure = False
        init_members(full, mcols, V0, VPL, MU, CS)
        autotune.tune_ensemble(full, use_cache=False)
        f_ms, f_acc = time_ensemble(full, args.steps, args.warmup, args.graph_steps)
        elision.update({"value_streaming_all_six": f_acc / (f_ms * 1e-3), "gemm_ms_all_six": full.gemm(reps=10),
                        "kc_all_six": full.kc})
        full.close()
        log(f"[bench] same ensemble streaming all six matrices: {elision['value_streaming_all_six']:.1f} "
            f"member-steps/s, GEMM {elision['gemm_ms_all_six']:.3f} ms")
    elision["note"] = ("a matrix holding only zeros (Bs and Bd of a planar vertical fault: the normal traction of "
                       "in-plane slip vanishes by mirror symmetry) is found by a device scan after assembly and "
                       "skipped by the dense product and the GEMM; results are bit-identical to streaming it")
    sim.close()

    # --- plain single simulation, FP64 storage (denominator of batched vs plain)
    s64 = DenseSim(N, _lib.MAT_F64)
    s64.assemble(*P, xg, ev, MU, LAM, True)
    init_state(s64, cols)
    tuning["matvec_f64"] = _tuned(s64, autotune.tune_dense(s64, use_cache=False), 256)
    k64 = max(100, args.plain_steps // 4)
    p64_ms = time_plain(s64, k64, args.warmup, args.graph_steps)
    plain64 = k64 / (p64_ms * 1e-3)
    mv64_ms = s64.time_matvec(10)
    log(f"[bench] plain FP64: {plain64:.1f} member-steps/s; matvec {mv64_ms:.3f} ms")
    terr = traction_error(s64, gemm_out, member_u)
    terr32 = traction_error(s64, gemm_out32, member_u32)
    fp32_gemm["traction_error"] = {k: terr32[k] for k in ("max_rel_l2", "max_rel_max")}
    del gemm_out, member_u, gemm_out32, member_u32
    log(f"[bench] batched traction rates vs FP64 product: relative L2 error {terr['max_rel_l2']:.2e}, "
        f"max-norm {terr['max_rel_max']:.2e} over {len(terr['members'])} members")
    cpu = None
    if not args.no_cpu_baseline:
        mats64 = [s64.download_matrix(k) for k in range(6)]
        s64.close()
        cpu = cpu_baseline_step(cols, mats64)
        del mats64
        log(f"[bench] CPU baseline {cpu['value']:.3f} member-steps/s on {cpu['cores']} threads")
        asm_cpu = cpu_baseline_assembly(P, xg)
        if asm_cpu:
            gf["cpu_baseline"] = asm_cpu
    else:
        s64.close()

    serving = None
    if args.serve_requests > 0:
        serving = serve(args.serve_requests, args.serve_rate)
        log(f"[bench] serving: {serving['requests']} requests, p50 {serving['p50_time_to_result_s']:.2f} s, "
            f"p99 {serving['p99_time_to_result_s']:.2f} s, {serving['member_steps_per_s']:.0f} member-steps/s, "
            f"occupancy {serving['mean_occupancy']:.2f} (30k mesh {serving['occupancy_per_mesh']['30k']:.2f})")

    config1 = None if args.no_config1 else config1_leg(graph_steps=args.graph_steps,
                                                       cpu_budget_s=0.0 if args.no_cpu_baseline else 8.0)

    hmat = None
    if not args.no_hmat:
        hmat = hmatrix_leg(steps=args.hmat_steps, warmup=args.warmup, graph_steps=args.graph_steps,
                           cpu_budget_s=0.0 if args.no_cpu_baseline else 20.0)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "member-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ens_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "batched ensemble: 64 members (seeded a-b, Dc, sigma_n) on a planar vertical fault, "
                               "40x40 km, 100x100 grid -> 20,000 TDs, half space; FP64 GF assembly, FP32 GF storage, "
                               "GEMM with FP32 operands split exactly into three BF16 parts (six products, FP32 "
                               "accumulation, BF16 MFMA); aging-law RSF with radiation damping, adaptive RK45 per "
                               "member; a step = one global step (one accepted step of every member)",
                   "n_elements": N, "members": MEMBERS, "gf_storage": "fp32", "state_dtype": "f64",
                   "parallelism": "single GPU"},
        "roofline": ({"bound": "hbm", "achieved": gemm_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": gemm_gbs / HBM_PEAK_GBS, "traffic": _pmc_traffic("ens_gemm3_20k_traffic.json"),
                      "kernel": f"ens_gemm3_kernel (FP32 operands split into three BF16 parts, six products on "
                                f"v_mfma_f32_16x16x32_bf16; {nmat} x 20k x 20k FP32 by 20k x 64; "
                                f"{6 - nmat} all-zero matrices skipped)",
                      "alg_bytes_per_launch": gemm_bytes, "avg_launch_ms": gemm_ms,
                      "product_tflops": gemm_tf, "pmc": _pmc_mfma("ens_gemm3_20k_mfma_pmc.json")}
                     if split else
                     {"bound": "mfma", "achieved": gemm_tf, "peak": FP32_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                      "frac": gemm_tf / FP32_MFMA_PEAK_TFLOPS, "traffic": _pmc_traffic("ens_gemm_20k_traffic.json"),
                      "kernel": f"ens_gemm_kernel (v_mfma_f32_16x16x4f32, {nmat} x 20k x 20k FP32 by 20k x 64; "
                                f"{6 - nmat} all-zero matrices skipped)",
                      "flops_per_launch": gemm_flops, "avg_launch_ms": gemm_ms,
                      "pmc": _pmc_mfma("ens_gemm_20k_mfma_pmc.json")}),
        "fp32_mfma_gemm": fp32_gemm,
        "cpu_baseline": cpu,
        "plain_fp32_member_steps_per_s": plain32,
        "plain_fp64_member_steps_per_s": plain64,
        "batched_vs_plain_fp64": value / plain64,
        "batched_vs_plain_fp32": value / plain32,
        "plain_roofline": {"bound": "hbm", "achieved": mv_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": mv_gbs / HBM_PEAK_GBS, "traffic": _pmc_traffic("matvec_f32_20k_traffic.json"),
                           "kernel": "matvec_kernel<float> (six FP32 traction matrices, FP64 accumulation)",
                           "alg_bytes_per_launch": mv_alg, "avg_launch_ms": mv_ms},
        "ensemble": ens_info,
        "zero_matrix_elision": elision,
        "traction_error": terr,
        "autotune": tuning,
        "gf_assembly": gf,
        "p99_solve_latency_s": serving["p99_time_to_result_s"] if serving else None,
        "serving": serving,
        "config1": config1,
        "hmatrix": hmat,
    }
    out["measured_peaks"] = peaks
    for r in (out["roofline"], out["plain_roofline"], (hmat or {}).get("roofline")):
        if r and r.get("unit") == "GB/s":
            r["frac_of_measured"] = r["achieved"] / peaks["hbm_read_gbs"]
    for g in (gf, gf.get("pmc_counted")):
        if g and g.get("fp64_tflops"):
            g["frac_of_measured_fp64"] = g["fp64_tflops"] / peaks["fp64_fma_tflops"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
