"""Benchmark of the PyQuake3D hot path on one MI355X.

Workload (N = 1): planar vertical strike-slip fault, 40 km x 40 km, 100 x 100
squares split into 20,000 triangular dislocations, half space; aging-law
rate-and-state friction (b = 0.03, Dc = 0.14 m with a seeded lognormal
perturbation of sigma 0.2, a = 0.004 in a central 30 x 15 km patch rising to
0.04 outside), sigma_n = 25 MPa, Vpl = V_init = 1e-9 m/s, state at steady state.

* Green's functions: FP64 triangular-dislocation assembly on the GPU (6 dense
  20k x 20k traction matrices), stored in FP32; reported as pairs/s and FP64
  FLOP/s.
* Stepping: adaptive Dormand-Prince RK45 with the reference's accept/reject
  rule; one *step* = one accepted step of the whole simulation (one member).
  value = accepted member-steps per second over K timed steps, device-resident,
  no host synchronisation inside the timed region.
* roofline: the traction matvec (dominant kernel): algorithmic bytes per launch
  (six 20k x 20k FP32 matrices + the slip-rate vectors) / its average launch
  time from HIP events; peak 8 TB/s.
* cpu_baseline: the NumPy FP64 restatement of the reference's dense step
  (oracle/qdsim_oracle.py, six np.dot per step, BLAS on all host cores) on the
  same 20k matrices, timed on a bounded number of steps.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
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


def time_steps(sim, steps, warmup, graph_steps):
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


def assembly_flops_per_pair():
    """FP64 operations per source-receiver pair (both slip directions) from the
    counting build of the kernel's own source (pyquake3d_amd/csrc/flopcount)."""
    p = os.path.join(ROOT, "pyquake3d_amd", "csrc", "flops_per_pair.txt")
    try:
        with open(p) as f:
            return float(f.read().split()[0])
    except (OSError, ValueError, IndexError):
        return None


def cpu_baseline_step(case, sim64_mats, budget_s=15.0, max_steps=40):
    """NumPy FP64 restatement of the dense step on the host (oracle)."""
    from oracle import qdsim_oracle as Q
    threads = len(os.sched_getaffinity(0))
    orc = Q.from_init_columns(sim64_mats, case[4], V0, VPL, MU, CS)
    dt = 1e-3
    orc.step(dt)  # untimed first step
    dt = 1.5e-3
    t0 = time.perf_counter()
    n = 0
    while n < max_steps and time.perf_counter() - t0 < budget_s:
        h, dt = orc.step(dt)
        n += 1
    el = time.perf_counter() - t0
    return {"value": n / el, "unit": "member-steps/s", "cores": threads, "kind": "port",
            "sample": f"{n} steps of the NumPy FP64 restatement of simu_forward (6 np.dot per step) on the "
                      f"same 20,000-element FP64 matrices, BLAS threads = {threads}, CPU {_cpu_model()}"}


def cpu_baseline_assembly(case, budget_s=10.0):
    """Reference C++ TDstressFS_C (built from source into oracle/_ref) over
    sampled source columns, one process per host core."""
    from oracle import refbind
    nodes, tri, ev, xg, cols, P = case
    if not refbind.available():
        return None
    import multiprocessing as mp
    threads = len(os.sched_getaffinity(0))
    ncol = 2 * threads
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(threads) as pool:
        pool.map(_ref_column, [(P[0][j], P[1][j], P[2][j], xg) for j in range(ncol)])
    el = time.perf_counter() - t0
    pairs = ncol * len(xg)
    return {"value": pairs / el, "unit": "pairs/s", "cores": threads, "kind": "reference",
            "sample": f"{ncol} source columns x {len(xg)} receivers through the reference TDstressFS_C "
                      f"(strike and dip calls), {threads} processes"}


def _ref_column(args):
    from oracle import refbind
    P1, P2, P3, xg = args
    for Ss, Ds in ((1.0, 0.0), (0.0, 1.0)):
        refbind.td_stress_fs_c(xg[:, 0], xg[:, 1], xg[:, 2], P1, P2, P3, Ss, Ds, 0.0, MU, LAM, True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--graph-steps", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--skip-fp64", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only; multi-GPU sharding is described in DESIGN.md",
              file=sys.stderr)
        sys.exit(2)

    from pyquake3d_amd import _lib
    from pyquake3d_amd.QDsim_gpu import DenseSim

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    case = build_case()
    nodes, tri, ev, xg, cols, P = case
    N = len(tri)
    log(f"[bench] N={N} elements")

    # --- Green's function assembly (FP64 arithmetic, FP32 storage)
    sim = DenseSim(N, _lib.MAT_F32)
    t = _lib.Timer()
    sim.assemble(*P, xg, ev, MU, LAM, True)   # first call includes code load
    t.start()
    sim.assemble(*P, xg, ev, MU, LAM, True)
    t.stop()
    asm_ms = t.ms()
    pairs = float(N) * N
    fpp = assembly_flops_per_pair()
    gf = {"pairs_per_s": pairs / (asm_ms * 1e-3), "ms": asm_ms, "pairs": pairs,
          "note": "pairs = receiver x source; each pair evaluates strike and dip slip"}
    if fpp:
        tf = pairs * fpp / (asm_ms * 1e-3) / 1e12
        gf.update({"fp64_flops_per_pair": fpp, "fp64_tflops": tf, "fp64_peak_tflops": FP64_PEAK_TFLOPS,
                   "frac_fp64_peak": tf / FP64_PEAK_TFLOPS})
    log(f"[bench] assembly {asm_ms:.1f} ms, {gf['pairs_per_s']:.3e} pairs/s")

    # --- plain single-simulation stepping, FP32 storage (the measured value)
    init_state(sim, cols)
    mv_ms = sim.time_matvec(20)
    step_ms = time_steps(sim, args.steps, args.warmup, args.graph_steps)
    mv_ms_after = sim.time_matvec(20)
    st = sim.state()
    value = args.steps / (step_ms * 1e-3)
    log(f"[bench] {args.steps} steps in {step_ms:.1f} ms -> {value:.1f} member-steps/s; "
        f"matvec {mv_ms:.3f}/{mv_ms_after:.3f} ms; sim time {st['time']:.3e} s")
    mv_avg = 0.5 * (mv_ms + mv_ms_after)
    alg_bytes = 6.0 * N * N * 4 + 4.0 * N * 8 + 3.0 * sim_chunks(sim) * N * 8
    achieved = alg_bytes / (mv_avg * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": _pmc_traffic(),
                "kernel": "matvec_kernel<float> (six FP32 traction matrices, FP64 accumulation)",
                "alg_bytes_per_launch": alg_bytes, "avg_launch_ms": mv_avg}

    # --- plain FP64 path (denominator of the batched-vs-plain ratio)
    extra = {}
    sim64_mats = None
    if not args.skip_fp64:
        sim.close()
        s64 = DenseSim(N, _lib.MAT_F64)
        s64.assemble(*P, xg, ev, MU, LAM, True)
        init_state(s64, cols)
        k64 = max(50, args.steps // 8)
        ms64 = time_steps(s64, k64, args.warmup, args.graph_steps)
        extra["plain_fp64_member_steps_per_s"] = k64 / (ms64 * 1e-3)
        extra["plain_fp64_matvec_ms"] = s64.time_matvec(10)
        log(f"[bench] FP64 storage: {extra['plain_fp64_member_steps_per_s']:.1f} member-steps/s")
        if not args.no_cpu_baseline:
            sim64_mats = [s64.download_matrix(k) for k in range(6)]
        s64.close()

    cpu = None
    if not args.no_cpu_baseline and sim64_mats is not None:
        cpu = cpu_baseline_step(case, sim64_mats)
        del sim64_mats
        log(f"[bench] CPU baseline {cpu['value']:.3f} member-steps/s on {cpu['cores']} threads")
        asm_cpu = cpu_baseline_assembly(case)
        if asm_cpu:
            gf["cpu_baseline"] = asm_cpu

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "member-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "planar vertical fault 40x40 km, 100x100 grid -> 20,000 TDs, half space; "
                               "FP64 GF assembly, FP32 GF storage (FP64 accumulation); aging-law RSF, "
                               "RK45; plain single simulation (1 member)",
                   "n_elements": N, "members": 1, "gf_storage": "fp32", "state_dtype": "f64",
                   "matrices_streamed_per_step": 6, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "gf_assembly": gf,
        "p99_solve_latency_s": None,
        "sim_time_s": st["time"],
        "rejections": None,
    }
    out.update(extra)
    print(json.dumps(out))


def sim_chunks(sim):
    return (sim.n + 255) // 256


def _pmc_traffic():
    """HBM bytes per matvec launch from a rocprofv3 --pmc summary committed
    under profiles/ (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel
    calibration recorded there), or None."""
    p = os.path.join(ROOT, "profiles", "matvec_f32_20k_traffic.json")
    try:
        with open(p) as f:
            return float(json.load(f)["hbm_bytes_per_launch"])
    except (OSError, ValueError, KeyError):
        return None


if __name__ == "__main__":
    main()
