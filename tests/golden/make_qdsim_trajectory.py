"""Write tests/golden/qdsim_trajectory.npz: a multi-step trajectory of the
reference's own CPU stepping path (QDsim_gpu.QDsim.simu_forward,
src/QDsim_gpu.py:1029-1122).

Multi-step trajectories are chaotic, so only the reference run itself pins
them.  This script imports the reference's Python in THIS container (the
reference tree is read, never copied) and records, per step, what its driver
(src/main.py:190) sees.

Case (the 1,600-element workload): planar vertical fault, 80 x 40 km, 40 x 20
squares of 2 km split into 1,600 TDs, half space, mu = lambda = 32038120320 Pa,
rho = mu / 3464^2, V0 = 1e-6, f0 = 0.6, Vpl = 1e-9; friction from
pyquake3d_amd.synth.friction_params(xg, dc_sigma=0.2, seed=3, rake=0.5) (a
mixed rake, so the dip traction is not rounding noise; with rake 0 the
reference's relative error estimate of that component is set by summation
order).  The Green's functions are the reference's C++ TDstressFS_C built from
source (oracle/_ref), projected with the element frames of readmsh.get_eleVec,
written as the reference's Corefunc .npy cache (float64, row = receiver,
column = source) so the reference loads them instead of recomputing; their
SHA-256 is stored so a test can check it rebuilt the same matrices.

Stored per step k (driver: h, dtnext = simu_forward(dt); dt = dtnext):
time, h, dtnext, max slip rate; plus the state and slip-rate vectors after
steps 50, 500 and the last step.
"""
import hashlib
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_SRC = "/root/reference/src"
sys.path.insert(0, ROOT)

from oracle import refbind  # noqa: E402
from pyquake3d_amd import readmsh, synth  # noqa: E402

MU = LAM = 32038120320.0
STEPS = 2000
SNAP = (50, 500, STEPS)


def _columns(args):
    nodes_m, tri, j0, j1 = args
    ev, xg = readmsh.get_eleVec(nodes_m, tri, False)
    e1, e2, nv = ev[:, 0:3], ev[:, 3:6], ev[:, 6:9]
    out = np.zeros((6, len(tri), j1 - j0))
    for j in range(j0, j1):
        P1, P2, P3 = (nodes_m[tri[j, k] - 1] for k in range(3))
        for b, (Ss, Ds) in enumerate(((1.0, 0.0), (0.0, 1.0))):
            st, _ = refbind.td_stress_fs_c(xg[:, 0], xg[:, 1], xg[:, 2], P1, P2, P3, Ss, Ds, 0.0, MU, LAM, True)
            sxx, syy, szz, sxy, sxz, syz = st.T
            t = np.stack([sxx * nv[:, 0] + sxy * nv[:, 1] + sxz * nv[:, 2],
                          sxy * nv[:, 0] + syy * nv[:, 1] + syz * nv[:, 2],
                          sxz * nv[:, 0] + syz * nv[:, 1] + szz * nv[:, 2]], 1)
            out[3 * b + 0, :, j - j0] = np.sum(t * e1, 1)
            out[3 * b + 1, :, j - j0] = np.sum(t * e2, 1)
            out[3 * b + 2, :, j - j0] = np.sum(t * nv, 1)
    return out


def traction_matrices(nodes_m, tri, workers=None, context="fork"):
    """Six traction Green's functions from the reference C++ (strike, dip),
    source columns split over worker processes (the result does not depend on
    the split).  A process that has opened the GPU passes context="spawn", so
    the workers start fresh instead of inheriting its device handles."""
    import multiprocessing as mp
    n = len(tri)
    workers = workers or min(16, len(os.sched_getaffinity(0)))
    cuts = np.linspace(0, n, workers + 1).astype(int)
    parts = [(nodes_m, tri, int(a), int(b)) for a, b in zip(cuts[:-1], cuts[1:]) if b > a]
    with mp.get_context(context).Pool(len(parts)) as pool:
        blocks = pool.map(_columns, parts)
    return np.concatenate(blocks, axis=2)


def events(rec, thr=1e-3):
    """(onset time, peak slip rate, steps above thr) of every excursion of the
    max slip rate (column 3 of per-step records time, h, dtnext, max V) above
    thr m/s."""
    t, v = rec[:, 0], rec[:, 3]
    out = []
    on = np.nonzero((v[1:] > thr) & (v[:-1] <= thr))[0] + 1
    for i in on:
        j = i
        while j < len(v) and v[j] > thr:
            j += 1
        out.append((t[i], v[i:j].max(), j - i))
    return out


def matrices_sha256(mats):
    return hashlib.sha256(np.ascontiguousarray(mats, dtype=np.float64).tobytes()).hexdigest()


def case_inputs():
    nodes_km, tri = synth.planar_fault_mesh(40, 20, 80.0, 40.0)
    nodes_m = nodes_km * 1e3
    ev, xg = readmsh.get_eleVec(nodes_m, tri, False)
    cols = synth.friction_params(xg, dc_sigma=0.2, seed=3, rake=0.5)
    return nodes_km, nodes_m, tri, cols


def main():
    if not refbind.available():
        raise SystemExit("build the reference C++ first: make -C oracle ref")
    os.environ.setdefault("MPLBACKEND", "Agg")
    nodes_km, nodes_m, tri, cols = case_inputs()
    mats = traction_matrices(nodes_m, tri)
    work = tempfile.mkdtemp(prefix="qdsim_golden_")
    core = os.path.join(work, "core")
    os.makedirs(core)
    for k, name in enumerate(("A1s", "A2s", "Bs", "A1d", "A2d", "Bd")):
        np.save(os.path.join(core, name + ".npy"), mats[k])
    synth.write_msh(os.path.join(work, "mesh.msh"), nodes_km, tri)
    synth.write_param_file(os.path.join(work, "param.dat"), cols)
    with open(os.path.join(work, "parameter.txt"), "w") as f:
        f.write(synth.PARAMETER_TXT.format(core=core, gpu=False, lam=LAM, mu=MU, rho=synth.RHO, param="param.dat",
                                           halfspace=True, v0=synth.V0, f0=synth.F0, vpl=synth.VPL, steps=STEPS))
    sys.path.insert(0, REF_SRC)
    cwd = os.getcwd()
    os.chdir(work)
    try:
        import QDsim_gpu as ref  # the reference module
        import readmsh as ref_readmsh
        nodelst, elelst = ref_readmsh.read_mshV2(os.path.join(work, "mesh.msh"))
        sim = ref.QDsim(elelst, nodelst, os.path.join(work, "parameter.txt"))
        rec = np.zeros((STEPS, 4))
        snaps = {}
        dt = sim.htry
        for k in range(STEPS):
            h, dtnext = sim.simu_forward(dt)
            rec[k] = (sim.time, h, dtnext, np.max(sim.slipv))
            dt = dtnext
            if k + 1 in SNAP:
                snaps[k + 1] = (np.copy(sim.state), np.copy(sim.slipv), np.copy(sim.Tt1o))
    finally:
        os.chdir(cwd)
    out = {"records": rec, "steps": np.array(STEPS), "gf_sha256": np.array(matrices_sha256(mats)),
           "snap_steps": np.array(SNAP)}
    for s, (st, v, t1) in snaps.items():
        out[f"state_{s}"] = st
        out[f"slipv_{s}"] = v
        out[f"Tt1o_{s}"] = t1
    np.savez_compressed(os.path.join(HERE, "qdsim_trajectory.npz"), **out)
    print("wrote", os.path.join(HERE, "qdsim_trajectory.npz"), "max V", rec[:, 3].max(), "t_end", rec[-1, 0])


if __name__ == "__main__":
    main()
