"""Write tests/golden/dilatancy_trajectory.npz: the reference's own CPU
stepping with dilatancy on (QDsim_gpu.QDsim.simu_forward followed by its
Calc_P_implicit after every accepted step, src/QDsim_gpu.py:1029-1122,
:611-671) on the 1,600-element case of make_qdsim_trajectory.py.

Same mesh, friction, Green's functions (the reference C++, fed through its
.npy cache, SHA-256 stored) and loading as that script; the parameter file
switches `If Dilatancy` on with the fluid values of the reference's BP5
example (dilatancy coefficient 3e-4, hydraulic diffusivity 1e-5, shearing
zone 3e-3, compressibility 8e-11, pore pressures 2 MPa).  The reference
module is imported in THIS container; only its outputs are stored: per step
time, h, dtnext and max slip rate, and the pore pressure at the fault after
steps 100 and the last.
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_SRC = "/root/reference/src"
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import make_qdsim_trajectory as T  # noqa: E402
from oracle import refbind  # noqa: E402
from pyquake3d_amd import synth  # noqa: E402

STEPS = 1500
SNAP = (100, STEPS)


def parameter_text(core):
    text = synth.PARAMETER_TXT.format(core=core, gpu=False, lam=T.LAM, mu=T.MU, rho=synth.RHO, param="param.dat",
                                      halfspace=True, v0=synth.V0, f0=synth.F0, vpl=synth.VPL, steps=STEPS)
    return text.replace("If Dilatancy:False", "If Dilatancy:True")


def main():
    if not refbind.available():
        raise SystemExit("build the reference C++ first: make -C oracle ref")
    os.environ.setdefault("MPLBACKEND", "Agg")
    nodes_km, nodes_m, tri, cols = T.case_inputs()
    mats = T.traction_matrices(nodes_m, tri)
    work = tempfile.mkdtemp(prefix="dila_golden_")
    core = os.path.join(work, "core")
    os.makedirs(core)
    for k, name in enumerate(("A1s", "A2s", "Bs", "A1d", "A2d", "Bd")):
        np.save(os.path.join(core, name + ".npy"), mats[k])
    synth.write_msh(os.path.join(work, "mesh.msh"), nodes_km, tri)
    synth.write_param_file(os.path.join(work, "param.dat"), cols)
    with open(os.path.join(work, "parameter.txt"), "w") as f:
        f.write(parameter_text(core))
    sys.path.insert(0, REF_SRC)
    cwd = os.getcwd()
    os.chdir(work)
    try:
        import QDsim_gpu as ref  # the reference module
        import readmsh as ref_readmsh
        nodelst, elelst = ref_readmsh.read_mshV2(os.path.join(work, "mesh.msh"))
        sim = ref.QDsim(elelst, nodelst, os.path.join(work, "parameter.txt"))
        assert sim.Ifdila
        rec = np.zeros((STEPS, 4))
        snaps = {}
        dt = sim.htry
        for k in range(STEPS):
            h, dtnext = sim.simu_forward(dt)
            rec[k] = (sim.time, h, dtnext, np.max(sim.slipv))
            dt = dtnext
            if k + 1 in SNAP:
                snaps[k + 1] = np.copy(sim.P)
    finally:
        os.chdir(cwd)
    out = {"records": rec, "steps": np.array(STEPS), "gf_sha256": np.array(T.matrices_sha256(mats)),
           "snap_steps": np.array(SNAP)}
    for s, p in snaps.items():
        out[f"P_{s}"] = p
    np.savez_compressed(os.path.join(HERE, "dilatancy_trajectory.npz"), **out)
    print("wrote dilatancy_trajectory.npz; max V", rec[:, 3].max(), "t_end", rec[-1, 0])


if __name__ == "__main__":
    main()
