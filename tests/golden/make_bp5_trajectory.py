"""Write tests/golden/bp5_trajectory.npz: the first steps of the reference's
own CPU stepping (QDsim_gpu.QDsim.simu_forward, src/QDsim_gpu.py:1029-1122) on
its BP5-QD example (examples/BP5-QD: 9,250 TDs, half space, per-element
parameter file with a nucleation patch).

The mesh, parameter file and parameter table are the ones stored in
init_golden.npz (make_init_golden.py).  The Green's functions are the
reference's C++ (TDstressFS_C built from source into oracle/_ref), fed
through the reference's own .npy cache; their SHA-256 is stored.  The
reference module is imported in THIS container; only its outputs are kept:
per step time, h, dtnext and max slip rate.
"""
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_SRC = "/root/reference/src"
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import make_qdsim_trajectory as T  # noqa: E402  (Green's functions from the reference C++)
from oracle import refbind  # noqa: E402

STEPS = 300


def bp5_inputs():
    g = np.load(os.path.join(HERE, "init_golden.npz"), allow_pickle=False)
    text = g["bp5_parameter_txt"].tobytes().decode()
    return g["bp5_nodes"], g["bp5_tri"], text, str(g["bp5_param_name"]), g["bp5_param_table"]


def with_settings(text, core_dir, save, gpu=False):
    """The example's parameter text with the cache directory, cache saving and
    GPU switch replaced (nothing else changes)."""
    out = []
    for line in text.splitlines():
        key = line.split(":", 1)[0].strip()
        if key == "Corefunc directory":
            line = f"Corefunc directory: {core_dir}"
        elif key == "save Corefunc":
            line = f"save Corefunc: {save}"
        elif key == "GPU":
            line = f"GPU:{gpu}"
        out.append(line)
    return "\n".join(out) + "\n"


def main():
    if not refbind.available():
        raise SystemExit("build the reference C++ first: make -C oracle ref")
    os.environ.setdefault("MPLBACKEND", "Agg")
    nodes, tri, text, pname, table = bp5_inputs()
    mats = T.traction_matrices(nodes, tri)
    work = tempfile.mkdtemp(prefix="bp5_golden_")
    try:
        core = os.path.join(work, "core")
        os.makedirs(core)
        for k, name in enumerate(("A1s", "A2s", "Bs", "A1d", "A2d", "Bd")):
            np.save(os.path.join(core, name + ".npy"), mats[k])
        with open(os.path.join(work, "parameter.txt"), "w") as f:
            f.write(with_settings(text, core, False))
        np.savetxt(os.path.join(work, pname), table, fmt="%.17g")
        sys.path.insert(0, REF_SRC)
        import QDsim_gpu as ref  # the reference module
        sim = ref.QDsim(tri, nodes, os.path.join(work, "parameter.txt"))
        rec = np.zeros((STEPS, 4))
        dt = sim.htry
        for k in range(STEPS):
            h, dtnext = sim.simu_forward(dt)
            rec[k] = (sim.time, h, dtnext, np.max(sim.slipv))
            dt = dtnext
    finally:
        shutil.rmtree(work, ignore_errors=True)
    np.savez_compressed(os.path.join(HERE, "bp5_trajectory.npz"), records=rec, steps=np.array(STEPS),
                        gf_sha256=np.array(T.matrices_sha256(mats)))
    print("wrote bp5_trajectory.npz; max V", rec[:, 3].max(), "t_end", rec[-1, 0])


if __name__ == "__main__":
    main()
