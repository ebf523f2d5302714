"""Write tests/golden/init_golden.npz: the reference's own initial state
(QDsim_gpu.QDsim.__init__ -> Init_condition, src/QDsim_gpu.py:119-186,
:319-609, :714-770) on three of its own example inputs, each taking a
different branch:

* BP5-QD: per-element 10-column parameter file (InputHetoparamter True);
* EAFZ-model: friction zones from the mesh boundary, depth-dependent stress,
  ChangefriA, half space;
* Lab-model: zones, normal traction and rake solved from the stress tensor,
  full space, metre coordinates (Scale_km False).

The reference module is imported in THIS container and constructed with its
Green's-function step (calc_corefunc) replaced by a no-op, so only the initial
state is computed.  Stored per case: the mesh as read by the reference reader
(nodes, connectivity), the parameter file text and, for BP5, the parameter
table (inputs); and the initial fields at 1,500 seeded element indices plus
per-field sums over all elements (outputs).  No reference code is stored.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"
CASES = {"bp5": ("examples/BP5-QD", "bp5t.msh"), "eafz": ("examples/EAFZ-model", "turkey.msh"),
         "lab": ("examples/Lab-model", "lab.msh")}
FIELDS = ("Tt1o", "Tt2o", "Tno", "Tt", "state", "a", "b", "dc", "f0", "slipv1", "slipv2", "slipv", "rake",
          "shear_loading", "normal_loading", "fric")
NSAMPLE = 1500


def main():
    os.environ.setdefault("MPLBACKEND", "Agg")
    sys.path.insert(0, os.path.join(REF, "src"))
    import QDsim_gpu as ref  # the reference module
    import readmsh as ref_readmsh
    ref.QDsim.calc_corefunc = lambda self: None  # initial state only
    out = {"fields": np.array(FIELDS)}
    rng = np.random.default_rng(11)
    for name, (d, msh) in CASES.items():
        folder = os.path.join(REF, d)
        nodelst, elelst = ref_readmsh.read_mshV2(os.path.join(folder, msh))
        param = os.path.join(folder, "parameter.txt")
        sim = ref.QDsim(elelst, nodelst, param)
        n = len(sim.xg)
        idx = np.sort(rng.choice(n, size=min(n, NSAMPLE), replace=False))
        out[f"{name}_nodes"] = np.asarray(nodelst, dtype=np.float64)
        out[f"{name}_tri"] = np.asarray(elelst, dtype=np.int32)
        with open(param, "rb") as f:
            out[f"{name}_parameter_txt"] = np.frombuffer(f.read(), dtype=np.uint8)
        if sim.InputHetoparamter:
            pfile = sim.Para0["Inputparamter file"]
            out[f"{name}_param_table"] = np.loadtxt(os.path.join(folder, pfile), ndmin=2)
            out[f"{name}_param_name"] = np.array(pfile)
        out[f"{name}_idx"] = idx
        for fld in FIELDS:
            v = np.asarray(getattr(sim, fld), dtype=np.float64)
            out[f"{name}_{fld}"] = v[idx]
            out[f"{name}_{fld}_sum"] = np.array(np.sum(v))
            out[f"{name}_{fld}_abssum"] = np.array(np.sum(np.abs(v)))
        print(name, "elements", n, "InputHetoparamter", sim.InputHetoparamter)
    np.savez_compressed(os.path.join(HERE, "init_golden.npz"), **out)
    print("wrote", os.path.join(HERE, "init_golden.npz"))


if __name__ == "__main__":
    main()
