"""Write tests/golden/dilatancy_golden.npz: input/output vectors of the
reference's own pore-pressure functions, QDsim.Calc_Pmatrix and
QDsim.Calc_P_implicit (src/QDsim_gpu.py:611-671).

The functions are imported from the reference in THIS container and called
on a plain attribute object carrying exactly the attributes they read (no
simulation is built).  The reference code is not copied; only its outputs are
stored.  Three cases of 40 elements each, with seeded states and slip rates
spanning the interseismic and coseismic range, and step sizes 1e-3 s, 1e5 s
and 3e7 s; each case runs two consecutive updates so the second starts from a
non-uniform profile.
"""
import os
import sys
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_SRC = "/root/reference/src"

N = 40
NUM = 60
C = 1e-2
PARAMS = {"DilatancyC": 3e-4, "hs": 3e-3, "EPermeability": 8e-11, "Chyd": 1e-5}
P_INIT = 2.0e6     # initial pore pressure (Pa), parameter file value 2 MPa
P_CONST = 2.0e6    # fixed far-field pore pressure (Pa)
V0 = 1e-6


def main():
    os.environ.setdefault("MPLBACKEND", "Agg")
    sys.path.insert(0, REF_SRC)
    import QDsim_gpu as ref  # the reference module
    rng = np.random.default_rng(21)
    yp = np.logspace(start=-3, stop=10, num=NUM, base=10)
    zp = np.log(1 + yp / C)
    out = {"zp": zp, "P0": np.array(P_CONST), "V0": np.array(V0)}
    for k, v in PARAMS.items():
        out[k] = np.array(v)
    base = types.SimpleNamespace(zp=zp, Chyd=PARAMS["Chyd"], Pmatrix=np.zeros([NUM, NUM]))
    ref.QDsim.Calc_Pmatrix(base)
    out["Pmatrix"] = base.Pmatrix.copy()
    for case, dt in enumerate((1e-3, 1e5, 3e7)):
        a = rng.uniform(0.004, 0.02, N)
        b = rng.uniform(0.01, 0.03, N)
        dc = 0.14 * np.exp(0.3 * rng.standard_normal(N))
        f0 = np.full(N, 0.6)
        slipv = 10.0 ** rng.uniform(-12, 0, N)
        state = a * np.log(2 * V0 / slipv) + rng.normal(0, 0.02, N) + 0.6
        ns = types.SimpleNamespace(
            DilatancyC=PARAMS["DilatancyC"], hs=PARAMS["hs"], EPermeability=PARAMS["EPermeability"],
            Chyd=PARAMS["Chyd"], dc=dc, zp=zp, Pmatrix=base.Pmatrix.copy(), useGPU=False, slipv=slipv,
            state=state, f0=f0, b=b, V0=V0, eleVec=np.zeros((N, 9)), Parr=np.ones([N, NUM]) * P_INIT,
            P0=P_CONST, dPdt0=np.zeros(N), P=np.ones(N) * P_INIT)
        pre = {"a": a, "b": b, "dc": dc, "f0": f0, "slipv": slipv, "state": state, "dt": np.array(dt)}
        for key, v in pre.items():
            out[f"c{case}_{key}"] = v
        for it in range(2):
            out[f"c{case}_Parr_in{it}"] = ns.Parr.copy()
            ref.QDsim.Calc_P_implicit(ns, dt)
            out[f"c{case}_Parr_out{it}"] = ns.Parr.copy()
            out[f"c{case}_P_out{it}"] = ns.P.copy()
            out[f"c{case}_dPdt0_out{it}"] = ns.dPdt0.copy()
    np.savez_compressed(os.path.join(HERE, "dilatancy_golden.npz"), **out)
    print("wrote", os.path.join(HERE, "dilatancy_golden.npz"))


if __name__ == "__main__":
    main()
