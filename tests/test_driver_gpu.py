"""The reference driver's loop (src/main_gpu.py:100-251) over the QDsim
mirror, both branches, restated here (not copied): every attribute and method
of ``sim0`` the driver uses must exist and behave, and the run directory it
leaves -- state.txt in the driver's row formats, out_vtk snapshots -- feeds the
event catalogue.

* GPU branch (``useGPU``): initGPUvariable, simu_forwardGPU, the *_gpu
  device fields' .get(), time, Para0 keys, ouputVTK every outsteps steps;
  its steps follow the NumPy FP64 oracle (1e-9).
* CPU branch: simu_forward, the host attributes it refreshes (slipv, Tt),
  Relerrormax1 / Relerrormax2 in every state.txt row.
* On this planar fault the two error estimates coincide (the normal traction
  does not change), so both branches take the same steps: equal time columns.
"""
import os

import numpy as np
import pytest

from oracle import qdsim_oracle as Q

pytestmark = pytest.mark.gpu

MU = 32038120320.0
V0, VPL = 1e-6, 1e-9
NSTEPS = 60


def _drive(sim0, use_gpu, run_dir, msh, par):
    """One pass of the driver loop; returns the per-step (dttry, dtnext)."""
    from pyquake3d_amd import outputs
    log = outputs.StateLog(os.path.join(run_dir, "state.txt"), msh, par)
    sim0.ouputVTK(os.path.join(run_dir, "Init.vtk"))
    log.columns()
    total = min(int(sim0.Para0["totaloutputsteps"]), NSTEPS)
    outsteps = int(sim0.Para0["outsteps"])
    vtk_dir = os.path.join(run_dir, "out_vtk")
    os.makedirs(vtk_dir, exist_ok=True)
    steps = []
    dtnext = None
    if use_gpu:
        sim0.initGPUvariable()
    for i in range(total):
        dttry = sim0.htry if i == 0 else dtnext
        if use_gpu:
            dttry, dtnext = sim0.simu_forwardGPU(dttry)
            for name in ("slipv1", "slipv2", "slipv", "slip", "Tt1o", "Tt2o", "Tt", "rake"):
                setattr(sim0, name, getattr(sim0, name + "_gpu").get())
            log.row_gpu(i, dttry, sim0)
            if i % outsteps == 0:
                for name in ("slipv1", "slipv2", "slip1", "slip2", "Tno", "Tt1o", "Tt2o", "P"):
                    setattr(sim0, name, getattr(sim0, name + "_gpu").get())
        else:
            dttry, dtnext = sim0.simu_forward(dttry)
            log.row_cpu(i, dttry, sim0)
            assert sim0.Tt.shape == sim0.slipv.shape
        if i % outsteps == 0 and sim0.Para0["outputvtu"] == "True":
            sim0.ouputVTK(os.path.join(vtk_dir, "step%d.vtk" % i))
        assert sim0.Para0["outputSLIPV"] in ("True", "False")
        steps.append((dttry, dtnext))
    log.close(1.0)
    return np.array(steps)


def test_reference_driver_both_branches(tmp_path):
    from pyquake3d_amd import QDsim_gpu, outputs, readmsh, synth
    msh, par = synth.write_case(str(tmp_path / "case"), 12, 6, 24.0, 12.0, dc_sigma=0.2, seed=1)
    nodelst, elelst = readmsh.read_mshV2(msh)
    runs = {}
    for use_gpu in (True, False):
        run_dir = str(tmp_path / ("gpu" if use_gpu else "cpu"))
        os.makedirs(run_dir)
        sim0 = QDsim_gpu.QDsim(elelst, nodelst, par)
        assert sim0.useGPU is True            # the parameter file's GPU key
        runs[use_gpu] = (_drive(sim0, use_gpu, run_dir, msh, par), outputs.read_state(
            os.path.join(run_dir, "state.txt")), sim0)
        assert sorted(os.listdir(os.path.join(run_dir, "out_vtk"))) == ["step0.vtk", "step50.vtk"]
        events, text = outputs.extract_slip_info(run_dir)
        assert events == [] and text.startswith("Evnt")     # interseismic: no event
    (sg, rg, simg), (sc, rc, simc) = runs[True], runs[False]
    assert rg.shape == (NSTEPS, 6) and rc.shape[0] == NSTEPS
    assert np.array_equal(rg[:, 0], np.arange(NSTEPS)) and np.array_equal(rc[:, 0], np.arange(NSTEPS))
    assert np.array_equal(sg, sc)                       # same steps in both branches
    assert np.allclose(rg[:, 4], rc[:, 3], rtol=0, atol=1e-6)   # time columns ('%f')
    # CPU rows carry the error measures (printed '%f'): below RelTol
    assert np.all(rc[:, 5] <= 1e-4) and np.all(rc[:, 6] <= 1e-4)
    assert simc.Relerrormax1 > 1e-10 and simc.Relerrormax2 > 1e-10
    # the GPU branch's steps against the oracle
    core = simg.Para0["Corefunc directory"]
    orc = Q.from_init_columns([np.load(os.path.join(core, n + ".npy")) for n in
                               ("A1s", "A2s", "Bs", "A1d", "A2d", "Bd")],
                              np.loadtxt(os.path.join(os.path.dirname(par), "param.dat")), V0, VPL, MU,
                              np.sqrt(MU / simg.density))
    dt = 1e-3
    for k in range(NSTEPS):
        h, dt = orc.step(dt)
        assert abs(h - sg[k, 0]) <= 1e-9 * h
    assert abs(simg.time - orc.time) <= 1e-9 * orc.time
    assert np.abs(simg.slipv1 - orc.slipv1).max() <= 1e-9 * np.abs(orc.slipv1).max()
