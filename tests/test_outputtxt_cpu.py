"""QDsim.outputtxt, the reference's surface-grid text output
(QDsim_gpu.py:2085-2151), on the host-side state of a synthetic case.

Pinned by what the mathematics fixes rather than by a second copy of the
code: the grid is 300 rows x 500 columns of equally spaced points inset by
the largest element size; cubic (Clough-Tocher) interpolation reproduces a
linear field up to its gradient estimate's iteration tolerance (1e-6), so
every interior value equals the linear function at the grid point to 1e-3 of
the field's range; a swapped axis or the wrong centroid column gives errors
of the order of the whole range."""
import os

import numpy as np
import pytest


def _linear(xg, c0, cx, cz):
    return c0 + cx * xg[:, 0] + cz * xg[:, 2]


def test_outputtxt_grids_and_linear_reproduction(tmp_path, monkeypatch):
    pytest.importorskip("scipy")
    from pyquake3d_amd import QDsim_gpu, readmsh, synth
    msh, par = synth.write_case(str(tmp_path / "case"), 12, 6, 24.0, 12.0, dc_sigma=0.2, seed=1)
    nodelst, elelst = readmsh.read_mshV2(msh)
    sim = QDsim_gpu.QDsim(elelst, nodelst, par, device=False)
    xg = np.asarray(sim.xg)
    coefs = {"slipv": (1.0, 2e-5, 3e-5), "slipv1": (0.5, -1e-5, 2e-5), "slipv2": (2.0, 3e-5, -1e-5),
             "Tt": (30.0, 1e-4, -2e-4)}
    for attr, c in coefs.items():
        setattr(sim, attr, _linear(xg, *c))
    monkeypatch.chdir(tmp_path)
    sim.outputtxt("step7")
    out = tmp_path / "out_txt"
    X = np.loadtxt(out / "X_grid.txt")
    Z = np.loadtxt(out / "Y_grid.txt")
    assert X.shape == Z.shape == (300, 500)
    x1 = np.linspace(xg[:, 0].min() + sim.maxsize, xg[:, 0].max() - sim.maxsize, 500)
    z1 = np.linspace(xg[:, 2].min() + sim.maxsize, xg[:, 2].max() - sim.maxsize, 300)
    assert np.abs(X - x1[None, :]).max() <= 5e-7 and np.abs(Z - z1[:, None]).max() <= 5e-7
    assert np.all(np.diff(X, axis=1) > 0) and np.all(np.diff(Z, axis=0) > 0)
    for attr, tag in (("slipv", "slipv"), ("slipv1", "slipv1"), ("slipv2", "slipv2"), ("Tt", "Traction")):
        g = np.loadtxt(out / f"step7{tag}.txt")
        assert g.shape == (300, 500)
        inside = ~np.isnan(g)
        assert inside.mean() > 0.9
        c0, cx, cz = coefs[attr]
        exact = c0 + cx * x1[None, :] + cz * z1[:, None]
        tol = 1e-3 * (exact.max() - exact.min()) + 5e-6
        assert np.abs(g - exact)[inside].max() <= tol, attr
    # the format: '%.5f ' per value, one grid row per line
    first = open(out / "step7slipv.txt").readline()
    assert first.endswith(" \n") and len(first.split()) == 500
    assert all(len(t.split(".")[1]) == 5 for t in first.split() if t != "nan")
