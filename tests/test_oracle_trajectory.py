"""The oracle's restatement of the reference's CPU step (simu_forward,
src/QDsim_gpu.py:1029-1122) against a 2,000-step trajectory of the reference
itself (tests/golden/qdsim_trajectory.npz, written by
tests/golden/make_qdsim_trajectory.py, which imports the reference's Python).

Both runs use bit-identical Green's functions (the reference C++ built from
source, checked by SHA-256).  Bars:
* steps 1-45: time, h and dtnext equal to 1e-13 and the max slip rate to
  1e-12 (the two runs do the same arithmetic until rounding differences in
  the slip rates first grow);
* steps 46-100: within 1e-9 (chaotic growth of those differences); the
  state after step 50 within 1e-11;
* the whole run: the same seismic events (max slip rate above 1e-3 m/s),
  onset times within 1e-3 (measured 1.9e-5), peak slip rates within 1%.
"""
import os
import sys

import numpy as np
import pytest

from oracle import qdsim_oracle as Q
from oracle import refbind

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
sys.path.insert(0, GOLDEN)

pytestmark = pytest.mark.skipif(not refbind.available(), reason="reference C++ (oracle/_ref) not built")


def compare(rec, ref):
    rel = np.abs(rec - ref) / np.abs(ref)
    return rel.max(axis=1)


@pytest.fixture(scope="module")
def golden_case():
    import make_qdsim_trajectory as G
    g = dict(np.load(os.path.join(GOLDEN, "qdsim_trajectory.npz"), allow_pickle=False))
    nodes_km, nodes_m, tri, cols = G.case_inputs()
    mats = G.traction_matrices(nodes_m, tri)
    assert G.matrices_sha256(mats) == str(g["gf_sha256"])
    return g, mats, cols


def test_oracle_follows_reference_trajectory(golden_case):
    import make_qdsim_trajectory as G
    g, mats, cols = golden_case
    ref = g["records"]
    orc = Q.from_init_columns([mats[k] for k in range(6)], cols, 1e-6, 1e-9, 32038120320.0, 3464.0,
                              gpu_semantics=False)
    rec = np.zeros_like(ref)
    dt = 1e-3
    for k in range(len(ref)):
        h, dtn = orc.step(dt)
        rec[k] = (orc.time, h, dtn, np.max(orc.slipv))
        dt = dtn
        if k + 1 == 50:
            assert np.abs(orc.state - g["state_50"]).max() <= 1e-11 * np.abs(g["state_50"]).max()
    rel = compare(rec, ref)
    step_rel = (np.abs(rec - ref) / np.abs(ref))[:45]
    assert step_rel[:, :3].max() <= 1e-13, step_rel[:, :3].max()
    assert step_rel[:, 3].max() <= 1e-12, step_rel[:, 3].max()
    assert rel[:100].max() <= 1e-9, rel[:100].max()
    er, eo = G.events(ref), G.events(rec)
    assert len(er) >= 1 and len(eo) == len(er)
    for a, b in zip(er, eo):
        assert abs(a[0] - b[0]) <= 1e-3 * a[0]
        assert abs(a[1] - b[1]) <= 1e-2 * a[1]
