"""GPU parity of the batched ensemble path through seismic events and on a
non-planar fault.

* Event times: the 1,600-element case of the reference's own 2,000-step run
  (tests/golden/qdsim_trajectory.npz, QDsim_gpu.simu_forward) loaded into four
  slots (0, 17, 42, 63: one per GEMM member subtile) of a 64-slot ensemble
  with FP32 Green's functions, stepped to 2,000 accepted steps with the
  BF16-split GEMM and with the FP32 MFMA GEMM.  Bars: the same seismic events
  as the reference run (max slip rate above 1e-3 m/s until it falls below
  1e-4), onset within 1e-3 relative and peak slip rate within 1 %; the four
  slots bit-identical to each other.  The Green's functions are
  this build's FP64 assembly rounded to FP32 (the reference stores FP64): the
  bar is the reference's event times, not its bits.
* Rough fault (Bs, Bd nonzero): the GEMM's three components against the FP64
  product of the same FP32 matrices (per row, 2e-6 of sum |a_ij u_j|) in both
  arithmetics -- the BF16-split GEMM then runs its normal pass -- and three
  members' trajectories against the NumPy FP64 oracle over 30 global steps
  from slip rates off the plate rate: time, state and shear traction within
  1e-6, the normal traction's change and the frozen dsigma/dt within 1e-3 of
  their largest magnitude (FP32 GEMM sums against FP64 ones).
* Fix_Tn (QDsim_gpu.py:950-953) on the rough fault: dsigma/dt is the normal
  loading alone, the normal traction stays put, and the members follow the
  oracle run with fix_Tn.
"""
import os
import sys

import numpy as np
import pytest

from oracle import qdsim_oracle as Q

pytestmark = pytest.mark.gpu

MU = LAM = 32038120320.0
CS = 3464.0
V0, VPL = 1e-6, 1e-9
B = 64
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PRECISIONS = ("bf16x3", "fp32")


@pytest.fixture(scope="module")
def golden_case():
    from pyquake3d_amd import _lib, greenfunc, readmsh
    from pyquake3d_amd.QDsim_gpu import DenseSim
    sys.path.insert(0, GOLDEN)
    import make_qdsim_trajectory as G
    g = dict(np.load(os.path.join(GOLDEN, "qdsim_trajectory.npz"), allow_pickle=False))
    nodes_km, nodes_m, tri, cols = G.case_inputs()
    ev, xg = readmsh.get_eleVec(nodes_m, tri, False)
    sim = DenseSim(len(tri), _lib.MAT_F32)
    sim.assemble(*greenfunc.source_vertices(nodes_m, tri), xg, ev, MU, LAM, True)
    yield sim, cols, g, G
    sim.close()


def events(rec, on=1e-3, off=1e-4):
    """(onset time, peak slip rate) of each excursion of the max slip rate
    (column 3 of per-step records time, h, dtnext, max V) that starts above
    `on` and ends when the rate falls below `off` (the hysteresis keeps a
    step that dips just under `on` during nucleation inside one event)."""
    t, v = rec[:, 0], rec[:, 3]
    out, i = [], 0
    while i < len(v):
        if v[i] > on and (i == 0 or v[i - 1] <= on):
            j = i
            while j < len(v) and v[j] >= off:
                j += 1
            out.append((t[i], v[i:j].max()))
            i = j
        else:
            i += 1
    return out


@pytest.mark.parametrize("precision", PRECISIONS)
def test_ensemble_slots_follow_reference_events(golden_case, precision):
    from pyquake3d_amd.ensemble import Ensemble, load_member
    sim, cols, g, G = golden_case
    ref = g["records"]
    slots = (0, 17, 42, 63)
    e = Ensemble(sim.n, B, matrices_from=sim, precision=precision)
    e.set_params(V0, VPL, MU / (2 * CS))
    for m in range(B):
        if m in slots:
            load_member(e, m, cols, V0, VPL)
        else:
            e.reset_member(m, active=False)
    rec = []
    last = 0
    gsteps = 0
    while len(rec) < len(ref):
        e.run(1, graph_steps=1)
        gsteps += 1
        mem = e.members()
        assert mem["status"][0] == 0
        if mem["steps"][0] > last:
            last = int(mem["steps"][0])
            rec.append((mem["time"][0], mem["h"][0], mem["dtnext"][0], e.get_member(0, "slipv").max()))
        assert gsteps < 3 * len(ref), "members stopped accepting steps"
    rec = np.array(rec)
    mem = e.members()
    for m in slots[1:]:   # one member per GEMM subtile: identical columns, identical bits
        assert mem["steps"][m] == mem["steps"][0] and mem["time"][m] == mem["time"][0]
        for name in ("state", "Tt1o", "slip"):
            assert np.array_equal(e.get_member(m, name), e.get_member(0, name)), (m, name)
    er, ed = events(ref), events(rec)
    print(f"{precision}: {gsteps} global steps for {len(rec)} accepted; reference events {len(er)}, "
          f"ensemble events {len(ed)}", flush=True)
    assert len(er) >= 1 and len(ed) == len(er)
    for a, b in zip(er, ed):
        print(f"  onset {a[0]:.6e} / {b[0]:.6e} s (rel {abs(a[0] - b[0]) / a[0]:.2e}), "
              f"peak V {a[1]:.4f} / {b[1]:.4f} m/s", flush=True)
        assert abs(a[0] - b[0]) <= 1e-3 * a[0]
        assert abs(a[1] - b[1]) <= 1e-2 * a[1]
    e.close()


@pytest.fixture(scope="module")
def rough():
    from pyquake3d_amd import _lib, greenfunc, readmsh, synth
    from pyquake3d_amd.QDsim_gpu import DenseSim
    nodes_km, tri = synth.rough_fault_mesh(40, 12, length_km=40.0, depth_km=20.0, amp_km=1.0,
                                           wavelength_km=20.0, seed=5)
    nodes = nodes_km * 1e3
    ev, xg = readmsh.get_eleVec(nodes, tri, False)
    sim = DenseSim(len(tri), _lib.MAT_F32)
    sim.assemble(*greenfunc.source_vertices(nodes, tri), xg, ev, MU, LAM, True)
    mats = [sim.download_matrix(k) for k in range(6)]
    assert np.abs(mats[2]).max() > 0 and np.abs(mats[5]).max() > 0   # a non-planar fault
    cols = np.stack([synth.member_params(xg, m) for m in range(B)])
    # slip rates off the plate rate (each element at steady state at its own
    # rate), so the traction rates, the normal one included, are not rounding noise
    cols[:, :, 7] = VPL * np.exp(0.5 * np.random.default_rng(9).standard_normal(cols.shape[:2]))
    yield sim, mats, cols
    sim.close()


def _rough_ens(rough, precision, fix_tn=False):
    from pyquake3d_amd.ensemble import Ensemble, init_members
    sim, mats, cols = rough
    e = Ensemble(sim.n, B, matrices_from=sim, kc=128, precision=precision)
    init_members(e, cols, V0, VPL, MU, CS, fix_tn=fix_tn)
    assert e.structure[1] is False                      # Bs, Bd streamed
    assert e.precision == (precision, precision == "bf16x3")
    return e


@pytest.mark.parametrize("precision", PRECISIONS)
def test_rough_fault_gemm_three_components(rough, precision):
    sim, mats, cols = rough
    e = _rough_ens(rough, precision)
    rng = np.random.default_rng(6)
    n = sim.n
    V1 = rng.lognormal(-18, 3, (B, n)) * rng.choice([-1, 1], (B, n))
    V2 = rng.lognormal(-18, 3, (B, n)) * rng.choice([-1, 1], (B, n))
    e.set("slipv1", V1)
    e.set("slipv2", V2)
    e.gemm()
    got = e.gemm_result()
    r0 = cols[:, :, 0]
    U1 = (V1 - VPL * np.cos(r0)).astype(np.float32).astype(np.float64)
    U2 = (V2 - VPL * np.sin(r0)).astype(np.float32).astype(np.float64)
    for q in range(3):
        want = U1 @ mats[q].T + U2 @ mats[q + 3].T
        mag = np.abs(U1) @ np.abs(mats[q]).T + np.abs(U2) @ np.abs(mats[q + 3]).T
        err = np.abs(got[q] - want) / np.where(mag > 0, mag, 1.0)
        print(f"{precision} component {q}: max error {err.max():.2e} of sum |a u|", flush=True)
        assert err.max() < 2e-6, (q, err.max())
        assert np.abs(want).max() > 0
    e.close()


def _oracle_member(mats, cols, fix_tn=False):
    orc = Q.from_init_columns(mats, cols, V0, VPL, MU, CS)
    orc.fix_Tn = fix_tn
    return orc


@pytest.mark.parametrize("precision", PRECISIONS)
def test_rough_fault_members_follow_oracle(rough, precision):
    sim, mats, cols = rough
    e = _rough_ens(rough, precision)
    K = 30
    Tno0 = e.get("Tno")
    e.run(K, graph_steps=4)
    mem = e.members()
    assert np.all(mem["steps"] >= K - 2) and np.all(mem["status"] == 0)
    got = {k: e.get(k) for k in ("state", "Tt1o", "Tno", "dsigmadt")}
    for m in (0, 17, 63):
        orc = _oracle_member(mats, cols[m])
        dt = 1e-3
        for _ in range(int(mem["steps"][m])):
            _, dt = orc.step(dt)
        assert abs(mem["time"][m] - orc.time) <= 1e-6 * orc.time
        assert np.abs(got["state"][m] - orc.state).max() <= 1e-6 * np.abs(orc.state).max()
        assert np.abs(got["Tt1o"][m] - orc.Tt1o).max() <= 1e-6 * np.abs(orc.Tt1o).max()
        # the normal traction moves on a rough fault: its change follows the oracle's
        dT = got["Tno"][m] - Tno0[m]
        dTo = orc.Tno - Tno0[m]
        assert np.abs(dTo).max() > 0
        rel = np.abs(dT - dTo).max() / np.abs(dTo).max()
        print(f"{precision} member {m}: normal traction change within {rel:.2e} of the oracle's", flush=True)
        assert rel <= 1e-3
        if mem["nrjct"][m] == 0:   # accepted in the last global step: same frozen rates as the oracle's last step
            assert np.abs(got["dsigmadt"][m] - orc.dsigmadt).max() <= 1e-3 * np.abs(orc.dsigmadt).max()
    e.close()


@pytest.mark.parametrize("precision", PRECISIONS)
def test_fix_tn_keeps_normal_traction(rough, precision):
    sim, mats, cols = rough
    e = _rough_ens(rough, precision, fix_tn=True)
    K = 20
    Tno0 = e.get("Tno")
    e.run(K, graph_steps=4)
    mem = e.members()
    assert np.all(mem["status"] == 0)
    assert np.array_equal(e.get("Tno"), Tno0)                  # dsigma/dt = normal loading = 0
    assert not np.any(e.get("dsigmadt"))
    st, tt = e.get("state"), e.get("Tt1o")
    for m in (0, 33):
        orc = _oracle_member(mats, cols[m], fix_tn=True)
        dt = 1e-3
        for _ in range(int(mem["steps"][m])):
            _, dt = orc.step(dt)
        assert abs(mem["time"][m] - orc.time) <= 1e-6 * orc.time
        assert np.abs(st[m] - orc.state).max() <= 1e-6 * np.abs(orc.state).max()
        assert np.abs(tt[m] - orc.Tt1o).max() <= 1e-6 * np.abs(orc.Tt1o).max()
    e.close()
