"""GPU tests of the ensemble mode (64 members on one mesh, MFMA GEMM).

Every test of the GEMM and the trajectories runs with both products: FP32
inputs on the FP32 MFMA units, and the same FP32 inputs split exactly into
three BF16 parts each with the six leading products on the BF16 MFMA units
(the Python default on a planar fault).

Bars:
* GEMM (FP32 accumulation, split-K partials summed in FP64): per row, error
  <= 2e-6 of sum_j |a_ij u_j| against an FP64 product of the same FP32
  matrices;
* member trajectories: each member against the NumPy FP64 oracle with that
  member's parameters and the same FP32-rounded matrices, 30 global steps:
  step sizes within 1e-6 and states within 1e-6 of their scale;
* isolation: a member forced to fail (NaN injected from the host) is masked
  and every other member's state is bit-identical to a run without it.
"""
import numpy as np
import pytest

from oracle import qdsim_oracle as Q

pytestmark = pytest.mark.gpu

MU = LAM = 32038120320.0
CS = 3464.0
V0, VPL = 1e-6, 1e-9
B = 64


@pytest.fixture(scope="module")
def mesh():
    from pyquake3d_amd import _lib, greenfunc, readmsh, synth
    from pyquake3d_amd.QDsim_gpu import DenseSim
    nodes_km, tri = synth.planar_fault_mesh(20, 10, 40.0, 20.0)
    nodes = nodes_km * 1e3
    ev, xg = readmsh.get_eleVec(nodes, tri, False)
    sim = DenseSim(len(tri), _lib.MAT_F32)
    sim.assemble(*greenfunc.source_vertices(nodes, tri), xg, ev, MU, LAM, True)
    mats = [sim.download_matrix(k) for k in range(6)]  # FP32 values in float64
    cols = np.stack([synth.member_params(xg, m) for m in range(B)])
    return sim, mats, cols, xg


PRECISIONS = ("fp32", "bf16x3")


def _ens(mesh, precision="bf16x3"):
    from pyquake3d_amd.ensemble import Ensemble, init_members
    sim, mats, cols, xg = mesh
    e = Ensemble(sim.n, B, matrices_from=sim, kc=128, precision=precision)
    assert e.precision == (precision, precision == "bf16x3")  # planar mesh: Bs = Bd = 0
    init_members(e, cols, V0, VPL, MU, CS)
    return e


@pytest.mark.parametrize("precision", PRECISIONS)
def test_gemm_matches_fp64_product(mesh, precision):
    sim, mats, cols, xg = mesh
    e = _ens(mesh, precision)
    rng = np.random.default_rng(4)
    n = sim.n
    V1 = rng.lognormal(-18, 3, (B, n)) * rng.choice([-1, 1], (B, n))
    V2 = rng.lognormal(-18, 3, (B, n))
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
        assert err.max() < 2e-6, (q, err.max())
    e.close()


@pytest.mark.parametrize("precision", PRECISIONS)
def test_members_follow_the_oracle(mesh, precision):
    sim, mats, cols, xg = mesh
    e = _ens(mesh, precision)
    K = 30
    e.run(K, graph_steps=4)
    mem = e.members()
    # a member still retrying after a step's attempt rounds continues at the
    # next global step, so a member has taken between K - 2 and K steps
    assert np.all(mem["steps"] >= K - 2) and np.all(mem["status"] == 0)
    states = e.get("state")
    tt1 = e.get("Tt1o")
    for m in (0, 17, 63):
        orc = Q.from_init_columns(mats, cols[m], V0, VPL, MU, CS)
        dt = 1e-3
        for _ in range(int(mem["steps"][m])):
            h, dt = orc.step(dt)
        assert abs(mem["time"][m] - orc.time) <= 1e-6 * orc.time, (m, mem["time"][m], orc.time)
        if mem["nrjct"][m] == 0:
            assert abs(mem["dttry"][m] - dt) <= 1e-6 * dt
        assert np.abs(states[m] - orc.state).max() <= 1e-6 * np.abs(orc.state).max()
        assert np.abs(tt1[m] - orc.Tt1o).max() <= 1e-6 * np.abs(orc.Tt1o).max()
    e.close()


def test_failed_member_is_masked_and_others_unchanged(mesh):
    sim, mats, cols, xg = mesh
    a = _ens(mesh)
    b = _ens(mesh)
    st = b.get("state")
    st[5, 11] = np.nan
    b.set("state", st)
    a.run(12, graph_steps=4)
    b.run(12, graph_steps=4)
    ma, mb = a.members(), b.members()
    assert mb["status"][5] == 1 and mb["active"][5] == 0
    keep = np.arange(B) != 5
    for name in ("state", "Tt1o", "slipv1", "slip"):
        xa, xb = a.get(name), b.get(name)
        assert np.array_equal(xa[keep], xb[keep]), name
    assert np.array_equal(ma["time"][keep], mb["time"][keep])
    a.close()
    b.close()


@pytest.mark.parametrize("precision", PRECISIONS)
def test_gemm_skips_idle_member_subtiles_exactly(mesh, precision):
    """Member subtile j holds members j, j + 4, ..., j + 60; the GEMM skips the
    subtiles above the highest one with an active member.  Active members'
    results must be bit-identical to the all-active product and inactive
    members' exactly zero, whichever subtiles are live."""
    sim, mats, cols, xg = mesh
    e = _ens(mesh, precision)
    rng = np.random.default_rng(8)
    n = sim.n
    e.set("slipv1", rng.lognormal(-18, 3, (B, n)) * rng.choice([-1, 1], (B, n)))
    e.set("slipv2", rng.lognormal(-18, 3, (B, n)))
    e.gemm()
    full = e.gemm_result()
    for active in ([0, 4, 60], [0, 1], [2], [5, 10], [3, 63], list(range(64))):
        for m in range(B):
            e.reset_member(m, active=m in active)
        e.gemm()
        got = e.gemm_result()
        on = np.isin(np.arange(B), active)
        assert np.array_equal(got[:, on], full[:, on]), active
        assert not np.any(got[:, ~on]), active
    e.close()
