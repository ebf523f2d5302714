"""GPU tests of the QDsim_gpu stepping calls one at a time, against the NumPy
FP64 oracle:

* ``derivative_gpu(Tno, Tt1o, Tt2o, state)`` -> 4-tuple (QDsim_gpu.py:830-891)
  at the current state and at perturbed host vectors, with the frozen traction
  rates of the last step: relative 1e-12 of each output's magnitude;
* ``RungeKutte_solve_Dormand_Prince_gpu(h)`` -> 8-tuple (QDsim_gpu.py:1128-1262)
  and the CPU branch's ``RungeKutte_solve_Dormand_Prince(h)`` (:1266-1397) at an
  accepted and a rejected step size (a patch slipping at 1 mm/s): vectors within 1e-12 of their magnitude,
  conditions, proposed step sizes and Relerrormax1 / Relerrormax2 (the
  conditions x 1e-4) within 1e-7 relative (error measures are differences of
  two solutions, so last-bit differences reach them amplified);
* the attempt commits nothing: the device state is unchanged after it;
* ``Relerrormax1`` / ``Relerrormax2`` after ``simu_forward`` are the accepted
  attempt's, as the CPU branch of the reference driver writes them to
  state.txt (main_gpu.py:214);
* a failed step ends a graph run: the state is left as it was, the records of
  the skipped launches are NaN, and reset_ctrl keeps time and step counters.
"""
import numpy as np
import pytest

from oracle import qdsim_oracle as Q

pytestmark = pytest.mark.gpu

MU = LAM = 32038120320.0
CS = 3464.0
V0, VPL = 1e-6, 1e-9


@pytest.fixture(scope="module")
def case():
    from pyquake3d_amd import greenfunc, readmsh, synth
    nodes_km, tri = synth.planar_fault_mesh(20, 10, 40.0, 20.0)
    nodes = nodes_km * 1e3
    ev, xg = readmsh.get_eleVec(nodes, tri, False)
    cols = synth.friction_params(xg, dc_sigma=0.2, seed=3, rake=0.5)
    return nodes, tri, ev, xg, cols, greenfunc.source_vertices(nodes, tri)


def _pair(case, gpu_semantics=True):
    """Device simulation and oracle on the same FP64 matrices and state."""
    from pyquake3d_amd import _lib
    from pyquake3d_amd.QDsim_gpu import DenseSim
    nodes, tri, ev, xg, cols, P = case
    s = DenseSim(len(tri), _lib.MAT_F64)
    s.assemble(*P, xg, ev, MU, LAM, True)
    mats = [s.download_matrix(k) for k in range(6)]
    orc = Q.from_init_columns(mats, cols, V0, VPL, MU, CS, gpu_semantics)
    for name in ("Tno", "Tt1o", "Tt2o", "state", "slipv1", "slipv2", "slipv", "a", "b", "dc", "f0",
                 "shear_loading", "normal_loading"):
        s.set(name, getattr(orc, name))
    s.set("rake0", orc.rake0)
    s.set_params(V0, VPL, MU / (2 * CS), False, gpu_semantics)
    return s, orc


class _Mirror:
    """The QDsim mirror's stepping calls bound to a bare DenseSim (the
    constructor needs mesh and parameter files; the methods need only
    ``sim`` and the ready flag)."""

    def __init__(self, sim):
        from pyquake3d_amd.QDsim_gpu import QDsim
        self.q = QDsim.__new__(QDsim)
        self.q.sim = sim
        self.q._gpu_ready = True


def _close(a, b, tol):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() <= tol * max(np.abs(b).max(), 1e-300)


def test_derivative_gpu_matches_oracle(case):
    s, orc = _pair(case)
    dt = 1e-3
    for _ in range(6):                      # frozen rates of a nonzero slip-rate field
        h, dt, _ = s.step(dt)
    q = _Mirror(s).q
    # the oracle's RHS with the device's frozen rates and current slip rates
    orc.AdotV1, orc.AdotV2, orc.dsigmadt = (s.get(k) for k in ("AdotV1", "AdotV2", "dsigmadt"))
    orc.slipv = s.get("slipv")
    y = [s.get(k) for k in ("Tno", "Tt1o", "Tt2o", "state")]
    rng = np.random.default_rng(2)
    for inputs in (
            (q.Tno_gpu, q.Tt1o_gpu, q.Tt2o_gpu, q.state_gpu),            # device fields
            tuple(x * (1 + 1e-3 * rng.standard_normal(len(x))) for x in y)):  # host vectors
        got = q.derivative_gpu(*inputs)
        assert len(got) == 4 and all(g.shape == (s.n,) for g in got)
        want = orc.derivative(*[np.asarray(x) for x in inputs])
        for k, (g, w) in enumerate(zip(got, want)):
            assert _close(g.get(), w, 1e-12), k
    # the CPU-branch twin returns host arrays with the same values
    cpu = q.derivative(*y)
    assert all(isinstance(c, np.ndarray) for c in cpu)
    for c, w in zip(cpu, orc.derivative(*y)):
        assert _close(c, w, 1e-12)
    s.close()


@pytest.mark.parametrize("gpu_semantics", (True, False))
def test_dormand_prince_attempt_matches_oracle(case, gpu_semantics):
    s, orc = _pair(case, gpu_semantics)
    dt = do = 1e-3
    for _ in range(8):
        h, dt, _ = s.step(dt)
        ho, do = orc.step(do)
    assert abs(dt - do) <= 1e-9 * do
    q = _Mirror(s).q
    # a patch slipping at 1 mm/s, so that a large step is rejected
    V1, V = s.get("slipv1"), s.get("slipv")
    V1[:40], V[:40] = 1e-3 * np.cos(orc.rake0[:40]), np.hypot(1e-3 * np.cos(orc.rake0[:40]), s.get("slipv2")[:40])
    s.set("slipv1", V1)
    s.set("slipv", V)
    s.matvec()                                          # frozen rates of the next step
    # the oracle attempts from the device's state and frozen rates, so the
    # comparison isolates the attempt's arithmetic
    for k in ("Tno", "Tt1o", "Tt2o", "state", "slipv1", "slipv2", "slipv", "AdotV1", "AdotV2", "dsigmadt"):
        setattr(orc, k, s.get(k))
    before = [s.get(k) for k in ("Tno", "Tt1o", "Tt2o", "state", "slipv1")]
    # a step the oracle rejects with finite error measures (too large a step
    # overflows the slip rates to NaN, which proves less)
    big = next(h for h in (1e-2, 1e-1, 1.0, 10.0, 100.0)
               if np.isfinite(orc.attempt(h)[1:3]).all() and max(orc.attempt(h)[1:3]) >= 1.0)
    conds = []
    for step in (1e-6, big):                            # accepted, then rejected
        if gpu_semantics:
            r = q.RungeKutte_solve_Dormand_Prince_gpu(step)
            vecs = [x.get() for x in r[:4]]
        else:
            r = q.RungeKutte_solve_Dormand_Prince(step)
            vecs = list(r[:4])
        assert len(r) == 8
        Y5, c1, c2, hn1, hn2 = orc.attempt(step)
        for k in range(4):
            assert _close(vecs[k], Y5[k], 1e-12), k
        # the error measures are differences of the fourth- and fifth-order
        # solutions, so last-bit differences of the two (FMA contraction)
        # reach them amplified: 1e-7 relative
        for g, w in zip(r[4:], (c1, c2, hn1, hn2)):
            assert abs(g - w) <= 1e-7 * abs(w), (g, w)
        assert abs(q.Relerrormax1 - c1 * 1e-4) <= 1e-7 * c1 * 1e-4
        assert abs(q.Relerrormax2 - c2 * 1e-4) <= 1e-7 * c2 * 1e-4
        conds.append(max(r[4], r[5]))
    assert conds[0] < 1.0 <= conds[1]                   # the small step accepted, the large one rejected
    after = [s.get(k) for k in ("Tno", "Tt1o", "Tt2o", "state", "slipv1")]
    assert all(np.array_equal(a, b) for a, b in zip(before, after))   # nothing committed
    s.close()


def test_simu_forward_sets_both_error_measures(case):
    """Relerrormax1 / 2 after a CPU-branch step are those of the accepted
    attempt (main_gpu.py:214 writes both to state.txt every step)."""
    s, orc = _pair(case, gpu_semantics=False)
    q = _Mirror(s).q
    q.V0, q.Vpl_con, q.mu, q.Cs, q.fix_Tn, q.time, q.Ifdila = V0, VPL, MU, CS, False, 0.0, False
    dt = do = 1e-3
    for _ in range(6):
        h, dt = q.simu_forward(dt)
        # the oracle's attempt of the accepted step size from the same state
        orc.AdotV1, orc.AdotV2, orc.dsigmadt = orc.frozen()
        _, c1, c2, _, _ = orc.attempt(h)
        assert abs(q.Relerrormax1 - c1 * 1e-4) <= 1e-6 * c1 * 1e-4
        assert abs(q.Relerrormax2 - c2 * 1e-4) <= 1e-6 * c2 * 1e-4
        assert q.Relerrormax1 < 1e-4 and q.Relerrormax2 < 1e-4    # accepted: both below RelTol
        ho, do = orc.step(do)
        assert abs(h - ho) <= 1e-9 * ho
        # simu_forward leaves the host attributes as the CPU branch does
        assert np.abs(q.slipv - orc.slipv).max() <= 1e-9 * np.abs(orc.slipv).max()
    s.close()


def test_failed_step_ends_a_graph_run(case):
    s, orc = _pair(case)
    h, dt, _ = s.step(1e-3)
    st0 = s.state()
    good = s.get("state")
    bad = good.copy()
    bad[3] = np.nan
    s.set("state", bad)
    before = s.get("Tt1o")
    status, rec = s.run(12, graph_steps=4, records=True)
    assert status == 1
    assert np.all(np.isnan(rec))                    # no step accepted, every row NaN
    assert np.array_equal(s.get("Tt1o"), before)    # the skipped launches left the state alone
    st = s.state()
    assert st["steps"] == st0["steps"] and st["time"] == st0["time"]
    s.reset_ctrl()                                  # clears the failure, keeps time and counters
    st = s.state()
    assert st["status"] == 0 and st["time"] == st0["time"] and st["steps"] == st0["steps"]
    s.set("state", good)
    status, rec = s.run(5, graph_steps=4, records=True)
    assert status == 0 and np.all(np.isfinite(rec)) and s.state()["steps"] == st0["steps"] + 5
    s.close()
