"""Pins of the stepping oracle (oracle/qdsim_oracle.py), CPU only."""
from fractions import Fraction as Fr

import numpy as np
import pytest

from oracle import qdsim_oracle as Q

V0, VPL, MU, CS = 1e-6, 1e-9, 32038120320.0, 3464.0


def test_dormand_prince_order_conditions():
    """The tableau (QDsim_gpu.py:1129-1159) against the published Dormand &
    Prince (1980) rationals and the order conditions of both weight rows.
    The reference negates a63 (QDsim_gpu.py:1147: -46732/5247 instead of
    +46732/5247); the oracle and the kernels reproduce that sign."""
    A = {(2, 1): Fr(1, 5), (3, 1): Fr(3, 40), (3, 2): Fr(9, 40), (4, 1): Fr(44, 45), (4, 2): Fr(-56, 15),
         (4, 3): Fr(32, 9), (5, 1): Fr(19372, 6561), (5, 2): Fr(-25360, 2187), (5, 3): Fr(64448, 6561),
         (5, 4): Fr(-212, 729), (6, 1): Fr(9017, 3168), (6, 2): Fr(-355, 33), (6, 3): Fr(46732, 5247),
         (6, 4): Fr(49, 176), (6, 5): Fr(-5103, 18656)}
    b5 = [Fr(35, 384), 0, Fr(500, 1113), Fr(125, 192), Fr(-2187, 6784), Fr(11, 84), 0]
    b4 = [Fr(5179, 57600), 0, Fr(7571, 16695), Fr(393, 640), Fr(-92097, 339200), Fr(187, 2100), Fr(1, 40)]
    got = {(2, 1): Q.B21, (3, 1): Q.B31, (3, 2): Q.B32, (4, 1): Q.B41, (4, 2): Q.B42, (4, 3): Q.B43,
           (5, 1): Q.B51, (5, 2): Q.B52, (5, 3): Q.B53, (5, 4): Q.B54, (6, 1): Q.B61, (6, 2): Q.B62,
           (6, 3): Q.B63, (6, 4): Q.B64, (6, 5): Q.B65}
    for k, v in A.items():
        assert got[k] == (float(-v) if k == (6, 3) else float(v)), k
    assert [Q.B71, 0, Q.B73, Q.B74, Q.B75, Q.B76, 0] == [float(x) for x in b5]
    assert [Q.B81, 0, Q.B83, Q.B84, Q.B85, Q.B86, Q.B87] == [float(x) for x in b4]
    # published tableau: row sums give the nodes, and the weights satisfy the
    # quadrature order conditions (5th order for b5, 4th for b4)
    c = [Fr(0)] + [sum(A[(i, j)] for j in range(1, i)) for i in range(2, 7)] + [Fr(1)]
    assert c == [0, Fr(1, 5), Fr(3, 10), Fr(4, 5), Fr(8, 9), 1, 1]
    for p in range(1, 6):
        assert sum(b5[i] * c[i] ** (p - 1) for i in range(7)) == Fr(1, p)
    for p in range(1, 5):
        assert sum(b4[i] * c[i] ** (p - 1) for i in range(7)) == Fr(1, p)
    # with the reference's a63 the sixth stage is evaluated off its node
    assert sum(A[(6, j)] for j in range(1, 6)) - 2 * A[(6, 3)] != 1


def _toy_mats(n, seed=0, diag=3e7):
    rng = np.random.default_rng(seed)
    A = [rng.normal(0, 1e5, (n, n)) for _ in range(6)]
    for k in (0, 4):
        A[k] += diag * np.eye(n)
    A[2] *= 0.1
    A[5] *= 0.1
    return A


def _cols(n, seed=0, V=VPL):
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.004, 0.04, n)
    cols = np.stack([np.zeros(n), a, np.full(n, 0.03), np.full(n, 0.14), np.full(n, 0.6), np.zeros(n),
                     np.full(n, 25.0), np.full(n, V), np.zeros(n), np.zeros(n)], 1)
    return cols


def test_steady_state_is_a_fixed_point_and_dt_grows_by_1p5():
    n = 12
    cols = _cols(n)
    cols[:, 0] = 0.3  # rake: both shear components non-zero
    sim = Q.from_init_columns(_toy_mats(n), cols, V0, VPL, MU, CS)
    # remove the 1e-16 offsets so that u = V - Vpl vanishes exactly
    sim.slipv1 = VPL * np.cos(sim.rake0)
    sim.slipv2 = VPL * np.sin(sim.rake0)
    y0 = (sim.Tno.copy(), sim.Tt1o.copy(), sim.state.copy())
    dt = 1e-3
    for k in range(6):
        h, dtn = sim.step(dt)
        assert h == dt and dtn == 1.5 * h   # dt_k = 1e-3 * 1.5**k (reference notebook traces)
        dt = dtn
    assert np.allclose(sim.Tno, y0[0], rtol=1e-14) and np.allclose(sim.Tt1o, y0[1], rtol=1e-12)
    assert np.allclose(sim.state, y0[2], rtol=1e-12)


def test_rhs_satisfies_quasidynamic_balance():
    """dtau/dt + eta dV/dt = -K u + loading, with dV/dt by central differences
    of V(tau, sigma, psi) = 2 V0 exp(-psi/a) sinh(tau/(a sigma)) -- pins the
    partial derivatives of the RHS (QDsim_gpu.py:813-826)."""
    n = 8
    rng = np.random.default_rng(1)
    sim = Q.from_init_columns(_toy_mats(n), _cols(n, V=1e-6), V0, VPL, MU, CS)
    sim.AdotV1, sim.AdotV2, sim.dsigmadt = rng.normal(0, 1e3, n), rng.normal(0, 1e3, n), rng.normal(0, 1e3, n)
    Tno, T1, T2, st = sim.Tno * 1.01, sim.Tt1o * 1.02, sim.Tt1o * 0.3 + 0.1, sim.state * 0.999
    dst, dsn, dt1, dt2 = sim.derivative(Tno, T1, T2, st)
    eta = MU / (2 * CS)

    def V(tau, sig, psi):
        return 2 * V0 * np.exp(-psi / sim.a) * np.sinh(tau / (sim.a * sig))

    for tau, dtau, AV in ((T1, dt1, sim.AdotV1), (T2, dt2, sim.AdotV2)):
        eps = 1e-6
        rates = (dtau * 1e6, dsn * 1e6, dst)
        base = (tau * 1e6, Tno * 1e6, st)
        dV = 0.0
        for k in range(3):
            hi = [b.copy() for b in base]
            lo = [b.copy() for b in base]
            step = eps * np.maximum(np.abs(base[k]), 1e-3)
            hi[k] += step
            lo[k] -= step
            dV = dV + (V(*hi) - V(*lo)) / (2 * step) * rates[k]
        lhs = dtau * 1e6 + eta * dV
        assert np.allclose(lhs, -AV + sim.shear_loading, rtol=1e-5, atol=1e-6 * np.abs(AV).max())


def test_reject_path_and_python_nan_semantics():
    assert Q.py_max(1.0, np.nan) == 1.0 and np.isnan(Q.py_max(np.nan, 1.0))
    assert Q.py_min(1.0, np.nan) == 1.0 and np.isnan(Q.py_min(np.nan, 1.0))
    n = 10
    sim = Q.from_init_columns(_toy_mats(n), _cols(n, V=1e-3), V0, VPL, MU, CS)
    h, dtn = sim.step(1e4)   # far too large: rejected and shrunk
    assert sim.nrjct >= 1 and h < 1e4
    sim2 = Q.from_init_columns(_toy_mats(n), _cols(n, V=1e-3), V0, VPL, MU, CS)
    sim2.state[3] = np.nan
    with pytest.raises(RuntimeError, match="dt is too small"):
        sim2.step(1.0)


def test_gpu_error_estimate_drops_normal_traction():
    n = 6
    mats = _toy_mats(n)
    cols = _cols(n, V=1e-4)
    g = Q.from_init_columns(mats, cols, V0, VPL, MU, CS, gpu_semantics=True)
    c = Q.from_init_columns(mats, cols, V0, VPL, MU, CS, gpu_semantics=False)
    g.AdotV1, g.AdotV2, g.dsigmadt = g.frozen()
    c.AdotV1, c.AdotV2, c.dsigmadt = c.frozen()
    _, _, c2g, _, _ = g.attempt(10.0)
    _, _, c2c, _, _ = c.attempt(10.0)
    assert c2c >= c2g
