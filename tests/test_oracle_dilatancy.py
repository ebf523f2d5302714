"""Pins of the dilatancy / pore-pressure oracle (oracle/dilatancy_oracle.py):

* golden vectors of the reference's own Calc_Pmatrix and Calc_P_implicit
  (tests/golden/dilatancy_golden.npz, written by the script beside it): the
  operator to 1e-15 and two consecutive updates of 40 elements at dt = 1e-3,
  1e5 and 3e7 s to 1e-12 (profiles, P, dP/dt);
* a uniform profile at the boundary pressure with no dilatancy source moves
  only by the solution driven by the reference's far-boundary coefficient
  mismatch (the operator itself annihilates constants);
* the dense solve equals an independent tridiagonal elimination.
"""
import os

import numpy as np
import pytest

from oracle import dilatancy_oracle as D

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dilatancy_golden.npz")


@pytest.fixture(scope="module")
def g():
    return dict(np.load(GOLDEN, allow_pickle=False))


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_grid_and_operator_match_reference(g):
    zp = D.pore_grid()
    assert np.array_equal(zp, g["zp"])
    Pm = D.pmatrix(zp, float(g["Chyd"]))
    assert _rel(Pm, g["Pmatrix"]) < 1e-15


def test_updates_match_reference(g):
    zp, Pm = g["zp"], g["Pmatrix"]
    for c in range(3):
        kw = {k: g[f"c{c}_{k}"] for k in ("state", "slipv", "dc", "f0", "b")}
        gg = D.source_g(kw["state"], kw["slipv"], kw["dc"], kw["f0"], kw["b"], float(g["V0"]),
                        float(g["DilatancyC"]), float(g["hs"]), float(g["EPermeability"]), float(g["Chyd"]))
        for it in range(2):
            Parr, P, dP = D.implicit_step(g[f"c{c}_Parr_in{it}"], gg, zp, Pm, float(g["Chyd"]), float(g["P0"]),
                                          float(g[f"c{c}_dt"]))
            assert _rel(Parr, g[f"c{c}_Parr_out{it}"]) < 1e-12, (c, it)
            assert _rel(P, g[f"c{c}_P_out{it}"]) < 1e-12, (c, it)
            assert np.abs(dP - g[f"c{c}_dPdt0_out{it}"]).max() <= 1e-10 * np.abs(g[f"c{c}_dPdt0_out{it}"]).max()


def test_uniform_boundary_pressure_and_the_boundary_coefficient():
    """A uniform profile at P0 with no source: the operator annihilates
    constants, so the only change comes from the far-boundary source term,
    whose coefficient exp(-(z[-2] - dz/2)) / dz^2 differs from the operator's
    own coupling exp(-(2 z[-2] + dz/2)) / dz^2 (Calc_P_implicit :654 vs
    Calc_Pmatrix :621).  The update equals the solution driven by exactly
    that mismatch (a reference quirk the oracle keeps)."""
    zp = D.pore_grid()
    c_hyd = 1e-5
    Pm = D.pmatrix(zp, c_hyd)
    dz = np.diff(zp)
    P0 = 3.5e6
    Parr = np.full((5, len(zp)), P0)
    for dt in (1e-3, 1e4, 1e9):
        out, P, dP = D.implicit_step(Parr, np.zeros(5), zp, Pm, c_hyd, P0, dt)
        M = Pm[:-1, :-1] * dt + np.eye(len(zp) - 1)
        q_src = c_hyd * np.exp(-(zp[-2] - dz[-1] / 2)) / dz[-1] ** 2
        q_op = c_hyd * np.exp(-(2 * zp[-2] + dz[-1] / 2)) / dz[-1] ** 2
        rhs = np.zeros(len(zp) - 1)
        rhs[-1] = dt * (q_src - q_op) * P0
        delta = np.linalg.solve(M, rhs)
        assert np.abs(out[:, :-1] - P0 - delta).max() <= 1e-12 * P0


def test_dense_solve_equals_tridiagonal_elimination():
    zp = D.pore_grid()
    Pm = D.pmatrix(zp, 1e-5)
    rng = np.random.default_rng(0)
    for dt in (1e-2, 1e6):
        M = Pm[:-1, :-1] * dt + np.eye(len(zp) - 1)
        rhs = rng.uniform(1e6, 3e6, len(zp) - 1)
        lower = np.r_[0.0, np.diag(M, -1)]
        upper = np.r_[np.diag(M, 1), 0.0]
        x = D.thomas(lower, np.diag(M).copy(), upper, rhs)
        assert _rel(x, np.linalg.solve(M, rhs)) < 1e-12
