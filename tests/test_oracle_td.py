"""Pins of the TD stress oracle (CPU only).

The oracle (oracle/td_oracle.py) is checked against
* the golden vectors produced by the reference's own C++ (tests/golden),
* the reference library built from source (oracle/_ref), when present,
* physics that fixes the answer independently of any implementation:
  traction-free free surface, static equilibrium, the Burgers-vector jump of
  displacement across the TD, and invariance under rotations about the
  vertical axis.
"""
import numpy as np
import pytest

from oracle import refbind, td_oracle as T

MU = 32038120320.0
LAM = 32038120320.0
NU = 0.25


def _rel(a, b):
    """Per-pair error relative to the pair's largest stress component."""
    sc = np.abs(b).max(axis=-1)
    return np.abs(a - b).max(axis=-1) / np.where(sc > 0, sc, 1.0)


@pytest.mark.parametrize("case,hs", [("random_fs", False), ("random_hs", True), ("surface", True)])
def test_oracle_matches_reference_golden(golden_td, case, hs):
    g = golden_td
    R = g[f"{case}_R"]
    P1, P2, P3 = g[f"{case}_P1"], g[f"{case}_P2"], g[f"{case}_P3"]
    for k, b in enumerate(g["burgers"]):
        s, _ = T.td_stress(R, P1, P2, P3, b[1], b[2], b[0], MU, LAM, hs)
        s = np.asarray(s, np.float64)
        e = _rel(s, g[f"{case}_ref_stress"][k])
        # the reference is double precision: cancellation-limited pairs reach
        # ~1e-8; a dropped term or wrong sign would show at O(1)
        assert np.quantile(e, 0.99) < 1e-10, (case, k, np.quantile(e, 0.99))
        assert e.max() < 1e-6, (case, k, e.max())
        assert np.abs(s - g[f"{case}_oracle_stress"][k]).max() <= 1e-12 * np.abs(s).max()


def test_oracle_main_demo_case(golden_td):
    g = golden_td
    R, P = g["main_R"], g["main_P"]
    s, _ = T.td_stress(R, np.tile(P[0], (3, 1)), np.tile(P[1], (3, 1)), np.tile(P[2], (3, 1)),
                       1.0, 0.5, 0.0, MU, LAM, True)
    assert _rel(np.asarray(s, np.float64), g["main_ref_stress"]).max() < 1e-10


def test_oracle_planar_gf_matches_reference(golden_td):
    g = golden_td
    mats = np.stack(T.traction_gf(g["planar_nodes"], g["planar_tri"], MU, LAM, True))
    ref = g["planar_gf_ref"]
    col = np.abs(ref).max(axis=(0, 1))
    assert (np.abs(mats - ref) / col).max() < 1e-12


@pytest.mark.skipif(not refbind.available(), reason="oracle/_ref not built")
def test_oracle_vs_reference_library_fresh_pairs():
    rng = np.random.default_rng(7)
    m = 150
    C = np.stack([rng.uniform(-3e3, 3e3, m), rng.uniform(-3e3, 3e3, m), rng.uniform(-15e3, -2e3, m)], 1)
    P = [C + rng.normal(0, 700, (m, 3)) for _ in range(3)]
    for p in P:
        p[:, 2] = np.minimum(p[:, 2], -5.0)
    R = C + rng.normal(0, 3e3, (m, 3))
    R[:, 2] = np.minimum(R[:, 2], -1.0)
    for hs in (False, True):
        s, _ = T.td_stress(R, *P, 0.7, -0.4, 0.2, MU, LAM, hs)
        r, _ = refbind.td_stress_pairs(R, *P, 0.7, -0.4, 0.2, MU, LAM, hs)
        e = _rel(np.asarray(s, np.float64), r)
        assert np.quantile(e, 0.99) < 1e-10 and e.max() < 1e-6


def _tri_at_depth(rng, m, depth=(-12e3, -3e3)):
    C = np.stack([rng.uniform(-2e3, 2e3, m), rng.uniform(-2e3, 2e3, m), rng.uniform(*depth, m)], 1)
    return [C + rng.normal(0, 800, (m, 3)) for _ in range(3)]


def test_free_surface_is_traction_free():
    rng = np.random.default_rng(11)
    m = 60
    P1, P2, P3 = _tri_at_depth(rng, m)
    R = np.stack([rng.uniform(-8e3, 8e3, m), rng.uniform(-8e3, 8e3, m), np.zeros(m)], 1)
    s, _ = T.td_stress(R, P1, P2, P3, 1.0, 0.6, 0.3, MU, LAM, True)
    s = np.asarray(s, np.float64)
    scale = np.abs(s).max(axis=1)
    # sigma_xz, sigma_yz, sigma_zz vanish on z = 0
    assert (np.abs(s[:, [2, 4, 5]]).max(axis=1) / scale).max() < 1e-9
    # ... and the full-space solution alone does not (the check has teeth)
    sf, _ = T.td_stress(R, P1, P2, P3, 1.0, 0.6, 0.3, MU, LAM, False)
    sf = np.asarray(sf, np.float64)
    assert (np.abs(sf[:, [2, 4, 5]]).max(axis=1) / np.abs(sf).max(axis=1)).max() > 1e-2


@pytest.mark.parametrize("hs", [False, True])
def test_static_equilibrium(hs):
    """div(sigma) = 0 away from the dislocation (central differences, h = 1 m)."""
    rng = np.random.default_rng(5)
    m = 20
    P1, P2, P3 = _tri_at_depth(rng, m)
    R0 = np.stack([rng.uniform(-6e3, 6e3, m), rng.uniform(-6e3, 6e3, m), rng.uniform(-9e3, -1e3, m)], 1)
    h = 1.0
    b = (0.0, 1.0, 0.5)

    def sig(R):
        s, _ = T.td_stress(R, P1, P2, P3, b[1], b[2], b[0], MU, LAM, hs)
        S = np.zeros((m, 3, 3), dtype=T.LD)
        idx = {(0, 0): 0, (1, 1): 1, (2, 2): 2, (0, 1): 3, (0, 2): 4, (1, 2): 5}
        for (i, j), k in idx.items():
            S[:, i, j] = S[:, j, i] = s[:, k]
        return S

    div = np.zeros((m, 3), dtype=T.LD)
    for j in range(3):
        d = np.zeros(3)
        d[j] = h
        div += (sig(R0 + d)[:, :, j] - sig(R0 - d)[:, :, j]) / (2 * h)
    dist = np.min([np.linalg.norm(R0 - p, axis=1) for p in (P1, P2, P3)], axis=0)
    scale = np.abs(sig(R0)).max(axis=(1, 2)) / dist
    assert (np.abs(div).max(axis=1) / scale).max() < 1e-5


def test_burgers_jump_across_td():
    """u(x+) - u(x-) = b across the TD, x+ on the side of the TDCS normal."""
    rng = np.random.default_rng(3)
    m = 25
    P1, P2, P3 = _tri_at_depth(rng, m, depth=(-20e3, -10e3))
    w = rng.dirichlet([2, 2, 2], m)
    C = w[:, :1] * P1 + w[:, 1:2] * P2 + w[:, 2:] * P3
    n = np.cross(P2 - P1, P3 - P1)
    n /= np.linalg.norm(n, axis=1)[:, None]
    b = np.array([0.3, 1.0, -0.5])  # (Ts, Ss, Ds)
    eps = 1e-6
    for hs in (False, True):
        up = T.td_displacement(C + eps * n, P1, P2, P3, b, NU, hs)
        um = T.td_displacement(C - eps * n, P1, P2, P3, b, NU, hs)
        Vn, Vs, Vd = T._td_frame(P1.astype(T.LD), P2.astype(T.LD), P3.astype(T.LD))
        bE = b[0] * Vn + b[1] * Vs + b[2] * Vd
        jump = np.asarray(up - um, np.float64)
        assert np.abs(jump - np.asarray(bE, np.float64)).max() < 1e-7


def test_rotation_about_vertical_axis():
    """Rotating sources and receivers about z rotates the stress tensor."""
    rng = np.random.default_rng(9)
    m = 40
    P1, P2, P3 = _tri_at_depth(rng, m)
    R = np.stack([rng.uniform(-6e3, 6e3, m), rng.uniform(-6e3, 6e3, m), rng.uniform(-9e3, -100, m)], 1)
    th = 0.7
    Q = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    for hs in (False, True):
        s0, _ = T.td_stress(R, P1, P2, P3, 1.0, 0.4, 0.0, MU, LAM, hs)
        s1, _ = T.td_stress(R @ Q.T, P1 @ Q.T, P2 @ Q.T, P3 @ Q.T, 1.0, 0.4, 0.0, MU, LAM, hs)

        def full(s):
            S = np.zeros((m, 3, 3))
            for (i, j), k in {(0, 0): 0, (1, 1): 1, (2, 2): 2, (0, 1): 3, (0, 2): 4, (1, 2): 5}.items():
                S[:, i, j] = S[:, j, i] = np.asarray(s[:, k], np.float64)
            return S

        rot = np.einsum("ij,mjk,lk->mil", Q, full(s0), Q)
        assert (np.abs(rot - full(s1)).max(axis=(1, 2)) / np.abs(rot).max(axis=(1, 2))).max() < 1e-9
