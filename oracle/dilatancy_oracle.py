"""CPU restatement of the reference's dilatancy / pore-pressure update (test
infrastructure only: imported by tests/ and nothing else).

Reference: src/QDsim_gpu.py
* grid (:336-349): yp = logspace(-3, 10, 60), zp = log(1 + yp / c), c = 1e-2;
  pore pressures in Pa, P0 (the fixed far boundary) and the initial profile
  from the parameter file in MPa times 1e6;
* Calc_Pmatrix (:611-624): the diffusion operator on the stretched grid, one
  row per node but the last; row 0 couples nodes 0 and 1, row i couples
  i - 1, i, i + 1 with weights exp(-(2 z_i -+ dz_i / 2)) / dz_i^2;
* Calc_P_implicit (:628-671), run after every accepted step with the step's
  h: theta = Dc / V0 exp((psi - f0) / b), dtheta/dt = 1 - V theta / Dc,
  g = -e hs / (2 beta c_hyd theta) dtheta/dt; one backward-Euler step
  (I + dt P) x = b of the first 59 nodes, with the flux source at node 0 and
  the fixed-pressure coupling at node 58; then P = x_0 and dP/dt from the
  flux expression at node 0.

Pins (tests/test_oracle_dilatancy.py): golden vectors of the reference's own
functions (tests/golden/dilatancy_golden.npz); a uniform profile at P0 with no
source stays at P0 (the discrete operator annihilates constants); the solve
against an independent tridiagonal elimination.
"""
import numpy as np


def pore_grid(num=60, c=1e-2):
    yp = np.logspace(start=-3, stop=10, num=num, base=10)
    return np.log(1 + yp / c)


def pmatrix(zp, c_hyd):
    """Dense num x num operator with the reference's entries (last row zero)."""
    K = len(zp)
    dz = np.diff(zp)
    P = np.zeros((K, K))
    for i in range(K - 1):
        lo = np.exp(-(2 * zp[i] - dz[i] / 2))
        hi = np.exp(-(2 * zp[i] + dz[i] / 2))
        P[i, i] = c_hyd * (lo + hi) / (dz[i] * dz[i])
        if i == 0:
            P[0, 1] = -c_hyd * (lo + hi) / (dz[0] * dz[0])
        else:
            P[i, i + 1] = -c_hyd * hi / (dz[i] * dz[i])
            P[i, i - 1] = -c_hyd * lo / (dz[i] * dz[i])
    return P


def source_g(state, slipv, dc, f0, b, V0, e, hs, beta, c_hyd):
    theta = dc / V0 * np.exp((state - f0) / b)
    dthetadt = 1.0 - slipv * theta / dc
    return -e * hs / (2.0 * beta * c_hyd * theta) * dthetadt


def implicit_step(Parr, g, zp, Pm, c_hyd, P0, dt):
    """One backward-Euler update of every element's profile.  Parr: (N, K)
    in Pa (last column the fixed boundary).  Returns (Parr_new, P, dPdt0)."""
    K = len(zp)
    dz = np.diff(zp)
    M = Pm[:-1, :-1] * dt + np.eye(K - 1)
    Bm = np.array(Parr[:, :-1], dtype=np.float64).T.copy()      # (K-1, N)
    Bm[0] -= 2.0 * c_hyd * np.exp(-(zp[0] - dz[0] / 2)) * g * dt / dz[0]
    Bm[-1] += dt / (dz[-1] * dz[-1]) * c_hyd * np.exp(-(zp[-2] - dz[-1] / 2)) * P0
    X = np.linalg.solve(M, Bm)
    out = np.array(Parr, dtype=np.float64, copy=True)
    out[:, :-1] = X.T
    term1 = -np.exp(-(zp[0] - dz[0] / 2)) * (out[:, 0] - out[:, 1] + 2 * dz[0] * g * np.exp(zp[0]))
    term2 = np.exp(-(zp[0] + dz[0] / 2)) * (out[:, 1] - out[:, 0])
    dPdt0 = c_hyd * np.exp(-zp[0]) * (term1 + term2) / dz[0] ** 2
    return out, out[:, 0].copy(), dPdt0


def thomas(lower, diag, upper, rhs):
    """Tridiagonal solve by elimination without pivoting (independent check of
    the dense solve for diagonally dominant systems)."""
    n = len(diag)
    c = np.zeros(n)
    d = np.zeros(n)
    c[0] = upper[0] / diag[0]
    d[0] = rhs[0] / diag[0]
    for i in range(1, n):
        den = diag[i] - lower[i] * c[i - 1]
        c[i] = upper[i] / den if i < n - 1 else 0.0
        d[i] = (rhs[i] - lower[i] * d[i - 1]) / den
    x = np.zeros(n)
    x[-1] = d[-1]
    for i in range(n - 2, -1, -1):
        x[i] = d[i] - c[i] * x[i + 1]
    return x
