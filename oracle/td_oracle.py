"""Long-double CPU oracle for triangular-dislocation (TD) strain and stress.

TEST INFRASTRUCTURE ONLY -- never imported by the product package.

What it restates
----------------
The TD displacement field of Nikkhoo & Walter (2015, GJI 201:1119) exactly as
the reference holds it in ``src/DH_greenfunction.py``:

* angular-dislocation displacement  ``AngDisDisp``      DH_greenfunction.py:71-104
* ADCS set-up / rotation             ``TDSetupD``        DH_greenfunction.py:40-69
* full-space TD                      ``TDdispFS``        DH_greenfunction.py:106-208
* free-surface correction            ``TDdisp_HarFunc``  DH_greenfunction.py:211-231
                                     ``AngSetupFSC``     DH_greenfunction.py:234-271
                                     ``AngDisDispFSC``   DH_greenfunction.py:273-351
* half space = main + image + FSC    ``TDdispHS``        DH_greenfunction.py:353-383

The field is evaluated on forward-mode dual numbers seeded with the receiver's
Cartesian coordinates, so every evaluation also carries the exact displacement
gradient ``G[i][j] = du_i/dX_j``.  Strain is ``(G + G^T)/2`` and stress follows
Hooke's law, as in ``TDstressFS`` (src/TDstressFS_C.cpp:506-513).

Where the reference's *stress* path (the C++ behind the ctypes boundary) takes
a different branch from the displacement code, the stress path is followed:

* trimode 0 points (receiver on a TD edge, in plane) give zero strain from the
  full-space part -- TDstressFS_C.cpp:446-477 leaves them at zero (the NumPy
  twin writes NaN, SH_greenfunction.py:296-302);
* a free-surface-correction side is skipped when |beta| or |pi-beta| < 1e-4
  (TDstressFS_C.cpp:1523-1552, SH_greenfunction.py:744; the displacement code
  uses 2.2e-10);
* for a TD with all vertices at z = 0 the image's xz and yz strain are negated
  (TDstressFS_C.cpp:1998-2007).

Precision: numpy ``longdouble`` (x87 80-bit on x86-64, eps ~1.1e-19) with the
exact pi, so differences against the double-precision reference measure the
reference's own rounding.

Pins (see tests/test_oracle_td.py): the reference C++ compiled from source
(oracle/_ref) on random and structured pairs, committed golden vectors made
from it, linearity in the Burgers vector, Hooke's trace relation, equilibrium
div(sigma) = 0 and a traction-free free surface by finite differences, and the
far-field point-source (double-couple) limit.
"""
import numpy as np

LD = np.longdouble
PI = np.arccos(LD(-1))

__all__ = ["td_stress", "td_grad", "td_displacement", "element_frames", "traction_gf", "PI"]


# ---------------------------------------------------------------------------
# Forward-mode dual numbers: value v (m,), gradient d (3, m)
# ---------------------------------------------------------------------------
class Dual:
    __slots__ = ("v", "d")
    __array_priority__ = 1000

    def __init__(self, v, d):
        self.v = v
        self.d = d

    def __add__(self, o):
        if isinstance(o, Dual):
            return Dual(self.v + o.v, self.d + o.d)
        return Dual(self.v + o, self.d)

    __radd__ = __add__

    def __sub__(self, o):
        if isinstance(o, Dual):
            return Dual(self.v - o.v, self.d - o.d)
        return Dual(self.v - o, self.d)

    def __rsub__(self, o):
        return Dual(o - self.v, -self.d)

    def __neg__(self):
        return Dual(-self.v, -self.d)

    def __mul__(self, o):
        if isinstance(o, Dual):
            return Dual(self.v * o.v, self.d * o.v + o.d * self.v)
        return Dual(self.v * o, self.d * o)

    __rmul__ = __mul__

    def __truediv__(self, o):
        if isinstance(o, Dual):
            q = self.v / o.v
            return Dual(q, (self.d - o.d * q) / o.v)
        return Dual(self.v / o, self.d / o)

    def __rtruediv__(self, o):
        q = o / self.v
        return Dual(q, -self.d * (q / self.v))


def _sqrt(a):
    s = np.sqrt(a.v)
    return Dual(s, a.d / (2 * s))


def _log(a):
    return Dual(np.log(a.v), a.d / a.v)


def _atan(a):
    return Dual(np.arctan(a.v), a.d / (1 + a.v * a.v))


def _atan2(n, d):
    r2 = n.v * n.v + d.v * d.v
    return Dual(np.arctan2(n.v, d.v), (n.d * d.v - d.d * n.v) / r2)


def _where(mask, a, b):
    return Dual(np.where(mask, a.v, b.v), np.where(mask, a.d, b.d))


def _zero_like(a):
    return Dual(np.zeros_like(a.v), np.zeros_like(a.d))


def _dot3(u, w):
    return u[..., 0] * w[..., 0] + u[..., 1] * w[..., 1] + u[..., 2] * w[..., 2]


def _cross(u, w):
    return np.stack([u[..., 1] * w[..., 2] - u[..., 2] * w[..., 1],
                     u[..., 2] * w[..., 0] - u[..., 0] * w[..., 2],
                     u[..., 0] * w[..., 1] - u[..., 1] * w[..., 0]], axis=-1)


def _unit(u):
    return u / np.sqrt(_dot3(u, u))[..., None]


# ---------------------------------------------------------------------------
# Angular dislocation in full space (DH_greenfunction.py:71-104)
# ---------------------------------------------------------------------------
def _angdis_disp(x, y, z, alpha, bx, by, bz, nu):
    cosA = np.cos(alpha)
    sinA = np.sin(alpha)
    eta = y * cosA - z * sinA
    zeta = y * sinA + z * cosA
    r = _sqrt(x * x + y * y + z * z)
    # keep the logarithms real (DH_greenfunction.py:81-82)
    zeta = _where(zeta.v > r.v, r, zeta)
    z = _where(z.v > r.v, r, z)
    k = 1 / (8 * PI * (1 - nu))
    k2 = 1 - 2 * nu
    rz = r - zeta
    rr = r - z
    lz = _log(rz)
    lr = _log(rr)
    ux = bx * k * (x * y / r / rr - x * eta / r / rz)
    vx = bx * k * (eta * sinA / rz - y * eta / r / rz + y * y / r / rr
                   + k2 * (cosA * lz - lr))
    wx = bx * k * (eta * cosA / rz - y / r - eta * z / r / rz - k2 * sinA * lz)
    uy = by * k * (x * x * cosA / r / rz - x * x / r / rr - k2 * (cosA * lz - lr))
    vy = by * k * x * (y * cosA / r / rz - sinA * cosA / rz - y / r / rr)
    wy = by * k * x * (z * cosA / r / rz - cosA * cosA / rz + 1 / r)
    uz = bz * sinA * k * (k2 * lz - x * x / r / rz)
    vz = bz * sinA * k * x * (sinA / rz - y / r / rz)
    wz = bz * sinA * k * x * (cosA / rz - z / r / rz)
    return ux + uy + uz, vx + vy + vz, wx + wy + wz


def _td_setup(X1, angle, bx, by, bz, nu, vy, vz, sy, sz):
    """ADCS of one TD vertex (DH_greenfunction.py:40-69)."""
    dy = X1[1] - vy
    dz = X1[2] - vz
    y1 = dy * sz - dz * sy
    z1 = dy * sy + dz * sz
    by1 = sz * by - sy * bz
    bz1 = sy * by + sz * bz
    u, v0, w0 = _angdis_disp(X1[0], y1, z1, -PI + angle, bx, by1, bz1, nu)
    return u, v0 * sz + w0 * sy, w0 * sz - v0 * sy


# ---------------------------------------------------------------------------
# TD frames and configuration (TDstressFS_C.cpp:300-392, :125-166)
# ---------------------------------------------------------------------------
def _td_frame(P1, P2, P3):
    Vn = _unit(_cross(P2 - P1, P3 - P1))
    Vs = np.stack([-Vn[..., 1], Vn[..., 0], np.zeros_like(Vn[..., 0])], axis=-1)
    ns = np.sqrt(_dot3(Vs, Vs))
    flat = ns == 0
    # horizontal TD: strike along +/-y (TDstressFS_C.cpp:1826-1829)
    Vs_flat = np.stack([np.zeros_like(ns), Vn[..., 2], np.zeros_like(ns)], axis=-1)
    Vs = np.where(flat[..., None], Vs_flat, Vs / np.where(flat, 1, ns)[..., None])
    Vd = _cross(Vn, Vs)
    return Vn, Vs, Vd


def _trimode(x, y, z, p1, p2, p3):
    """+1 / -1 / 0 configuration (TDstressFS_C.cpp:125-166)."""
    x1, x2, x3 = p1[..., 1], p2[..., 1], p3[..., 1]
    y1, y2, y3 = p1[..., 2], p2[..., 2], p3[..., 2]
    den = (x1 - x3) * (y2 - y3) - (x2 - x3) * (y1 - y3)
    a = ((y - x3) * (y2 - y3) - (x2 - x3) * (z - y3)) / den
    b = ((x1 - x3) * (z - y3) - (y - x3) * (y1 - y3)) / den
    c = 1 - a - b
    tm = np.ones(a.shape, dtype=np.int64)
    neg = (((a <= 0) & (b > c) & (c > a)) | ((b <= 0) & (c > a) & (a > b))
           | ((c <= 0) & (a > b) & (b > c)))
    tm[neg] = -1
    zero = (((a == 0) & (b >= 0) & (c >= 0)) | ((a >= 0) & (b == 0) & (c >= 0))
            | ((a >= 0) & (b >= 0) & (c == 0)))
    tm[zero & ~neg] = 0
    tm[(tm == 0) & (x != 0)] = 1
    return tm


def _seed(R):
    m = R.shape[0]
    out = []
    for j in range(3):
        d = np.zeros((3, m), dtype=LD)
        d[j] = 1
        out.append(Dual(R[:, j].astype(LD), d))
    return out


def _grad_fs(Xd, P1, P2, P3, bx, by, bz, nu):
    """Full-space TD displacement (TDdispFS) with gradient; returns uE (3 Duals)."""
    Vn, Vs, Vd = _td_frame(P1, P2, P3)
    axes = (Vn, Vs, Vd)
    X1 = [(Xd[0] - P2[:, 0]) * ax[:, 0] + (Xd[1] - P2[:, 1]) * ax[:, 1]
          + (Xd[2] - P2[:, 2]) * ax[:, 2] for ax in axes]
    p1 = np.stack([_dot3(P1 - P2, ax) for ax in axes], axis=-1)
    p3 = np.stack([_dot3(P3 - P2, ax) for ax in axes], axis=-1)
    p2 = np.zeros_like(p1)
    e12 = _unit(p2 - p1)
    e13 = _unit(p3 - p1)
    e23 = _unit(p3 - p2)
    A = np.arccos(_dot3(e12, e13))
    B = np.arccos(_dot3(-e12, e23))
    C = np.arccos(_dot3(e23, e13))
    tm = _trimode(X1[0].v, X1[1].v, X1[2].v, p1, p2, p3)
    s = np.where(tm == -1, LD(-1), LD(1))
    u = v = w = None
    for vert, side, ang in ((p1, -e13, A), (p2, e12, B), (p3, e23, C)):
        du, dv, dw = _td_setup(X1, ang, bx, by, bz, nu, vert[:, 1], vert[:, 2],
                               s * side[:, 1], s * side[:, 2])
        u = du if u is None else u + du
        v = dv if v is None else v + dv
        w = dw if w is None else w + dw
    # Burgers function of the whole TD (DH_greenfunction.py:182-195)
    a = [p1[:, k] - X1[k] for k in range(3)]
    b = [-X1[k] for k in range(3)]
    c = [p3[:, k] - X1[k] for k in range(3)]
    na = _sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])
    nb = _sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2])
    nc = _sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2])
    num = (a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0])
           + a[2] * (b[0] * c[1] - b[1] * c[0]))
    ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2]
    ac = a[0] * c[0] + a[1] * c[1] + a[2] * c[2]
    bc = b[0] * c[0] + b[1] * c[1] + b[2] * c[2]
    den = na * nb * nc + ab * nc + ac * nb + bc * na
    Fi = _atan2(num, den) * (-2 / (4 * PI))
    u = u + Fi * bx
    v = v + Fi * by
    w = w + Fi * bz
    uE = [u * Vn[:, j] + v * Vs[:, j] + w * Vd[:, j] for j in range(3)]
    keep = tm != 0
    return [_where(keep, q, _zero_like(q)) for q in uE]


# ---------------------------------------------------------------------------
# Free-surface correction (DH_greenfunction.py:211-351)
# ---------------------------------------------------------------------------
def _angdis_disp_fsc(y1, y2, y3, beta, b1, b2, b3, nu, a):
    sinB = np.sin(beta)
    cosB = np.cos(beta)
    cotB = 1 / np.tan(beta)
    y3b = y3 + 2 * a
    z1b = y1 * cosB + y3b * sinB
    z3b = y3b * cosB - y1 * sinB
    r2b = y1 * y1 + y2 * y2 + y3b * y3b
    rb = _sqrt(r2b)
    r3b = rb * rb * rb
    Fib = 2 * _atan(-y2 / (y1 - (rb + y3b) / np.tan(beta / 2)))
    k = 1 / (4 * PI * (1 - nu))
    n = nu
    ry = rb + y3b
    rz = rb + z3b
    W = y3b - a
    lry = _log(ry)
    lrz = _log(rz)
    v1cb1 = b1 * k * (-2 * (1 - n) * (1 - 2 * n) * Fib * cotB * cotB
                      + (1 - 2 * n) * y2 / ry * ((1 - 2 * n - a / rb) * cotB - y1 / ry * (n + a / rb))
                      + (1 - 2 * n) * y2 * cosB * cotB / rz * (cosB + a / rb)
                      + a * y2 * W * cotB / r3b
                      + y2 * W / (rb * ry) * (-(1 - 2 * n) * cotB + y1 / ry * (2 * n + a / rb) + a * y1 / r2b)
                      + y2 * W / (rb * rz) * (cosB / rz * ((rb * cosB + y3b) * ((1 - 2 * n) * cosB - a / rb) * cotB
                                                          + 2 * (1 - n) * (rb * sinB - y1) * cosB)
                                              - a * y3b * cosB * cotB / r2b))
    v2cb1 = b1 * k * ((1 - 2 * n) * ((2 * (1 - n) * cotB * cotB - n) * lry
                                     - (2 * (1 - n) * cotB * cotB + 1 - 2 * n) * cosB * lrz)
                      - (1 - 2 * n) / ry * (y1 * cotB * (1 - 2 * n - a / rb) + n * y3b - a
                                            + y2 * y2 / ry * (n + a / rb))
                      - (1 - 2 * n) * z1b * cotB / rz * (cosB + a / rb)
                      - a * y1 * W * cotB / r3b
                      + W / ry * (-2 * n + 1 / rb * ((1 - 2 * n) * y1 * cotB - a)
                                  + y2 * y2 / (rb * ry) * (2 * n + a / rb) + a * y2 * y2 / r3b)
                      + W / rz * (cosB * cosB - 1 / rb * ((1 - 2 * n) * z1b * cotB + a * cosB)
                                  + a * y3b * z1b * cotB / r3b
                                  - 1 / (rb * rz) * (y2 * y2 * cosB * cosB - a * z1b * cotB / rb * (rb * cosB + y3b))))
    v3cb1 = b1 * k * (2 * (1 - n) * (((1 - 2 * n) * Fib * cotB) + (y2 / ry * (2 * n + a / rb))
                                     - (y2 * cosB / rz * (cosB + a / rb)))
                      + y2 * W / rb * (2 * n / ry + a / (rb * rb))
                      + y2 * W * cosB / (rb * rz) * (1 - 2 * n - (rb * cosB + y3b) / rz * (cosB + a / rb)
                                                     - a * y3b / rb / rb))
    v1cb2 = b2 * k * ((1 - 2 * n) * ((2 * (1 - n) * cotB * cotB + n) * lry
                                     - (2 * (1 - n) * cotB * cotB + 1) * cosB * lrz)
                      + (1 - 2 * n) / ry * (-(1 - 2 * n) * y1 * cotB + n * y3b - a + a * y1 * cotB / rb
                                            + y1 * y1 / ry * (n + a / rb))
                      - (1 - 2 * n) * cotB / rz * (z1b * cosB - a * (rb * sinB - y1) / (rb * cosB))
                      - a * y1 * W * cotB / r3b
                      + W / ry * (2 * n + 1 / rb * ((1 - 2 * n) * y1 * cotB + a)
                                  - y1 * y1 / (rb * ry) * (2 * n + a / rb) - a * y1 * y1 / r3b)
                      + W * cotB / rz * (-cosB * sinB + a * y1 * y3b / (r3b * cosB)
                                         + (rb * sinB - y1) / rb * (2 * (1 - n) * cosB
                                                                    - (rb * cosB + y3b) / rz * (1 + a / (rb * cosB)))))
    v2cb2 = b2 * k * (2 * (1 - n) * (1 - 2 * n) * Fib * cotB * cotB
                      + (1 - 2 * n) * y2 / ry * (-(1 - 2 * n - a / rb) * cotB + y1 / ry * (n + a / rb))
                      - (1 - 2 * n) * y2 * cotB / rz * (1 + a / (rb * cosB))
                      - a * y2 * W * cotB / r3b
                      + y2 * W / (rb * ry) * ((1 - 2 * n) * cotB - 2 * n * y1 / ry - a * y1 / rb * (1 / rb + 1 / ry))
                      + y2 * W * cotB / (rb * rz) * (-2 * (1 - n) * cosB + (rb * cosB + y3b) / rz * (1 + a / (rb * cosB))
                                                     + a * y3b / (r2b * cosB)))
    v3cb2 = b2 * k * (-2 * (1 - n) * (1 - 2 * n) * cotB * (lry - cosB * lrz)
                      - 2 * (1 - n) * y1 / ry * (2 * n + a / rb)
                      + 2 * (1 - n) * z1b / rz * (cosB + a / rb)
                      + W / rb * ((1 - 2 * n) * cotB - 2 * n * y1 / ry - a * y1 / r2b)
                      - W / rz * (cosB * sinB + (rb * cosB + y3b) * cotB / rb * (2 * (1 - n) * cosB - (rb * cosB + y3b) / rz)
                                  + a / rb * (sinB - y3b * z1b / r2b - z1b * (rb * cosB + y3b) / (rb * rz))))
    v1cb3 = b3 * k * ((1 - 2 * n) * (y2 / ry * (1 + a / rb) - y2 * cosB / rz * (cosB + a / rb))
                      - y2 * W / rb * (a / r2b + 1 / ry)
                      + y2 * W * cosB / (rb * rz) * ((rb * cosB + y3b) / rz * (cosB + a / rb) + a * y3b / r2b))
    v2cb3 = b3 * k * ((1 - 2 * n) * (-sinB * lrz - y1 / ry * (1 + a / rb) + z1b / rz * (cosB + a / rb))
                      + y1 * W / rb * (a / r2b + 1 / ry)
                      - W / rz * (sinB * (cosB - a / rb) + z1b / rb * (1 + a * y3b / r2b)
                                  - 1 / (rb * rz) * (y2 * y2 * cosB * sinB - a * z1b / rb * (rb * cosB + y3b))))
    v3cb3 = b3 * k * (2 * (1 - n) * Fib + 2 * (1 - n) * (y2 * sinB / rz * (cosB + a / rb))
                      + y2 * W * sinB / (rb * rz) * (1 + (rb * cosB + y3b) / rz * (cosB + a / rb) + a * y3b / r2b))
    return v1cb1 + v1cb2 + v1cb3, v2cb1 + v2cb2 + v2cb3, v3cb1 + v3cb2 + v3cb3


def _fsc_side(Xd, Bv, PA, PB, nu):
    """One TD side of the free-surface correction (AngSetupFSC), EFCS Duals."""
    side = PB - PA
    beta = np.arccos(-side[:, 2] / np.sqrt(_dot3(side, side)))
    eps = LD("1e-4")
    skip = (np.abs(beta) < eps) | (np.abs(PI - beta) < eps)
    h = np.sqrt(side[:, 0] ** 2 + side[:, 1] ** 2)
    hs = np.where(h == 0, 1, h)
    ey1 = np.stack([side[:, 0] / hs, side[:, 1] / hs, np.zeros_like(h)], axis=-1)
    ey3 = np.zeros_like(ey1)
    ey3[:, 2] = -1
    ey2 = _cross(ey3, ey1)
    eys = (ey1, ey2, ey3)
    yA = [(Xd[0] - PA[:, 0]) * e[:, 0] + (Xd[1] - PA[:, 1]) * e[:, 1]
          + (Xd[2] - PA[:, 2]) * e[:, 2] for e in eys]
    yAB = [_dot3(side, e) for e in eys]
    yB = [yA[k] - yAB[k] for k in range(3)]
    bv = [_dot3(Bv, e) for e in eys]
    cfgI = beta * yA[0].v >= 0
    ang = np.where(cfgI, -PI + beta, beta)
    # keep degenerate sides finite before masking them out
    ang = np.where(skip, PI / 2, ang)
    vA = _angdis_disp_fsc(yA[0], yA[1], yA[2], ang, bv[0], bv[1], bv[2], nu, -PA[:, 2])
    vB = _angdis_disp_fsc(yB[0], yB[1], yB[2], ang, bv[0], bv[1], bv[2], nu, -PB[:, 2])
    v = [vB[k] - vA[k] for k in range(3)]
    uE = [v[0] * ey1[:, j] + v[1] * ey2[:, j] + v[2] * ey3[:, j] for j in range(3)]
    return [_where(~skip, q, _zero_like(q)) for q in uE]


def _grad_fsc(Xd, P1, P2, P3, bx, by, bz, nu):
    Vn, Vs, Vd = _td_frame(P1, P2, P3)
    Bv = Vn * bx[:, None] + Vs * by[:, None] + Vd * bz[:, None]
    out = None
    for PA, PB in ((P1, P2), (P2, P3), (P3, P1)):
        u = _fsc_side(Xd, Bv, PA, PB, nu)
        out = u if out is None else [out[j] + u[j] for j in range(3)]
    return out


def _G(uE):
    """Stack the gradients of the 3 displacement components -> (m, 3, 3)."""
    return np.stack([q.d.T for q in uE], axis=1)


def td_grad(R, P1, P2, P3, b, nu, halfspace=True):
    """Displacement gradient G[m, i, j] = du_i/dX_j of a TD with TDCS Burgers
    vector b = (Ts, Ss, Ds) per pair (TDstressHS / TDstressFS semantics)."""
    R = np.asarray(R, dtype=LD)
    P1 = np.asarray(P1, dtype=LD)
    P2 = np.asarray(P2, dtype=LD)
    P3 = np.asarray(P3, dtype=LD)
    b = np.broadcast_to(np.asarray(b, dtype=LD), R.shape)
    bx, by, bz = b[:, 0], b[:, 1], b[:, 2]
    nu = LD(nu)
    Xd = _seed(R)
    G = _G(_grad_fs(Xd, P1, P2, P3, bx, by, bz, nu))
    if halfspace:
        G = G + _G(_grad_fsc(Xd, P1, P2, P3, bx, by, bz, nu))
        M = np.array([1, 1, -1], dtype=LD)
        Gi = _G(_grad_fs(Xd, P1 * M, P2 * M, P3 * M, bx, by, bz, nu))
        flush = (P1[:, 2] == 0) & (P2[:, 2] == 0) & (P3[:, 2] == 0)
        if np.any(flush):
            F = np.ones((3, 3), dtype=LD)
            F[0, 2] = F[2, 0] = F[1, 2] = F[2, 1] = -1
            Gi = np.where(flush[:, None, None], Gi * F, Gi)
        G = G + Gi
    return G


def td_displacement(R, P1, P2, P3, b, nu, halfspace=True):
    """Displacement (m, 3) in EFCS -- the values the gradient is taken of
    (TDdispHS / TDdispFS, DH_greenfunction.py:106-208, :353-383; the
    surface-flush sign conventions of the displacement code are not applied)."""
    R = np.asarray(R, dtype=LD)
    P1 = np.asarray(P1, dtype=LD)
    P2 = np.asarray(P2, dtype=LD)
    P3 = np.asarray(P3, dtype=LD)
    b = np.broadcast_to(np.asarray(b, dtype=LD), R.shape)
    bx, by, bz = b[:, 0], b[:, 1], b[:, 2]
    Xd = _seed(R)
    u = _grad_fs(Xd, P1, P2, P3, bx, by, bz, LD(nu))
    if halfspace:
        M = np.array([1, 1, -1], dtype=LD)
        f = _grad_fsc(Xd, P1, P2, P3, bx, by, bz, LD(nu))
        im = _grad_fs(Xd, P1 * M, P2 * M, P3 * M, bx, by, bz, LD(nu))
        u = [u[j] + f[j] + im[j] for j in range(3)]
    return np.stack([q.v for q in u], axis=1)


def strain_stress(G, mu, lam):
    """(G + G^T)/2 and Hooke's law, components [xx, yy, zz, xy, xz, yz]."""
    E = 0.5 * (G + np.swapaxes(G, 1, 2))
    eps = np.stack([E[:, 0, 0], E[:, 1, 1], E[:, 2, 2], E[:, 0, 1], E[:, 0, 2], E[:, 1, 2]], axis=1)
    tr = eps[:, 0] + eps[:, 1] + eps[:, 2]
    sig = 2 * LD(mu) * eps
    sig[:, :3] += LD(lam) * tr[:, None]
    return sig, eps


def td_stress(R, P1, P2, P3, Ss, Ds, Ts, mu, lam, halfspace=True):
    """Stress and strain (m, 6) for pairs (receiver R[i], TD P1..P3[i])."""
    nu = 1 / (1 + LD(lam) / LD(mu)) / 2
    b = np.array([Ts, Ss, Ds], dtype=LD)
    G = td_grad(R, P1, P2, P3, b, nu, halfspace)
    return strain_stress(G, mu, lam)


# ---------------------------------------------------------------------------
# Element frames and dense traction Green's functions
# ---------------------------------------------------------------------------
def element_frames(nodes, tri):
    """Centroids and local frames (readmsh.py:75-134, counter-clockwise order).

    ``tri`` is 1-based connectivity.  Returns xg (N,3), eleVec (N,9) with rows
    [e1, e2, n]."""
    nodes = np.asarray(nodes, dtype=np.float64)
    tri = np.asarray(tri, dtype=np.int64) - 1
    xa, xb, xc = nodes[tri[:, 0]], nodes[tri[:, 1]], nodes[tri[:, 2]]
    xg = ((xa + xb) + xc) / 3.0
    n = np.cross(xb - xa, xc - xa)
    n = n / np.sqrt((n * n).sum(1))[:, None]
    e1 = np.stack([n[:, 1], -n[:, 0], np.zeros(len(n))], axis=1)
    steep = np.abs(n[:, 2]) < 1
    e1n = np.sqrt(e1[:, 0] ** 2 + e1[:, 1] ** 2)
    e1 = np.where(steep[:, None], e1 / np.where(steep, e1n, 1)[:, None],
                  np.array([1.0, 0.0, 0.0]))
    e2 = np.stack([n[:, 1] * e1[:, 2] - n[:, 2] * e1[:, 1],
                   n[:, 2] * e1[:, 0] - n[:, 0] * e1[:, 2],
                   n[:, 0] * e1[:, 1] - n[:, 1] * e1[:, 0]], axis=1)
    return xg, np.concatenate([e1, e2, n], axis=1)


def traction_gf(nodes, tri, mu, lam, halfspace=True, rows=None, cols=None):
    """Dense traction Green's functions A1s, A2s, Bs, A1d, A2d, Bd [rows, cols]
    (QDsim_gpu.py:1745-1785 worker + GetTtstress :1400-1417): row = receiver
    centroid, column = source TD with unit strike (Ss=1) or dip (Ds=1) slip."""
    nodes = np.asarray(nodes, dtype=np.float64)
    tri = np.asarray(tri, dtype=np.int64)
    xg, ev = element_frames(nodes, tri)
    N = len(tri)
    rows = np.arange(N) if rows is None else np.asarray(rows)
    cols = np.arange(N) if cols is None else np.asarray(cols)
    rr, cc = np.meshgrid(rows, cols, indexing="ij")
    rr = rr.ravel()
    cc = cc.ravel()
    t0 = tri[cc] - 1
    P1, P2, P3 = nodes[t0[:, 0]], nodes[t0[:, 1]], nodes[t0[:, 2]]
    R = xg[rr]
    out = []
    for Ss, Ds in ((1.0, 0.0), (0.0, 1.0)):
        sig, _ = td_stress(R, P1, P2, P3, Ss, Ds, 0.0, mu, lam, halfspace)
        e1 = ev[rr, 0:3].astype(LD)
        e2 = ev[rr, 3:6].astype(LD)
        nn = ev[rr, 6:9].astype(LD)
        t = np.stack([sig[:, 0] * nn[:, 0] + sig[:, 3] * nn[:, 1] + sig[:, 4] * nn[:, 2],
                      sig[:, 3] * nn[:, 0] + sig[:, 1] * nn[:, 1] + sig[:, 5] * nn[:, 2],
                      sig[:, 4] * nn[:, 0] + sig[:, 5] * nn[:, 1] + sig[:, 2] * nn[:, 2]], axis=1)
        shape = (len(rows), len(cols))
        out.append(_dot3(t, e1).reshape(shape))
        out.append(_dot3(t, e2).reshape(shape))
        out.append(_dot3(t, nn).reshape(shape))
    return tuple(np.asarray(m, dtype=np.float64) for m in out)
