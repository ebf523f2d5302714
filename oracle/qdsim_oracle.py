"""NumPy FP64 oracle of the dense quasi-dynamic stepping.

TEST INFRASTRUCTURE ONLY -- never imported by the product package.

Restates, operation for operation where it matters, the reference's
``QDsim_gpu.QDsim`` stepping:

* frozen stiffness terms per step          QDsim_gpu.py:942-957 (GPU), :1029-1042 (CPU)
* rate-and-state RHS ``derivative[_gpu]``  QDsim_gpu.py:830-891, :772-828
* Dormand-Prince 5(4) attempt              QDsim_gpu.py:1128-1262, :1266-1397
* accept / reject / step-size control      QDsim_gpu.py:959-985, :1052-1075
* post-step update                         QDsim_gpu.py:987-1027, :1077-1122
* steady-state initial condition           QDsim_gpu.py:714-770 (read_parameter)

``gpu_semantics=True`` reproduces the GPU path's error estimate, where
``cp.maximum(a, b, c)`` writes into ``c`` and so drops the normal-traction
error (QDsim_gpu.py:1238); ``False`` is the CPU path's ``max`` of all three
(QDsim_gpu.py:1373).  Python's builtin ``min``/``max`` NaN semantics are kept.

Pins (tests/test_oracle_qdsim.py): the Dormand-Prince tableau against its
order conditions and an ODE with a closed-form solution, steady state as a
fixed point, finite-difference checks of the RHS partial derivatives, the
early-step closed form dt_k = 1e-3 * 1.5^k seen in the reference's notebook
traces, and accept/reject arithmetic on constructed error values.
"""
import numpy as np

# Dormand-Prince coefficients (QDsim_gpu.py:1129-1159)
B21 = .2
B31, B32 = 3. / 40, 9. / 40.
B41, B42, B43 = 44. / 45., -56. / 15, 32. / 9
B51, B52, B53, B54 = 19372. / 6561., -25360 / 2187., 64448. / 6561., -212. / 729.
B61, B62, B63, B64, B65 = 9017. / 3168., -355. / 33., -46732. / 5247., 49. / 176., -5103. / 18656.
B71, B73, B74, B75, B76 = 35. / 384., 500. / 1113., 125. / 192., -2187. / 6784., 11. / 84.
B81, B83, B84, B85, B86, B87 = 5179. / 57600., 7571. / 16695., 393. / 640., -92097. / 339200., 187. / 2100., 1. / 40.

RELTOL = 1e-4


def py_max(a, b):
    """Python's builtin max(a, b) (keeps a unless b > a)."""
    return b if b > a else a


def py_min(a, b):
    return b if b < a else a


def steady_state(a, b, f0, Tn, V, V0, Vpl, P=0.0):
    """Tt and state at steady state for slip rate V (QDsim_gpu.py:760-766)."""
    Tt = (Tn - P * 1e-6) * a * np.arcsinh(V / (2.0 * V0) * np.exp((f0 + b * np.log(V0 / Vpl)) / a))
    state = np.log(np.sinh(Tt / (Tn - P * 1e-6) / a) * 2.0 * V0 / V) * a
    return Tt, state


class DenseSim:
    """Dense single-simulation stepper on host FP64 arrays."""

    def __init__(self, mats, a, b, dc, f0, Tno, Tt1o, Tt2o, state, slipv1, slipv2, rake0, V0, Vpl, mu, Cs,
                 shear_loading=None, normal_loading=None, fix_Tn=False, gpu_semantics=True):
        self.A1s, self.A2s, self.Bs, self.A1d, self.A2d, self.Bd = [np.asarray(m, np.float64) for m in mats]
        n = len(a)
        z = np.zeros(n)
        self.a, self.b, self.dc, self.f0 = (np.asarray(x, np.float64).copy() for x in (a, b, dc, f0))
        self.Tno, self.Tt1o, self.Tt2o, self.state = (np.asarray(x, np.float64).copy()
                                                      for x in (Tno, Tt1o, Tt2o, state))
        self.slipv1 = np.asarray(slipv1, np.float64).copy()
        self.slipv2 = np.asarray(slipv2, np.float64).copy()
        self.slipv = np.sqrt(self.slipv1 ** 2 + self.slipv2 ** 2)
        self.rake0 = np.asarray(rake0, np.float64).copy()
        self.shear_loading = z.copy() if shear_loading is None else np.asarray(shear_loading, np.float64)
        self.normal_loading = z.copy() if normal_loading is None else np.asarray(normal_loading, np.float64)
        self.P = z.copy()
        self.dPdt0 = z.copy()
        self.slip1, self.slip2, self.slip = z.copy(), z.copy(), z.copy()
        self.V0, self.Vpl_con, self.mu, self.Cs = V0, Vpl, mu, Cs
        self.fix_Tn = fix_Tn
        self.gpu = gpu_semantics
        self.time = 0.0
        self.maxslipv0 = float(np.max(self.slipv))

    def frozen(self):
        """AdotV1, AdotV2, dsigmadt from the current slip rates."""
        u1 = self.slipv1 - self.Vpl_con * np.cos(self.rake0)
        u2 = self.slipv2 - self.Vpl_con * np.sin(self.rake0)
        if self.fix_Tn:
            dsig = self.normal_loading
        else:
            dsig = np.dot(self.Bs, u1) + np.dot(self.Bd, u2) + self.normal_loading
        return np.dot(self.A1s, u1) + np.dot(self.A1d, u2), np.dot(self.A2s, u1) + np.dot(self.A2d, u2), dsig

    def derivative(self, Tno, Tt1o, Tt2o, state):
        Tno = Tno * 1e6
        Tt1o = Tt1o * 1e6
        Tt2o = Tt2o * 1e6
        P = self.P
        dPdt = self.dPdt0

        def safe_exp(x):
            return np.exp(np.clip(x, -700, 700))

        def safe_cosh(x):
            x = np.clip(x, -700, 700)
            return (np.exp(x) + np.exp(-x)) / 2

        def safe_sinh(x):
            x = np.clip(x, -700, 700)
            return (np.exp(x) - np.exp(-x)) / 2

        V0, a, b, dc, f0 = self.V0, self.a, self.b, self.dc, self.f0
        dV1dtau = 2 * V0 / (a * (Tno - P)) * safe_exp(-state / a) * safe_cosh(Tt1o / (a * (Tno - P)))
        dV2dtau = 2 * V0 / (a * (Tno - P)) * safe_exp(-state / a) * safe_cosh(Tt2o / (a * (Tno - P)))
        dV1dsigma = -2 * V0 * Tt1o / (a * (Tno - P) ** 2) * safe_exp(-state / a) * safe_cosh(Tt1o / (a * (Tno - P)))
        dV2dsigma = -2 * V0 * Tt2o / (a * (Tno - P) ** 2) * safe_exp(-state / a) * safe_cosh(Tt2o / (a * (Tno - P)))
        dV1dstate = -2 * V0 / a * safe_exp(-state / a) * safe_sinh(Tt1o / (a * (Tno - P)))
        dV2dstate = -2 * V0 / a * safe_exp(-state / a) * safe_sinh(Tt2o / (a * (Tno - P)))
        dstatedt = b / dc * (V0 * safe_exp((f0 - state) / b) - self.slipv)
        eta = self.mu / (2.0 * self.Cs)
        dtau1dt = (-self.AdotV1 + self.shear_loading - eta * (dV1dsigma * (self.dsigmadt - dPdt)
                                                              + dV1dstate * dstatedt)) / (1.0 + eta * dV1dtau)
        dtau2dt = (-self.AdotV2 + self.shear_loading - eta * (dV2dsigma * (self.dsigmadt - dPdt)
                                                              + dV2dstate * dstatedt)) / (1.0 + eta * dV2dtau)
        return dstatedt, self.dsigmadt / 1e6, dtau1dt / 1e6, dtau2dt / 1e6

    def attempt(self, h):
        """One Dormand-Prince attempt; returns y5 and the control quantities."""
        y = (self.Tno, self.Tt1o, self.Tt2o, self.state)
        d = self.derivative
        k1 = d(*y)
        s1, n1, t1, u1 = k1

        def comb(*terms):
            return [y[j] + h * sum(c * kk[j] for c, kk in terms) if len(terms) > 1 else y[j] + h * terms[0][0] * terms[0][1][j]
                    for j in range(4)]

        # order in the tuples: (Tno, Tt1o, Tt2o, state); derivative returns (state, Tno, Tt1, Tt2)
        def reorder(k):
            return (k[1], k[2], k[3], k[0])

        K1 = reorder(k1)
        Y = [y[j] + h * B21 * K1[j] for j in range(4)]
        K2 = reorder(d(*Y))
        Y = [y[j] + h * (B31 * K1[j] + B32 * K2[j]) for j in range(4)]
        K3 = reorder(d(*Y))
        Y = [y[j] + h * (B41 * K1[j] + B42 * K2[j] + B43 * K3[j]) for j in range(4)]
        K4 = reorder(d(*Y))
        Y = [y[j] + h * (B51 * K1[j] + B52 * K2[j] + B53 * K3[j] + B54 * K4[j]) for j in range(4)]
        K5 = reorder(d(*Y))
        Y = [y[j] + h * (B61 * K1[j] + B62 * K2[j] + B63 * K3[j] + B64 * K4[j] + B65 * K5[j]) for j in range(4)]
        K6 = reorder(d(*Y))
        Y5 = [y[j] + h * (B71 * K1[j] + B73 * K3[j] + B74 * K4[j] + B75 * K5[j] + B76 * K6[j]) for j in range(4)]
        K7 = reorder(d(*Y5))
        Y4 = [y[j] + h * (B81 * K1[j] + B83 * K3[j] + B84 * K4[j] + B85 * K5[j] + B86 * K6[j] + B87 * K7[j])
              for j in range(4)]
        del comb, s1, n1, t1, u1
        errs = [np.abs(Y4[j] - Y5[j]) for j in range(4)]
        rel1 = np.max(np.abs(errs[3] / Y4[3])) + 1e-10
        r_t1 = np.max(np.abs(errs[1] / Y4[1]))
        r_t2 = np.max(np.abs(errs[2] / Y4[2]))
        r_n = np.max(np.abs(errs[0] / Y4[0]))
        if self.gpu:
            rel2 = np.maximum(r_t1, r_t2) + 1e-10
        else:
            rel2 = max(r_t1, r_t2, r_n) + 1e-10
        c1 = rel1 / RELTOL
        c2 = rel2 / RELTOL
        hnew1 = h * 0.9 * (RELTOL / rel1) ** 0.2
        hnew2 = h * 0.9 * (RELTOL / rel2) ** 0.2
        return Y5, c1, c2, hnew1, hnew2

    def step(self, dttry):
        """One accepted step (QDsim_gpu.simu_forwardGPU); returns (h, dtnext).
        Raises RuntimeError where the reference calls sys.exit()."""
        self.AdotV1, self.AdotV2, self.dsigmadt = self.frozen()
        nrjct = 0
        h = dttry
        while True:
            Y5, c1, c2, hnew1, hnew2 = self.attempt(h)
            if py_max(c1, c2) < 1.0 and not (np.isnan(c1) or np.isnan(c2)):
                dtnext = py_min(hnew1, hnew2)
                dtnext = py_min(1.5 * h, dtnext)
                break
            nrjct += 1
            dtnext = py_min(hnew1, hnew2)
            h = py_max(0.5 * h, dtnext)
            if h < 1.e-15 or nrjct > 20:
                raise RuntimeError("error: dt is too small")
        self.nrjct = nrjct
        self.time = self.time + h
        self.Tno, self.Tt1o, self.Tt2o, self.state = Y5
        self.Tt = np.sqrt(self.Tt1o * self.Tt1o + self.Tt2o * self.Tt2o)
        self.slipv1 = (2.0 * self.V0) * np.exp(-self.state / self.a) * np.sinh(self.Tt1o / self.Tno / self.a)
        self.slipv2 = (2.0 * self.V0) * np.exp(-self.state / self.a) * np.sinh(self.Tt2o / self.Tno / self.a)
        self.slipv = np.sqrt(self.slipv1 * self.slipv1 + self.slipv2 * self.slipv2)
        self.rake = np.arctan2(self.slipv2, self.slipv1)
        self.slipv = np.where(self.slipv < 1e-30, 1e-30, self.slipv)
        self.maxslipv0 = float(np.max(self.slipv))
        self.slip1 = self.slip1 + self.slipv1 * h
        self.slip2 = self.slip2 + self.slipv2 * h
        self.slip = np.sqrt(self.slip1 * self.slip1 + self.slip2 * self.slip2)
        self.fric = self.Tt / (self.Tno - self.P * 1e-6)
        return h, dtnext


def from_init_columns(mats, cols, V0, Vpl, mu, Cs, gpu_semantics=True):
    """DenseSim from the 10-column init file layout (QDsim_gpu.py:714-770)."""
    cols = np.asarray(cols, np.float64)
    rake, a, b, dc, f0, _, Tn, V, sl, nl = cols.T
    Tt, state = steady_state(a, b, f0, Tn, V, V0, Vpl)
    x, y = np.cos(rake), np.sin(rake)
    return DenseSim(mats, a, b, dc, f0, Tn, Tt * x, Tt * y, state, V * x + 1e-16, V * y + 1e-16, rake,
                    V0, Vpl, mu, Cs, sl, nl, gpu_semantics=gpu_semantics)
