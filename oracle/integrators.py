"""NumPy restatement of the reference's other time integrators: ForwardEuler,
AB2 and the embedded Runge-Kutta driver adaptiveRK with the Bogacki-Shampine
5(4) tableau (BS5_fixed, BS5_adaptive), plus the time loop they run in.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

What it restates (reference file:line):
  * adaptiveRK (FSAL offset bookkeeping, stage sums, solution and error sums,
    errnorm "2", step-size controller, rejection) ... spectralDNS/maths/integrators.py:15-147
  * ForwardEuler, AB2 ................................ integrators.py:161-175
  * BS5 tableau, err_order 4, fsal, TOL as aTOL/rTOL . integrators.py:193-224
  * time loop with the integrator's returned dt ...... spectralDNS/__init__.py:94-112
  * last-step adjustment ............................. spectralDNS/solvers/NS.py:112-122
  * est /= sqrt(prod(T.shape(True))): T.shape(True) is the spectral shape
    (N0, N1, N2//2+1), as K2 = np.zeros(T.shape(True)) in solvers/Bq2D.py:42.

How it is pinned (tests/test_oracle_integrators.py):
  * linear shear modes, where the RHS is exactly -nu k^2 u_hat: ForwardEuler's
    amplification (1 + z)^n, AB2's two-step recursion and its second-order
    convergence, and BS5's stability function R(z) = 1 + z b^T (I - zA)^-1 1
    (computed by a linear solve, not stage by stage), which must also agree
    with exp(z) to O(z^6) and whose embedded error is O(z^5);
  * the reference's own Taylor-Green regression values k(0.1), w(0.1) at the
    tolerances of tests/test_NSVV.py:76-91 (ntol 4 for ForwardEuler and AB2,
    7 for BS5) on the uniform and the non-uniform box.
"""
import numpy as np

# Bogacki-Shampine 5(4) (integrators.py:200-209): the published coefficients.
BS5_A = np.array([[0, 0, 0, 0, 0, 0, 0, 0],
                  [1 / 6, 0, 0, 0, 0, 0, 0, 0],
                  [2 / 27, 4 / 27, 0, 0, 0, 0, 0, 0],
                  [183 / 1372, -162 / 343, 1053 / 1372, 0, 0, 0, 0, 0],
                  [68 / 297, -4 / 11, 42 / 143, 1960 / 3861, 0, 0, 0, 0],
                  [597 / 22528, 81 / 352, 63099 / 585728, 58653 / 366080, 4617 / 20480, 0, 0, 0],
                  [174197 / 959244, -30942 / 79937, 8152137 / 19744439, 666106 / 1039181, -29421 / 29068,
                   482048 / 414219, 0, 0],
                  [587 / 8064, 0, 4440339 / 15491840, 24353 / 124800, 387 / 44800, 2152 / 5985, 7267 / 94080, 0]])
BS5_B = np.array([587 / 8064, 0, 4440339 / 15491840, 24353 / 124800, 387 / 44800, 2152 / 5985, 7267 / 94080, 0])
BS5_BHAT = np.array([2479 / 34992, 0, 123 / 416, 612941 / 3411720, 43 / 1440, 2272 / 6561, 79937 / 1113912,
                     3293 / 556956])


class ForwardEuler:
    """u0 += rhs*dt (integrators.py:161-165)."""

    def __init__(self, rhs_fn, fl):
        self.rhs_fn, self.fl = rhs_fn, fl

    def step(self, u0, dt, tstep):
        dt = self.fl(dt)
        u0 = u0 + self.rhs_fn(u0) * dt
        return u0, dt, dt


class AB2:
    """Adams-Bashforth 2 with a ForwardEuler first step (integrators.py:167-175);
    u1 holds rhs*dt of the previous step."""

    def __init__(self, rhs_fn, fl):
        self.rhs_fn, self.fl = rhs_fn, fl
        self.u1 = None

    def step(self, u0, dt, tstep):
        dt = self.fl(dt)
        rhs = self.rhs_fn(u0)
        u0 = u0.copy()
        if tstep == 0:
            u0 += rhs * dt
        else:
            u0 += (1.5 * rhs * dt - 0.5 * self.u1)
        self.u1 = rhs * dt
        return u0, dt, dt


class AdaptiveRK:
    """adaptiveRK with the BS5 tableau, errnorm "2", no predictive controller
    (integrators.py:15-147, 193-224)."""

    def __init__(self, rhs_fn, fl, spectral_shape, tol, adaptive, A=BS5_A, b=BS5_B, bhat=BS5_BHAT,
                 err_order=4, fsal=True):
        self.rhs_fn, self.fl = rhs_fn, fl
        self.A, self.b, self.bhat = A.astype(fl), b.astype(fl), bhat.astype(fl)
        self.err_order, self.fsal, self.adaptive = err_order, fsal, adaptive
        self.tol = tol
        self.nspec = np.sqrt(np.array(spectral_shape).prod())
        self.offset = 0
        self.fY = None
        self.rejected = 0

    def step(self, u0, dt, tstep):
        A, b, bhat = self.A, self.b, self.bhat
        s = A.shape[0]
        if self.fY is None:
            self.fY = np.zeros((s,) + u0.shape, dtype=u0.dtype)
        fY = self.fY
        facmax, fac, facmin = 2, 0.8, 0.01
        dt = self.fl(dt)
        while True:
            dt_prev = dt
            if self.fsal:
                self.offset = (self.offset - 1) % s
            off = self.offset
            for i in range(s):
                if not self.fsal or (tstep == 0 or i != 0):
                    y = u0.copy()
                    for j in range(i):
                        y += dt * A[i, j] * fY[(j + off) % s]
                    fY[(i + off) % s] = self.rhs_fn(y)
            u0_new = u0.copy()
            u0_new += dt * b[0] * fY[off % s]
            err = dt * (b[0] - bhat[0]) * fY[off % s]
            for j in range(1, s):
                u0_new += dt * b[j] * fY[(j + off) % s]
                err += dt * (b[j] - bhat[j]) * fY[(j + off) % s]
            sc = self.tol + np.maximum(np.abs(u0), np.abs(u0_new)) * self.tol
            nsq = np.array([np.sum(np.abs(err[k] / sc[k]).astype(np.float64) ** 2) for k in range(u0.shape[0])])
            est = np.max(np.sqrt(nsq)) / self.nspec
            exponent = 1.0 / (self.err_order + 1)
            factor = min(facmax, max(facmin, fac * pow((1 / est), exponent)))
            if self.adaptive:
                dt = dt * factor
                if est > 1.0:
                    facmax = 1
                    self.rejected += 1
                    if self.fsal:
                        self.offset += 1
                    continue
            break
        return u0_new, dt, dt_prev


def make(name, grid, nu, tol=1e-6):
    """Integrator object of the reference's name over grid.compute_rhs."""
    rhs = lambda u: grid.compute_rhs(u, nu)  # noqa: E731
    fl = grid.float
    if name == "ForwardEuler":
        return ForwardEuler(rhs, fl)
    if name == "AB2":
        return AB2(rhs, fl)
    if name in ("BS5_adaptive", "BS5_fixed"):
        return AdaptiveRK(rhs, fl, (grid.N[0], grid.N[1], grid.Nh), tol, name == "BS5_adaptive")
    raise ValueError(name)


def solve(integ, u_hat, T, dt, t=0.0, dt_precision="double"):
    """The reference time loop (spectralDNS/__init__.py:94-112 with
    NS.end_of_tstep, NS.py:112-122) around integ.step; returns (u_hat, t,
    tstep, dts taken).  With dt_precision "single" every read of params.dt is
    cast to float32 (config.py:123-127) and t accumulates the float32 dt_took
    in float32, as the host loop of a single-precision run does (the last
    step can then be a tiny one that closes the float32 rounding gap to T)."""
    rd = np.float32 if dt_precision == "single" else np.float64
    stored = dt
    tstep = 0
    taken = []
    while t + rd(stored) <= T + 1e-12:
        u_hat, stored, dt_took = integ.step(u_hat, rd(stored), tstep)
        t += rd(dt_took)
        tstep += 1
        taken.append(float(dt_took))
        if abs(t - T) < 1e-12:
            break
        if abs(t + rd(stored) - T) < 1e-12 or t + rd(stored) >= T + 1e-12:
            stored = T - t
    return u_hat, t, tstep, taken
