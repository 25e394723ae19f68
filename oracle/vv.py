"""NumPy restatement of the velocity-vorticity solver (spectralDNS/solvers/VV.py).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

What it restates (reference file:line), on oracle.ns.Grid's transforms:
  * compute_velocity: u_hat = 1j*(K/K^2 x w_hat) ........ solvers/VV.py:57-73
  * Conv (Vortex): F(u_dealias x w_dealias), then
    rhs = 1j*(K x v_hat) ................................. VV.py:91-108
  * Nyquist mask, add_linear (rhs -= nu K2 w_hat + Source) VV.py:110-146
  * RK4 on w_hat ........................................ maths/integrators.py:149-159
  * Taylor-Green start W_hat = 1j*(K x U_hat) ........... tests/TG.py:29-35

How it is pinned (tests/test_oracle_vv.py):
  * the reference's Taylor-Green regression values k(0.1), w(0.1) at ntol 7
    for the VV solver (tests/test_NSVV.py:9-16, 36-51: VV on the uniform and
    the non-uniform box), for the 2/3 and the 3/2 rule;
  * the NS oracle (pinned on its own, oracle/ns.py) carries the same flow:
    w_hat(t) of VV equals 1j*(K x u_hat(t)) of NS after the same RK4 steps,
    to rounding, since the curl commutes with the RK4 stages and removes the
    pressure gradient.
"""
import numpy as np


def cross2_k_over_k2(g, w_hat):
    """1j*(K_over_K2 x w_hat) (maths/cross.py:27-32 with a = K_over_K2)."""
    o = g.K_over_K2
    c = np.empty_like(w_hat)
    c[0] = o[1] * w_hat[2] - o[2] * w_hat[1]
    c[1] = o[2] * w_hat[0] - o[0] * w_hat[2]
    c[2] = o[0] * w_hat[1] - o[1] * w_hat[0]
    c *= 1j
    return c


def compute_rhs(g, w_hat, nu, source=None):
    """ComputeRHS of VV.py:117-146."""
    nu = g.float(nu)
    u_d = g.backward(cross2_k_over_k2(g, w_hat), dealiased=True)
    w_d = g.backward(w_hat, dealiased=True)
    v_hat = g.forward(g.cross1(u_d, w_d), dealiased=True)
    rhs = g.cross2(v_hat)
    if g.mask_nyquist_on:
        rhs = g.mask_nyquist(rhs)
    rhs -= nu * g.K2 * w_hat
    if source is not None:
        rhs += source
    return rhs


def rk4_step(g, w0, dt, nu, source=None):
    """RK4 (integrators.py:149-159) with the VV right-hand side; every stage
    adds `source` when given."""
    a = np.array([1. / 6., 1. / 3., 1. / 3., 1. / 6.], dtype=g.float)
    b = np.array([0.5, 0.5, 1.], dtype=g.float)
    dt = g.float(dt)
    w0 = w0.copy()
    w1 = w0.copy()
    w2 = w0.copy()
    for rk in range(4):
        rhs = compute_rhs(g, w0, nu, source)
        if rk < 3:
            w0[:] = w1 + b[rk] * dt * rhs
        w2 += a[rk] * dt * rhs
    w0[:] = w2
    return w0


def taylor_green(g):
    """W_hat = 1j*(K x U_hat) of the Taylor-Green field (tests/TG.py:29-35)."""
    return g.cross2(g.taylor_green())


def regression_values(g, w_hat):
    """k, w of the TG regression hook for VV: U from compute_velocity, curl =
    backward(W_hat) (tests/TG.py:116-126 with VV.py:75-89)."""
    U = g.backward(cross2_k_over_k2(g, w_hat))
    curl = g.backward(w_hat)
    n = np.prod(g.N)
    w = np.sum(curl.astype(np.float64) ** 2) / n / 2
    k = np.sum(U.astype(np.float64) ** 2) / n / 2
    return float(k), float(w)
