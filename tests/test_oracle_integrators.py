"""Pins of the oracle's ForwardEuler, AB2 and BS5 (adaptiveRK) restatement
(oracle/integrators.py), on the CPU.

* Linear shear modes u = (f(y), 0, 0): u x curl(u) = grad(f^2/2) is removed by
  the projection, so the RHS is exactly -nu k^2 u_hat and each mode is
  multiplied per step by the method's stability function R(z), z = -nu k^2 dt:
  1 + z for ForwardEuler, the AB2 two-step recursion, and for BS5
  R(z) = 1 + z b^T (I - zA)^-1 1 from a linear solve (not stage by stage).
  BS5's R(z) must also match exp(z) to O(z^6) (order 5 of the tableau) and AB2
  must converge at second order.
* The reference's Taylor-Green regression values k(0.1) = 0.124953117517 and
  w(0.1) = 0.375249930801 (tests/TG.py:125-126) at ntol 4 for ForwardEuler
  and AB2 and 7 for BS5 (tests/test_NSVV.py:76-91), on the uniform 16^3 box and
  the non-uniform 64x32x16 box on [0,6pi)x[0,4pi)x[0,2pi).
"""
import numpy as np
import pytest

from oracle import integrators as OI
from oracle.ns import Grid, tg_regression_values


def shear_field(g):
    """u = (sum_m a_m sin(m y + p_m), 0, 0) on the grid, in spectral space."""
    y = g.X[1][None, :, None]
    U = np.zeros((3,) + g.N)
    for m, a, p in ((1, 1.0, 0.3), (2, 0.5, 1.1), (3, 0.25, -0.7)):
        U[0] += a * np.sin(m * y + p) * np.ones(g.N)
    return g.forward(U)


def stability_bs5(z):
    A, b = OI.BS5_A, OI.BS5_B
    s = A.shape[0]
    return 1 + z * b @ np.linalg.solve(np.eye(s) - z * A, np.ones(s))


@pytest.mark.parametrize("name", ["ForwardEuler", "AB2", "BS5_fixed"])
def test_shear_mode_amplification(name):
    g = Grid((16, 16, 16), precision="double", dealias="2/3-rule")
    nu, dt, n = 0.05, 0.02, 6
    u0 = shear_field(g)
    integ = OI.make(name, g, nu)
    u = u0.copy()
    for tstep in range(n):
        u, _, _ = integ.step(u, dt, tstep)
    z = -nu * g.K2 * dt
    if name == "ForwardEuler":
        expect = (1 + z) ** n * u0
    elif name == "BS5_fixed":
        R = np.vectorize(stability_bs5)(z)
        expect = R ** n * u0
    else:
        prev, cur = u0, (1 + z) * u0
        for _ in range(n - 1):
            prev, cur = cur, cur + 1.5 * z * cur - 0.5 * z * prev
        expect = cur
    err = np.linalg.norm((u - expect).ravel()) / np.linalg.norm(expect.ravel())
    assert err < 1e-13, err


def test_bs5_tableau_order_and_embedded_error():
    for z in (-0.02, -0.05, -0.1):
        assert abs(stability_bs5(z) - np.exp(z)) < 2e-3 * abs(z) ** 6
    A, b, bh = OI.BS5_A, OI.BS5_B, OI.BS5_BHAT
    s = A.shape[0]
    one = np.ones(s)
    # embedded method: b_hat is order 4, so (b - b_hat) R-difference is O(z^5), not smaller
    d = [abs(z * (b - bh) @ np.linalg.solve(np.eye(s) - z * A, one)) for z in (-0.1, -0.05)]
    assert 20 < d[0] / d[1] < 44  # 2^5 = 32
    assert abs(np.sum(b) - 1) < 1e-15 and abs(np.sum(bh) - 1) < 1e-15
    assert np.allclose(A[-1], b)  # first same as last


def test_ab2_second_order():
    g = Grid((16, 16, 16), precision="double", dealias="2/3-rule")
    nu, T = 0.05, 0.2
    u0 = shear_field(g)
    exact = np.exp(-nu * g.K2 * T) * u0
    errs = []
    for dt in (0.02, 0.01):
        integ = OI.make("AB2", g, nu)
        u = u0.copy()
        for tstep in range(int(round(T / dt))):
            u, _, _ = integ.step(u, dt, tstep)
        errs.append(np.linalg.norm((u - exact).ravel()))
    assert 3.5 < errs[0] / errs[1] < 4.5, errs


def tg_run(name, M, L):
    N = tuple(2 ** m for m in M)
    g = Grid(N, L, "double", "2/3-rule")
    integ = OI.make(name, g, 0.000625, tol=1e-6)
    u, t, tstep, taken = OI.solve(integ, g.taylor_green(), 0.1, 0.01)
    return g, u, t, taken


@pytest.mark.parametrize("name,ntol", [("ForwardEuler", 4), ("AB2", 4), ("BS5_fixed", 7), ("BS5_adaptive", 7)])
@pytest.mark.parametrize("M,L", [((4, 4, 4), (2 * np.pi,) * 3), ((6, 5, 4), (6 * np.pi, 4 * np.pi, 2 * np.pi))])
def test_taylor_green_pins(name, ntol, M, L):
    g, u, t, taken = tg_run(name, M, L)
    assert abs(t - 0.1) < 1e-12
    k, w = tg_regression_values(g, u)
    assert round(w - 0.375249930801, ntol) == 0, (w, taken)
    assert round(k - 0.124953117517, ntol) == 0, (k, taken)
    if name == "BS5_adaptive":
        assert len(taken) < 10  # the controller lengthened the steps
