"""Pins of the velocity-vorticity oracle (oracle/vv.py), on the CPU: the
reference's Taylor-Green regression values for the VV solver
(tests/test_NSVV.py, tests/TG.py:116-126) at ntol 7 on the uniform and the
non-uniform box with both dealiasing rules, and the identity
w_hat_VV(t) = 1j*(K x u_hat_NS(t)) against the separately pinned NS oracle."""
import numpy as np
import pytest

from oracle import vv as OV
from oracle.ns import Grid, dt_schedule

BOXES = [((4, 4, 4), (2 * np.pi,) * 3), ((6, 5, 4), (6 * np.pi, 4 * np.pi, 2 * np.pi))]


@pytest.mark.parametrize("dealias", ["2/3-rule", "3/2-rule"])
@pytest.mark.parametrize("M,L", BOXES)
def test_vv_taylor_green_pins(M, L, dealias):
    g = Grid(tuple(2 ** m for m in M), L, "double", dealias)
    w = OV.taylor_green(g)
    for dt in dt_schedule(0.1, 0.01):
        w = OV.rk4_step(g, w, dt, 0.000625)
    k, ww = OV.regression_values(g, w)
    assert round(ww - 0.375249930801, 7) == 0, ww
    assert round(k - 0.124953117517, 7) == 0, k


def test_vv_is_the_curl_of_ns():
    from spectraldns_amd.isotropic import random_solenoidal_field
    N = (16, 16, 16)
    g = Grid(N, (2 * np.pi,) * 3, "double", "2/3-rule")
    u = random_solenoidal_field(N, seed=3, k0=3.0, energy=0.5)
    w = g.cross2(u)
    for _ in range(3):
        u = g.rk4_step(u, 0.01, 0.01)
        w = OV.rk4_step(g, w, 0.01, 0.01)
    expect = g.cross2(u)
    err = np.linalg.norm((w - expect).ravel()) / np.linalg.norm(expect.ravel())
    assert err < 1e-13, err
    # the VV right-hand side is the curl of the NS right-hand side (pressure gradient removed)
    assert np.linalg.norm((OV.compute_rhs(g, w, 0.01) - g.cross2(g.compute_rhs(u, 0.01))).ravel()) < \
        1e-12 * np.linalg.norm(expect.ravel())


def test_vv_mean_vorticity_enters_the_convection():
    """A constant vorticity w0 (the mean mode) adds u x w0 to the product
    while leaving the velocity 1j*(K/K^2) x w_hat unchanged (K = 0 there):
    F(u x (w + w0)) - F(u x w) = u_hat x w0 exactly on the modes of u_hat
    (a constant factor creates no new modes, so nothing aliases).  The
    difference of the two right-hand sides is therefore 1j*K x (u_hat x w0)
    on every mode, in closed form."""
    from spectraldns_amd.isotropic import random_solenoidal_field
    N = (12, 12, 12)
    for rule in ("2/3-rule", "3/2-rule"):
        g = Grid(N, (2 * np.pi,) * 3, "double", rule)
        u = random_solenoidal_field(N, seed=5, k0=2.0, energy=0.5)
        u = u * (np.abs(g.K[0]) < 4) * (np.abs(g.K[1]) < 4) * (g.K[2] < 4)  # inside every keep set
        w = g.cross2(u)
        w0 = np.array([0.3, -0.2, 0.5])
        wm = w.copy()
        wm[:, 0, 0, 0] += w0
        d = OV.compute_rhs(g, wm, 0.0) - OV.compute_rhs(g, w, 0.0)
        uxw0 = np.empty_like(u)
        uxw0[0] = u[1] * w0[2] - u[2] * w0[1]
        uxw0[1] = u[2] * w0[0] - u[0] * w0[2]
        uxw0[2] = u[0] * w0[1] - u[1] * w0[0]
        expect = g.mask_nyquist(g.cross2(uxw0))
        assert np.max(np.abs(d - expect)) < 1e-13 * np.max(np.abs(expect)), rule
