"""Pins of the MHD oracle (oracle/mhd.py), on the CPU: the reference's
regression values k(0.1), b(0.1) of tests/TGMHD.py:16-25 at 7 decimals on the
two boxes of tests/test_MHD.py:15-30 (2/3 rule, nu = 0.000625, eta = 0.01),
the same pins with the 3/2 rule, and the b = 0 reduction to the NS oracle on
the modes the 2/3 rule keeps and, with the 3/2 rule, on every mode."""
import numpy as np
import pytest

from oracle import mhd as OM
from oracle.ns import Grid, dt_schedule

BOXES = [((4, 4, 4), (2 * np.pi,) * 3), ((6, 5, 4), (6 * np.pi, 4 * np.pi, 2 * np.pi))]


@pytest.mark.parametrize("rule", ["2/3-rule", "3/2-rule"])
@pytest.mark.parametrize("M,L", BOXES)
def test_mhd_taylor_green_pins(M, L, rule):
    g = Grid(tuple(2 ** m for m in M), L, "double", rule)
    ub = OM.taylor_green(g)
    for dt in dt_schedule(0.1, 0.01):
        ub = OM.rk4_step(g, ub, dt, 0.000625, 0.01)
    k, b = OM.regression_values(g, ub)
    assert round(k - 0.124565408177, 7) == 0, k
    assert round(b - 0.124637762143, 7) == 0, b


def test_mhd_without_field_is_ns():
    from spectraldns_amd.isotropic import random_solenoidal_field
    N = (16, 16, 16)
    g = Grid(N, (2 * np.pi,) * 3, "double", "2/3-rule")
    u = random_solenoidal_field(N, seed=4, k0=3.0, energy=0.5)
    ub = np.concatenate([u, np.zeros_like(u)])
    for _ in range(2):
        u = g.rk4_step(u, 0.01, 0.01)
        ub = OM.rk4_step(g, ub, 0.01, 0.01, 0.02)
    # the two convection forms differ by a gradient (removed by the projection) and in the
    # aliasing they leave on the modes the 2/3 rule drops from every backward transform;
    # the kept modes 3|k_i| < N_i evolve identically
    kx, ky, kz = (np.abs(k) for k in g.kint)
    keep = (3 * kx[:, None, None] < N[0]) & (3 * ky[None, :, None] < N[1]) & (3 * kz[None, None, :] < N[2])
    err = np.linalg.norm((ub[:3] - u)[:, keep]) / np.linalg.norm(u[:, keep])
    assert err < 1e-12, err
    assert np.max(np.abs(ub[3:])) == 0


def test_mhd_without_field_is_ns_32_rule():
    """With padding the Elsasser products u_i u_j are exact on the kept modes, so
    the divergence form equals the rotational one up to the gradient that the
    projection removes: every mode follows the NS oracle."""
    from spectraldns_amd.isotropic import random_solenoidal_field
    N = (16, 16, 16)
    g = Grid(N, (2 * np.pi,) * 3, "double", "3/2-rule")
    u = random_solenoidal_field(N, seed=5, k0=3.0, energy=0.5)
    ub = np.concatenate([u, np.zeros_like(u)])
    for _ in range(2):
        u = g.rk4_step(u, 0.01, 0.01)
        ub = OM.rk4_step(g, ub, 0.01, 0.01, 0.02)
    err = np.linalg.norm((ub[:3] - u).ravel()) / np.linalg.norm(u.ravel())
    assert err < 1e-13, err
    assert np.abs(ub[3:]).max() == 0.0
