"""Pins of the Standard / Divergence / Skewed convection forms of the oracle
(oracle/convection.py), on the CPU: projected RHS equal to the Vortex form's
(the NS oracle) on the kept modes with the 2/3 rule and on every mode without
aliasing, equal to it on every mode with the 3/2 rule (padding makes the
quadratic products exact on the kept modes, so the forms differ only by the
gradient the projection removes), and the reference's Taylor-Green pins for
each form with the 2/3 and the 3/2 rule."""
import numpy as np
import pytest

from oracle import convection as OCV
from oracle.ns import Grid, dt_schedule, tg_regression_values

FORMS = ["Standard", "Divergence", "Skewed"]


@pytest.mark.parametrize("form", FORMS)
def test_forms_match_vortex_on_kept_modes(form):
    from spectraldns_amd.isotropic import random_solenoidal_field
    N = (16, 16, 16)
    g = Grid(N, (2 * np.pi,) * 3, "double", "2/3-rule")
    u = random_solenoidal_field(N, seed=3, k0=3.0, energy=0.5)
    r = OCV.compute_rhs(g, u, 0.01, form)
    v = g.compute_rhs(u, 0.01)
    kx, ky, kz = (np.abs(k) for k in g.kint)
    keep = (3 * kx[:, None, None] < N[0]) & (3 * ky[None, :, None] < N[1]) & (3 * kz[None, None, :] < N[2])
    err = np.linalg.norm((r - v)[:, keep]) / np.linalg.norm(v[:, keep])
    assert err < 1e-13, err


@pytest.mark.parametrize("form", FORMS)
def test_forms_match_vortex_without_aliasing(form):
    from spectraldns_amd.isotropic import random_solenoidal_field
    N = (16, 16, 16)
    g = Grid(N, (2 * np.pi,) * 3, "double", "None")
    u = random_solenoidal_field(N, seed=4, k0=2.0, energy=0.5)
    kx, ky, kz = (np.abs(k) for k in g.kint)
    low = (kx[:, None, None] < 4) & (ky[None, :, None] < 4) & (kz[None, None, :] < 4)
    u = u * low
    # keep it solenoidal after the cut: the cut is per mode, so k . u stays 0
    r = OCV.compute_rhs(g, u, 0.01, form)
    v = g.compute_rhs(u, 0.01)
    err = np.linalg.norm((r - v).ravel()) / np.linalg.norm(v.ravel())
    assert err < 1e-13, err


@pytest.mark.parametrize("form", FORMS)
def test_forms_match_vortex_32_rule(form):
    from spectraldns_amd.isotropic import random_solenoidal_field
    N = (16, 16, 16)
    g = Grid(N, (2 * np.pi,) * 3, "double", "3/2-rule")
    u = random_solenoidal_field(N, seed=5, k0=3.0, energy=0.5)
    r = OCV.compute_rhs(g, u, 0.01, form)
    v = g.compute_rhs(u, 0.01)
    err = np.linalg.norm((r - v).ravel()) / np.linalg.norm(v.ravel())
    assert err < 1e-13, err


@pytest.mark.parametrize("rule", ["2/3-rule", "3/2-rule"])
@pytest.mark.parametrize("form", FORMS)
def test_taylor_green_pins(form, rule):
    g = Grid((16, 16, 16), (2 * np.pi,) * 3, "double", rule)
    u = g.taylor_green()
    for dt in dt_schedule(0.1, 0.01):
        u = OCV.rk4_step(g, u, dt, 0.000625, form)
    k, w = tg_regression_values(g, u)
    assert round(w - 0.375249930801, 7) == 0, w
    assert round(k - 0.124953117517, 7) == 0, k
