"""A Source term and a given vorticity on the GPU against the oracles.

  * NS: the Source filled by a HIP launch through its device pointer (the
    plan's forward transform, as demo/Vortices.py:17-26 fills it with
    U.forward(Source)) is added to every RK4 stage, through the fused step
    (sdns_rk4_steps_source) and through the reference-shaped solver API
    (get_solver, getintegrator, ComputeRHS), FP64 within 1e-12 of the oracle's
    RK4 with that Source.
  * VV: the vorticity is used as given (VV.py:95-97), also when it is not a
    curl and has a mean (the random start of demo/Vortices.py:5-15), with a
    Source with a mean: FP64 within 1e-12 of the VV oracle after RK4 steps.
"""
import numpy as np
import pytest

from oracle import vv as OV
from oracle.ns import Grid, dt_schedule

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
TWO_PI = 2 * np.pi


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def physical_source(g, seed):
    """A smooth physical forcing with a mean, like demo/Vortices.py:17-26."""
    X = np.meshgrid(*g.X, indexing="ij")
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.5, 1.5, 3)
    f = np.empty((3,) + g.N)
    f[0] = a[0] * np.exp(-((X[1] - np.pi) ** 2 + (X[2] - np.pi) ** 2) / 0.5) + 0.1
    f[1] = a[1] * np.sin(X[0]) * np.cos(2 * X[2])
    f[2] = a[2] * np.cos(X[1]) - 0.05
    return f


@pytest.mark.parametrize("rule", ["2/3-rule", "3/2-rule"])
def test_ns_rk4_with_source_fp64(rule):
    from spectraldns_amd._abi import Plan
    from spectraldns_amd.isotropic import random_solenoidal_field
    N, L, nu, dt = (32, 32, 32), (TWO_PI,) * 3, 0.01, 0.005
    print("\n[case] NS RK4 + Source %s" % rule, flush=True)
    g = Grid(N, L, "double", rule)
    p = Plan(N, L, "double", rule)
    u = random_solenoidal_field(N, L, seed=41, k0=3.0, energy=0.5)
    f = physical_source(g, 41)
    src = torch.empty((3,) + p.spectral_shape, dtype=torch.complex128, device="cuda")
    p.forward(torch.from_numpy(f).cuda(), src)   # written by a HIP kernel through data_ptr
    s_np = g.forward(f)
    assert rel(src.cpu().numpy(), s_np) < 1e-14
    rhs = torch.empty_like(src)
    ud = torch.from_numpy(u).cuda()
    p.compute_rhs(ud, rhs, nu, source=src)
    assert rel(rhs.cpu().numpy(), g.compute_rhs(u, nu, s_np)) < 1e-13
    ref = u
    for _ in range(3):
        ref = g.rk4_step(ref, dt, nu, s_np)
    p.rk4_steps(ud, [dt] * 3, nu, source=src)
    assert rel(ud.cpu().numpy(), ref) < 1e-12
    # without the Source the result differs at O(dt |S|): the Source is not optional
    assert rel(g.rk4_step(u, dt, nu), g.rk4_step(u, dt, nu, s_np)) > 1e-4


def test_ns_solver_api_source_fp64():
    """get_solver / getintegrator / solve with a Source filled by
    VT.forward (a HIP launch) in the set-up, as demo/Vortices.py does."""
    from spectraldns_amd import config, get_solver, solve
    from spectraldns_amd.isotropic import random_solenoidal_field
    print("\n[case] NS solver API RK4 + Source", flush=True)
    config.update({"nu": 0.01, "dt": 0.005, "T": 0.015, "convection": "Vortex", "dealias": "2/3-rule",
                   "precision": "double", "integrator": "RK4"})
    solver = get_solver(parse_args=["--M", "5", "5", "5", "NS"])
    c = solver.get_context()
    N, L = (32, 32, 32), (TWO_PI,) * 3
    g = Grid(N, L, "double", "2/3-rule")
    u = random_solenoidal_field(N, L, seed=43, k0=3.0, energy=0.5)
    c.U_hat.copy_(torch.from_numpy(u))
    f = physical_source(g, 43)
    c.U.copy_(torch.from_numpy(f))
    c.Source = c.VT.forward(c.U, c.Source)
    config.params.t = 0.0
    config.params.tstep = 0
    solve(solver, c)
    ref = u
    s_np = g.forward(f)
    for dt in dt_schedule(0.015, 0.005):
        ref = g.rk4_step(ref, dt, 0.01, s_np)
    assert config.params.tstep == 3
    assert rel(c.U_hat.cpu().numpy(), ref) < 1e-12
    # the same with a stage-by-stage integrator (ComputeRHS per stage)
    config.update({"integrator": "ForwardEuler", "T": 0.01})
    solver = get_solver(parse_args=["--M", "5", "5", "5", "NS"])
    c = solver.get_context()
    c.U_hat.copy_(torch.from_numpy(u))
    c.U.copy_(torch.from_numpy(f))
    c.Source = c.VT.forward(c.U, c.Source)
    config.params.t = 0.0
    config.params.tstep = 0
    solve(solver, c)
    ref = u
    for dt in dt_schedule(0.01, 0.005):
        ref = ref + dt * g.compute_rhs(ref, 0.01, s_np)
    assert rel(c.U_hat.cpu().numpy(), ref) < 1e-12
    config.update({"integrator": "RK4"})


@pytest.mark.parametrize("rule", ["2/3-rule", "3/2-rule"])
def test_vv_given_vorticity_with_mean_fp64(rule):
    """W_hat not a curl (random physical vorticity with a mean) and a Source
    with a mean: the RHS and two RK4 steps of the VV solver against the VV
    oracle, which takes w_dealias = VTp.backward(w_hat) literally."""
    from spectraldns_amd import config, get_solver, solve
    print("\n[case] VV random vorticity + Source %s" % rule, flush=True)
    N, L, nu, dt = (32, 32, 32), (TWO_PI,) * 3, 0.01, 0.002
    config.update({"nu": nu, "dt": dt, "T": 2 * dt, "convection": "Vortex", "dealias": rule,
                   "precision": "double", "integrator": "RK4"})
    solver = get_solver(parse_args=["--M", "5", "5", "5", "VV"])
    c = solver.get_context()
    g = Grid(N, L, "double", rule)
    rng = np.random.default_rng(45)
    curl = 0.3 * rng.standard_normal((3,) + N)
    w = g.forward(curl)
    w = g.mask_nyquist(w)
    assert np.max(np.abs(w[:, 0, 0, 0])) > 1e-3          # a mean vorticity
    div = g.K[0] * w[0] + g.K[1] * w[1] + g.K[2] * w[2]
    assert np.max(np.abs(div)) > 1e-2                     # not a curl
    f = physical_source(g, 45)
    s_np = g.forward(f)
    c.W_hat.copy_(torch.from_numpy(w))
    c.U.copy_(torch.from_numpy(f))
    c.Source = c.VT.forward(c.U, c.Source)
    rhs = torch.empty_like(c.W_hat)
    solver.conv = solver.getConvection("Vortex")   # as solve() installs it
    solver.ComputeRHS(rhs, c.W_hat, solver, **c)
    assert rel(rhs.cpu().numpy(), OV.compute_rhs(g, w, nu, s_np)) < 1e-13
    config.params.t = 0.0
    config.params.tstep = 0
    solve(solver, c)
    ref = w
    for d in dt_schedule(2 * dt, dt):
        ref = OV.rk4_step(g, ref, d, nu, s_np)
    assert rel(c.W_hat.cpu().numpy(), ref) < 1e-12
