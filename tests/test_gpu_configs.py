"""The benchmark configurations as tests, on the GPU through the C-ABI.

  * Taylor-Green 32^3 FP64, 2/3 rule, 100 RK4 steps (T = 1.0): u_hat, kinetic
    energy and dissipation after every step against the oracle (FP64 1e-12
    relative L2 in u_hat, 1e-12 relative in E and eps).
  * Decaying isotropic turbulence 256^3, 3/2 rule (384^3 padded), nu = 0.001,
    dt = 0.001, the seeded field of the benchmark: two RK4 steps in FP64
    (<= 1e-12 relative L2 vs the FP64 oracle) and FP32 (<= 1e-5 vs the FP64
    oracle), with the in-step FP64 energies.
  * Forced isotropic turbulence 1024^3, FP32, 2/3 rule, compensated u_hat,
    forcing at |k| <= 2 after every step (nu = 0.0002, dt = 0.0005): three
    steps; the reference's energy-hold check after forcing (Isotropic.py:184),
    zero divergence and Nyquist planes, and the in-step FP64 energy equal to
    an independent FP64 reduction of the compensated state.
"""
import time

import numpy as np
import pytest

from oracle.ns import Grid, dt_schedule

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
TWO_PI = 2 * np.pi


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def energy_pair(hi, lo=None):
    """0.5 * sum w |hi + lo|^2 in FP64, one component at a time (torch on the GPU)."""
    Nh = hi.shape[-1]
    w = torch.full((Nh,), 2.0, dtype=torch.float64, device=hi.device)
    w[0] = 1.0
    w[-1] = 1.0
    s = 0.0
    for c in range(hi.shape[0]):
        v = hi[c].to(torch.complex128)
        if lo is not None:
            v = v + lo[c].to(torch.complex128)
        s += float(torch.sum((v.real * v.real + v.imag * v.imag) * w).item())
    return 0.5 * s


def test_taylor_green_32_fp64_100_steps():
    from spectraldns_amd._abi import Plan
    N, L, nu = (32, 32, 32), (TWO_PI,) * 3, 0.000625
    print("\n[case] TG 32^3 FP64 2/3-rule, 100 steps", flush=True)
    g = Grid(N, L, "double", "2/3-rule")
    p = Plan(N, L, "double", "2/3-rule")
    ref = g.taylor_green()
    ud = torch.from_numpy(ref).cuda()
    dts = dt_schedule(1.0, 0.01)
    assert len(dts) == 100
    diag = torch.zeros(2, dtype=torch.float64, device="cuda")
    worst = 0.0
    for s, dt in enumerate(dts):
        ref = g.rk4_step(ref, dt, nu)
        p.rk4_steps(ud, [dt], nu, diag=diag)
        e, eps = diag.cpu().numpy()
        err = rel(ud.cpu().numpy(), ref)
        worst = max(worst, err)
        assert err < 1e-12, (s, err)
        ek, epsk = g.kinetic_energy(ref), g.dissipation(ref, nu)
        assert abs(e - ek) < 1e-12 * ek, (s, e, ek)
        assert abs(eps - epsk) < 1e-12 * epsk, (s, eps, epsk)
    print("worst per-step rel-L2 %.3e; E(1.0) = %.12f" % (worst, e))


@pytest.mark.timeout(600)
def test_isotropic_256_32rule_two_steps():
    from spectraldns_amd._abi import Plan
    from spectraldns_amd.isotropic import random_solenoidal_field
    N, L, nu, dt = (256, 256, 256), (TWO_PI,) * 3, 0.001, 0.001
    print("\n[case] isotropic 256^3 3/2-rule, 2 steps, oracle FP64", flush=True)
    t0 = time.time()
    u = random_solenoidal_field(N, L, seed=1234, k0=4.0, energy=0.5)
    g = Grid(N, L, "double", "3/2-rule")
    ref = g.rk4_step(g.rk4_step(u, dt, nu), dt, nu)
    print("oracle: %.1f s" % (time.time() - t0), flush=True)
    e_ref = g.kinetic_energy(ref)
    for precision, tol in (("double", 1e-12), ("single", 1e-5)):
        p = Plan(N, L, precision, "3/2-rule")
        cd = torch.complex128 if precision == "double" else torch.complex64
        ud = torch.from_numpy(u).to(cd).cuda()
        diag = torch.zeros(4, dtype=torch.float64, device="cuda")
        p.rk4_steps(ud, [dt, dt], nu, diag=diag)
        err = rel(ud.cpu().numpy().astype(np.complex128), ref)
        e = diag.cpu().numpy()[2]
        print("%s: rel-L2 vs FP64 oracle %.3e, energy %.15f (oracle %.15f)" % (precision, err, e, e_ref),
              flush=True)
        assert err < tol, (precision, err)
        assert abs(e - e_ref) < (1e-13 if precision == "double" else 1e-6) * e_ref
        del ud, p
        torch.cuda.empty_cache()


@pytest.mark.timeout(300)
def test_forced_1024_fp32_compensated():
    from spectraldns_amd._abi import Plan
    from spectraldns_amd.isotropic import random_solenoidal_field_gpu
    N, L = (1024, 1024, 1024), (TWO_PI,) * 3
    nu, dt, kf = 0.0002, 0.0005, 2.0
    print("\n[case] forced 1024^3 FP32 2/3-rule compensated, 3 steps", flush=True)
    p = Plan(N, L, "single", "2/3-rule")
    u = random_solenoidal_field_gpu(p, seed=1234, k0=4.0, energy=0.5)
    lo = torch.zeros_like(u)
    target = 2.0 * energy_pair(u)
    diag = torch.zeros(2, dtype=torch.float64, device="cuda")
    sums = torch.zeros(2, dtype=torch.float64, device="cuda")
    e_prev = target / 2
    for s in range(3):
        p.rk4_steps(u, [dt], nu, diag=diag, u_lo=lo)
        e_step, eps = diag.cpu().numpy()
        e_ind = energy_pair(u, lo)
        print("step %d: in-step E %.15f, independent %.15f, eps %.6e" % (s, e_step, e_ind, eps), flush=True)
        assert abs(e_step - e_ind) < 1e-12 * e_ind
        assert e_step < e_prev and eps > 0          # viscous decay within the step
        p.force(u, target, kf, sums=sums, u_lo=lo)
        e_forced = 2.0 * energy_pair(u, lo)
        assert abs(e_forced - target) < 1e-7 * target, (e_forced, target)   # Isotropic.py:184
        e_prev = e_forced / 2
    h = u[:, :, :, 512]
    assert torch.count_nonzero(h).item() == 0
    assert torch.count_nonzero(u[:, 512]).item() == 0 and torch.count_nonzero(u[:, :, 512]).item() == 0
    dv = torch.empty(p.spectral_shape, dtype=torch.complex64, device="cuda")
    p.divergence(dv, u)
    assert dv.abs().max().item() < 1e-5 * u.abs().max().item() * 512
