"""GPU parity of the HIP path against the CPU oracle, through the C-ABI.

Bars: FP64 results within 1e-13 (one RHS) and 1e-12 (after N RK4 steps)
relative L2 of the oracle on the same inputs; FP32 within 1e-5 relative L2 of
the FP64 oracle (energy / dissipation reductions in FP64).
"""
import numpy as np
import pytest

from oracle.ns import Grid, dt_schedule

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
TWO_PI = 2 * np.pi


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def field(N, L=(TWO_PI,) * 3, seed=11, energy=0.5):
    from spectraldns_amd.isotropic import random_solenoidal_field
    return random_solenoidal_field(N, L, seed=seed, k0=min(N) / 8 + 1, energy=energy)


def to_gpu(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dtype).cuda()


def plan_of(N, L, precision, rule, mode=0, mask=True):
    from spectraldns_amd._abi import Plan
    print("\n[case] N=%s L=%s %s %s mode=%d mask=%s" % (N, tuple(np.round(L, 3)), precision, rule, mode, mask),
          flush=True)
    return Plan(N, L, precision, rule, mask, mode)


PLAIN_SIZES = [(8, 8, 8), (12, 16, 8), (16, 16, 16), (32, 24, 48), (64, 32, 16), (48, 48, 48),
               (72, 96, 64), (128, 128, 128), (256, 256, 256), (384, 192, 384)]


@pytest.mark.parametrize("N", PLAIN_SIZES)
@pytest.mark.parametrize("precision", ["double", "single"])
def test_plain_transforms(N, precision):
    rng = np.random.default_rng(0)
    L = (TWO_PI, 3.0, 5.0)
    p = plan_of(N, L, precision, "None")
    g = Grid(N, L, "double", "None")
    u = rng.standard_normal((3,) + N)
    cd = torch.complex128 if precision == "double" else torch.complex64
    rd = torch.float64 if precision == "double" else torch.float32
    ud = to_gpu(u, rd)
    uh = torch.empty((3,) + p.spectral_shape, dtype=cd, device="cuda")
    p.forward(ud, uh)
    ref = g.forward(u)
    tol = 1e-14 if precision == "double" else 2e-6
    assert rel(uh.cpu().numpy(), ref) < tol
    back = torch.empty_like(ud)
    p.backward(uh, back)
    assert rel(back.cpu().numpy(), u) < tol
    # inverse of a spectrum with non-Hermitian noise on kz = 0: imaginary DC/Nyquist dropped
    assert rel(g.backward(uh.cpu().numpy().astype(np.complex128)), back.cpu().numpy()) < tol


RHS_CASES = [
    ((8, 8, 8), (TWO_PI,) * 3, "2/3-rule", 0),
    ((16, 16, 16), (TWO_PI,) * 3, "2/3-rule", 0),
    ((16, 16, 16), (TWO_PI,) * 3, "2/3-rule", 1),
    ((16, 16, 16), (TWO_PI,) * 3, "3/2-rule", 0),
    ((16, 16, 16), (TWO_PI,) * 3, "None", 0),
    ((32, 32, 32), (TWO_PI,) * 3, "2/3-rule", 0),
    ((32, 32, 32), (TWO_PI,) * 3, "3/2-rule", 0),
    ((64, 32, 16), (6 * np.pi, 4 * np.pi, 2 * np.pi), "2/3-rule", 0),
    ((64, 32, 16), (6 * np.pi, 4 * np.pi, 2 * np.pi), "3/2-rule", 0),
    ((48, 48, 48), (TWO_PI,) * 3, "2/3-rule", 1),
    ((48, 32, 64), (TWO_PI, 3.0, 4.0), "3/2-rule", 0),
    ((64, 64, 64), (TWO_PI,) * 3, "2/3-rule", 0),
]


@pytest.mark.parametrize("N,L,rule,mode", RHS_CASES)
def test_compute_rhs_fp64(N, L, rule, mode):
    nu = 0.01
    p = plan_of(N, L, "double", rule, mode)
    g = Grid(N, L, "double", rule, True, mode)
    u = field(N, L)
    ud = to_gpu(u, torch.complex128)
    rhs = torch.empty_like(ud)
    ph = torch.empty(p.spectral_shape, dtype=torch.complex128, device="cuda")
    p.compute_rhs(ud, rhs, nu, p_hat=ph)
    ref = g.compute_rhs(u, nu)
    assert rel(rhs.cpu().numpy(), ref) < 1e-13
    conv = torch.empty_like(ud)
    p.convection(ud, conv)
    assert rel(conv.cpu().numpy(), g.conv(u)) < 1e-13
    _, P_ref = g.add_pressure_diffusion(g.mask_nyquist(g.conv(u)), u, nu)
    assert rel(ph.cpu().numpy(), P_ref) < 1e-12


@pytest.mark.parametrize("N,L,rule,mode", RHS_CASES[1:8])
def test_rk4_fp64_ten_steps(N, L, rule, mode):
    nu = 0.005
    p = plan_of(N, L, "double", rule, mode)
    g = Grid(N, L, "double", rule, True, mode)
    u = field(N, L, seed=3)
    ud = to_gpu(u, torch.complex128)
    dts = dt_schedule(0.1, 0.01)
    diag = torch.zeros(2 * len(dts), dtype=torch.float64, device="cuda")
    p.rk4_steps(ud, dts, nu, diag=diag)
    E = []
    for d in dts:
        u = g.rk4_step(u, d, nu)
        E.append((g.kinetic_energy(u), g.dissipation(u, nu)))
    assert rel(ud.cpu().numpy(), u) < 1e-12
    dg = diag.cpu().numpy().reshape(-1, 2)
    assert np.max(abs(dg - np.array(E)) / np.abs(np.array(E))) < 1e-12


@pytest.mark.parametrize("mask", [True, False])
@pytest.mark.parametrize("N,L", [((64, 48, 32), (TWO_PI, 3.0, 5.0)), ((96, 64, 32), (4 * np.pi, TWO_PI, 3.0))])
def test_rk4_32rule_stages(N, L, mask):
    """RK4 stages under the 3/2 rule on x lines of 96 and 144 points, with the
    run-time row maps (x-forward variant 0) and with the stage kernels whose
    kept x rows are fixed at compile time (variants 2 and 3): ten steps in
    FP64 within 1e-12 of the oracle with the per-step energy and dissipation,
    with and without the Nyquist mask, and FP32 within 1e-5 of the FP64
    oracle."""
    nu = 0.005
    g = Grid(N, L, "double", "3/2-rule", mask)
    u = field(N, L, seed=17)
    dts = dt_schedule(0.1, 0.01)
    got = {}
    for var in (0, 2, 3):
        p = plan_of(N, L, "double", "3/2-rule", mask=mask)
        p.set_variant(4, var)
        ud = to_gpu(u, torch.complex128)
        diag = torch.zeros(2 * len(dts), dtype=torch.float64, device="cuda")
        p.rk4_steps(ud, dts, nu, diag=diag)
        ps = plan_of(N, L, "single", "3/2-rule", mask=mask)
        ps.set_variant(4, var)
        us = to_gpu(u, torch.complex64)
        ps.rk4_steps(us, dts, np.float32(nu))
        got[var] = (ud.cpu().numpy(), diag.cpu().numpy().reshape(-1, 2), us.cpu().numpy().astype(np.complex128))
    E = []
    for d in dts:
        u = g.rk4_step(u, d, nu)
        E.append((g.kinetic_energy(u), g.dissipation(u, nu)))
    for var, (u64, dg, u32) in got.items():
        assert rel(u64, u) < 1e-12, var
        assert np.max(abs(dg - np.array(E)) / np.abs(np.array(E))) < 1e-12, var
        assert rel(u32, u) < 1e-5, var


@pytest.mark.parametrize("N,L,rule", [((32, 32, 32), (TWO_PI,) * 3, "2/3-rule"),
                                      ((32, 32, 32), (TWO_PI,) * 3, "3/2-rule"),
                                      ((64, 32, 16), (6 * np.pi, 4 * np.pi, 2 * np.pi), "2/3-rule")])
def test_rk4_fp32_vs_fp64_oracle(N, L, rule):
    nu = 0.005
    p = plan_of(N, L, "single", rule)
    g = Grid(N, L, "double", rule)
    u = field(N, L, seed=5)
    ud = to_gpu(u, torch.complex64)
    dts = dt_schedule(0.1, 0.01)
    p.rk4_steps(ud, dts, np.float32(nu))
    for d in dts:
        u = g.rk4_step(u, d, nu)
    err = rel(ud.cpu().numpy().astype(np.complex128), u)
    print("fp32 vs fp64 oracle rel-L2 after %d steps: %.3e" % (len(dts), err))
    assert err < 1e-5


def test_elementwise_kernels():
    N, L = (16, 12, 8), (TWO_PI, 3.0, 2.0)
    p = plan_of(N, L, "double", "2/3-rule")
    g = Grid(N, L, "double", "2/3-rule")
    rng = np.random.default_rng(2)
    a = rng.standard_normal((3,) + N)
    b = rng.standard_normal((3,) + N)
    c = torch.empty((3,) + N, dtype=torch.float64, device="cuda")
    p.cross1(c, to_gpu(a, torch.float64), to_gpu(b, torch.float64))
    assert rel(c.cpu().numpy(), g.cross1(a, b)) < 1e-15
    u = field(N, L)
    ud = to_gpu(u, torch.complex128)
    ch = torch.empty_like(ud)
    p.cross2(ch, ud)
    assert rel(ch.cpu().numpy(), g.cross2(u)) < 1e-15
    r = g.forward(rng.standard_normal((3,) + N))
    rd = to_gpu(r, torch.complex128)
    ph = torch.empty(p.spectral_shape, dtype=torch.complex128, device="cuda")
    p.add_pressure_diffusion(rd, ud, 0.03, ph)
    ref, pref = g.add_pressure_diffusion(r.copy(), u, 0.03)
    assert rel(rd.cpu().numpy(), ref) < 1e-14 and rel(ph.cpu().numpy(), pref) < 1e-14
    dv = torch.empty(p.spectral_shape, dtype=torch.complex128, device="cuda")
    p.divergence(dv, rd)
    assert np.max(abs(dv.cpu().numpy())) < 1e-13 * np.max(abs(r))
    out = torch.empty(2, dtype=torch.float64, device="cuda")
    p.energy(ud, 0.02, out)
    e, d = out.cpu().numpy()
    assert abs(e - g.kinetic_energy(u)) < 1e-14 * g.kinetic_energy(u)
    assert abs(d - g.dissipation(u, 0.02)) < 1e-13 * g.dissipation(u, 0.02)


@pytest.mark.parametrize("precision,rule", [("single", "3/2-rule"), ("single", "2/3-rule"),
                                            ("double", "3/2-rule")])
def test_256_cube_invariants(precision, rule):
    """Full-size properties the oracle cannot afford per step: the state stays
    divergence free with zero Nyquist planes, and the in-step FP64 energy
    reduction matches an independent reduction of the final state."""
    N, L = (256, 256, 256), (TWO_PI,) * 3
    nu = 0.001
    p = plan_of(N, L, precision, rule)
    cd = torch.complex64 if precision == "single" else torch.complex128
    ud = to_gpu(field(N, L, seed=1234, energy=0.5), cd)
    dts = [0.001] * 3
    diag = torch.zeros(6, dtype=torch.float64, device="cuda")
    p.rk4_steps(ud, dts, nu, diag=diag)
    out = torch.empty(2, dtype=torch.float64, device="cuda")
    p.energy(ud, nu, out)
    dg = diag.cpu().numpy().reshape(-1, 2)
    e = out.cpu().numpy()
    assert np.allclose(dg[-1], e, rtol=1e-12)
    assert 0.3 < dg[-1][0] < 0.5 and dg[0][0] > dg[1][0] > dg[2][0]   # decaying energy
    dv = torch.empty(p.spectral_shape, dtype=cd, device="cuda")
    p.divergence(dv, ud)
    umax = ud.abs().max().item()
    tol = 1e-5 if precision == "single" else 1e-12
    assert dv.abs().max().item() < tol * umax * 128
    h = ud.cpu().numpy()
    assert np.all(h[:, 128] == 0) and np.all(h[:, :, 128] == 0) and np.all(h[..., 128] == 0)


@pytest.mark.parametrize("N,rule", [((128, 64, 256), "2/3-rule"), ((128, 128, 128), "3/2-rule"),
                                    ((64, 32, 256), "3/2-rule"), ((64, 64, 64), "2/3-rule")])
def test_alternative_schedules_fp64(N, rule):
    """Every pass run with its alternative line-FFT schedule (variant 1) keeps
    FP64 parity with the oracle, for the RHS and the plain transforms (64^3:
    the 4-point schedules of the 64- and 32-point lines, on the small-grid
    tile divisors)."""
    L = (TWO_PI, 3.0, 5.0)
    nu = 0.01
    p = plan_of(N, L, "double", rule)
    for pid in range(6):
        p.set_variant(pid, 1)
    g = Grid(N, L, "double", rule)
    u = field(N, L, seed=21)
    ud = to_gpu(u, torch.complex128)
    rhs = torch.empty_like(ud)
    p.compute_rhs(ud, rhs, nu)
    assert rel(rhs.cpu().numpy(), g.compute_rhs(u, nu)) < 1e-13
    phys = torch.empty((3,) + N, dtype=torch.float64, device="cuda")
    p.backward(ud, phys)
    assert rel(phys.cpu().numpy(), g.backward(u)) < 1e-14
    back = torch.empty_like(ud)
    p.forward(phys, back)
    assert rel(back.cpu().numpy(), u) < 1e-14
    p.rk4_steps(ud, [0.01, 0.01], nu)
    assert rel(ud.cpu().numpy(), g.rk4_step(g.rk4_step(u, 0.01, nu), 0.01, nu)) < 1e-12


@pytest.mark.parametrize("N,rule", [((64, 48, 128), "2/3-rule"), ((96, 64, 128), "3/2-rule"),
                                    ((32, 32, 256), "3/2-rule"), ((64, 64, 96), "None")])
def test_z_pass_variants_fp64(N, rule):
    """Every storage scheme / schedule of the z pass (variants 0..5) keeps FP64
    parity with the oracle (variants 2/3 skip the bins that dealiasing zeroes,
    so their sums are grouped differently from variants 4/5 and agree with
    them to rounding, not bit for bit)."""
    L = (TWO_PI, 4.0, 3.0)
    nu = 0.01
    g = Grid(N, L, "double", rule)
    u = field(N, L, seed=23)
    ref = g.compute_rhs(u, nu)
    out = {}
    for var in range(6):
        p = plan_of(N, L, "double", rule)
        p.set_variant(2, var)
        ud = to_gpu(u, torch.complex128)
        rhs = torch.empty_like(ud)
        p.compute_rhs(ud, rhs, nu)
        out[var] = rhs.cpu().numpy()
        assert rel(out[var], ref) < 1e-13, var
    assert rel(out[2], out[4]) < 1e-14 and rel(out[3], out[5]) < 1e-14


@pytest.mark.parametrize("N", [(1024, 32, 64), (32, 64, 1024)])
def test_long_lines_fp64(N):
    """1024-point x lines and 512-point half z lines (the line lengths of a
    1024^3 box), 2/3 rule: RHS and two RK4 steps against the oracle, and every
    z-pass variant for the z case."""
    L = (TWO_PI,) * 3
    nu = 0.005
    g = Grid(N, L, "double", "2/3-rule")
    u = field(N, L, seed=29)
    ref = g.compute_rhs(u, nu)
    for var in (range(6) if N[2] == 1024 else (0,)):
        p = plan_of(N, L, "double", "2/3-rule")
        p.set_variant(2, var)
        ud = to_gpu(u, torch.complex128)
        rhs = torch.empty_like(ud)
        p.compute_rhs(ud, rhs, nu)
        assert rel(rhs.cpu().numpy(), ref) < 1e-13, var
    p = plan_of(N, L, "double", "2/3-rule")
    ud = to_gpu(u, torch.complex128)
    p.rk4_steps(ud, [0.002, 0.002], nu)
    assert rel(ud.cpu().numpy(), g.rk4_step(g.rk4_step(u, 0.002, nu), 0.002, nu)) < 1e-12
    # single precision through the same lines
    p = plan_of(N, L, "single", "2/3-rule")
    uf = to_gpu(u, torch.complex64)
    rhs = torch.empty_like(uf)
    p.compute_rhs(uf, rhs, nu)
    assert rel(rhs.cpu().numpy(), ref) < 1e-5


def test_compensated_fp32_accumulation():
    """Many small steps: plain FP32 accumulation of u_hat drifts by about one
    rounding per step, the compensated (hi, lo) pair stays at FP32 resolution
    of the increments; both are compared with the FP64 oracle."""
    N, L = (32, 32, 32), (TWO_PI,) * 3
    nu, dt, nsteps = 0.01, 2e-4, 100
    g = Grid(N, L, "double", "2/3-rule")
    u = field(N, L, seed=9)
    ref = u.copy()
    for _ in range(nsteps):
        ref = g.rk4_step(ref, dt, nu)
    p = plan_of(N, L, "single", "2/3-rule")
    plain = to_gpu(u, torch.complex64)
    p.rk4_steps(plain, [dt] * nsteps, nu)
    hi = to_gpu(u, torch.complex64)
    lo = to_gpu(u - hi.cpu().numpy().astype(np.complex128), torch.complex64)
    diag = torch.zeros(2 * nsteps, dtype=torch.float64, device="cuda")
    p.rk4_steps(hi, [dt] * nsteps, nu, diag=diag, u_lo=lo)
    e_plain = rel(plain.cpu().numpy().astype(np.complex128), ref)
    e_comp = rel(hi.cpu().numpy().astype(np.complex128) + lo.cpu().numpy(), ref)
    print("after %d steps: plain fp32 %.3e, compensated %.3e" % (nsteps, e_plain, e_comp))
    assert e_comp < 0.5 * e_plain and e_comp < 2e-6
    assert abs(diag[-2].item() - g.kinetic_energy(ref)) < 1e-6 * g.kinetic_energy(ref)


@pytest.mark.parametrize("N,rule", [((256, 16, 32), "2/3-rule"), ((256, 16, 16), "3/2-rule")])
def test_x_forward_rk_variants(N, rule):
    """Every variant of the x-forward + RK pass on 256- and 384-point x lines
    (the headline boxes' line lengths): both schedules, and under the 3/2 rule
    both the run-time row maps (variants 0, 1) and the stage kernels with the
    kept rows fixed at compile time (2, 3).  Two RK4 steps within 1e-12 of
    the oracle in FP64 with the in-step energies, the two row handlings of a
    schedule within 1e-15 of each other, and FP32, plain and compensated,
    within 1e-5 of the FP64 oracle."""
    L = (TWO_PI, 3.0, 5.0)
    nu, dt = 0.01, 0.002
    g = Grid(N, L, "double", rule)
    u = field(N, L, seed=31)
    ref = g.rk4_step(g.rk4_step(u, dt, nu), dt, nu)
    got64 = []
    for var in range(4):
        p = plan_of(N, L, "double", rule)
        p.set_variant(4, var)
        ud = to_gpu(u, torch.complex128)
        diag = torch.zeros(4, dtype=torch.float64, device="cuda")
        p.rk4_steps(ud, [dt, dt], nu, diag=diag)
        got64.append(ud.cpu().numpy())
        assert rel(got64[-1], ref) < 1e-12, var
        assert abs(diag[2].item() - g.kinetic_energy(ref)) < 1e-13 * g.kinetic_energy(ref), var
        if var >= 2:
            assert rel(got64[var], got64[var - 2]) < 1e-15, var
        for comp in (False, True):
            p = plan_of(N, L, "single", rule)
            p.set_variant(4, var)
            us = to_gpu(u, torch.complex64)
            lo = torch.zeros_like(us) if comp else None
            p.rk4_steps(us, [dt, dt], nu, u_lo=lo)
            got = us.cpu().numpy().astype(np.complex128)
            if comp:
                got = got + lo.cpu().numpy().astype(np.complex128)
            assert rel(got, ref) < 1e-5, (var, comp)
