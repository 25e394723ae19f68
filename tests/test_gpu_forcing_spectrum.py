"""GPU parity of the forcing (demo/Isotropic.py:154-184) and of the shell
spectrum (demo/Isotropic.py:88-115) against the CPU oracle, through the C-ABI.

Bars: FP64 within 1e-13 relative of the oracle (forced field, the two FP64
reductions, every shell of the spectrum) and exact shell mode counts; the
reference's own energy-hold check |E - E_target| < 1e-7 (Isotropic.py:184)
in FP64, FP32 and for the compensated FP32 pair.
"""
import numpy as np
import pytest

from oracle.ns import Grid

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
TWO_PI = 2 * np.pi

BOXES = [((32, 32, 32), (TWO_PI,) * 3), ((64, 32, 48), (TWO_PI, 2 * TWO_PI, TWO_PI))]


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def field(N, L, seed):
    from spectraldns_amd.isotropic import random_solenoidal_field
    return random_solenoidal_field(N, L, seed=seed, k0=3.0, energy=0.5)


def plan_of(N, L, precision):
    from spectraldns_amd._abi import Plan
    print("\n[case] N=%s L=%s %s" % (N, tuple(np.round(L, 3)), precision), flush=True)
    return Plan(N, L, precision, "2/3-rule")


def shell_counts(g, nb):
    """Modes per shell z = 1..nb (z - 1/2 < |K| <= z + 1/2), counted with NumPy."""
    k = np.sqrt(g.K2.astype(np.float64))
    return np.array([np.count_nonzero((k > z - 0.5) & (k <= z + 0.5)) for z in range(1, nb + 1)], dtype=float)


@pytest.mark.parametrize("N,L", BOXES)
def test_force_fp64(N, L):
    g = Grid(N, L, "double", "2/3-rule")
    u = field(N, L, seed=31)
    u[:, 0, 0, 0] = 0.01 * (1 + 1j)  # a mean mode the forcing must remove
    kf = 2.0
    target = 1.1 * g.energy_fourier(u)
    ref = g.force(u, target, kf)
    p = plan_of(N, L, "double")
    ud = torch.from_numpy(u).cuda()
    sums = torch.zeros(2, dtype=torch.float64, device="cuda")
    p.force(ud, target, kf, sums=sums)
    out = ud.cpu().numpy()
    assert rel(out, ref) < 1e-13
    u0 = u.copy()
    u0[:, 0, 0, 0] = 0
    low = g.energy_fourier(u0 * (g.K2 <= kf ** 2))
    s = sums.cpu().numpy()
    assert abs(s[0] - g.energy_fourier(u0)) < 1e-13 * s[0] and abs(s[1] - low) < 1e-13 * low
    assert abs(g.energy_fourier(out) - target) < 1e-7           # Isotropic.py:184
    assert np.all(out[:, 0, 0, 0] == 0)
    hi = g.K2 > kf ** 2
    assert np.array_equal(out[:, hi], u[:, hi])                 # only |k| <= kf is rescaled


@pytest.mark.parametrize("N,L", BOXES)
def test_force_fp32_and_compensated(N, L):
    g = Grid(N, L, "double", "2/3-rule")
    u = field(N, L, seed=32)
    kf = 2.0
    target = 1.05 * g.energy_fourier(u)
    ref = g.force(u, target, kf)
    p = plan_of(N, L, "single")
    uf = torch.from_numpy(u).to(torch.complex64).cuda()
    hi = uf.clone()
    lo = torch.from_numpy(u - uf.cpu().numpy().astype(np.complex128)).to(torch.complex64).cuda()
    p.force(uf, target, kf)
    f = uf.cpu().numpy().astype(np.complex128)
    assert rel(f, ref) < 1e-6
    assert abs(g.energy_fourier(f) - target) < 1e-7
    # the compensated pair hi + lo carries u to ~1e-14 and is forced as one value
    p.force(hi, target, kf, u_lo=lo)
    pair = hi.cpu().numpy().astype(np.complex128) + lo.cpu().numpy().astype(np.complex128)
    assert rel(pair, ref) < 1e-12
    assert abs(g.energy_fourier(pair) - target) < 1e-12 * target


@pytest.mark.parametrize("N,L", BOXES)
@pytest.mark.parametrize("precision", ["double", "single"])
def test_spectrum(N, L, precision):
    from spectraldns_amd.fourier import FourierSpace, spectrum
    g = Grid(N, L, "double", "2/3-rule")
    u = field(N, L, seed=33)
    Ek_ref, bins_ref = g.spectrum(u)
    p = plan_of(N, L, precision)
    cd = torch.complex128 if precision == "double" else torch.complex64
    ud = torch.from_numpy(u).to(cd).cuda()
    nb = len(bins_ref) - 1
    sums, counts = p.spectrum_sums(ud, nb)
    assert np.array_equal(counts.cpu().numpy(), shell_counts(g, nb))
    Ek, bins = spectrum(ud, FourierSpace(p, 1))
    assert np.array_equal(bins, bins_ref)
    tol = 1e-13 if precision == "double" else 2e-7
    assert np.max(np.abs(Ek - Ek_ref)) <= tol * np.max(np.abs(Ek_ref))
    assert np.count_nonzero(Ek_ref) > 5
