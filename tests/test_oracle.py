"""Pins of the CPU oracle (oracle/ns.py) against what the reference and the
mathematics fix, chosen so that a plausible slip (a sign, a dropped term, a
wrong index or coefficient) fails one of them.

  * TG pins of the reference (tests/TG.py:125-126) -- end to end.
  * brute-force DFT sums -- transforms, normalisation, wavenumber order.
  * shear mode closed forms -- curl sign (cross2), projection, diffusion sign.
  * ABC (Beltrami) field -- cross1/curl consistency and the RK4 coefficients.
  * triad sums -- the dealiasing keep sets and padding.
  * Parseval, TG E(0) = 1/8, Omega(0) = 3/8 -- diagnostics weights.
  * dt sequence of the solve loop, including the last-step adjustment.
"""
import numpy as np
import pytest

from oracle.ns import Grid, dt_schedule, keep_ranges, solve, tg_regression_values

TG_K, TG_W = 0.124953117517, 0.375249930801
NU, DT = 0.000625, 0.01
TWO_PI = 2 * np.pi


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


# ------------------------------------------------------------------ end to end
@pytest.mark.parametrize("N,L", [((16, 16, 16), (TWO_PI,) * 3),
                                 ((32, 32, 32), (TWO_PI,) * 3),
                                 ((64, 32, 16), (6 * np.pi, 4 * np.pi, 2 * np.pi))])
@pytest.mark.parametrize("rule,mode", [("2/3-rule", 0), ("2/3-rule", 1), ("3/2-rule", 0), ("None", 0)])
def test_taylor_green_pins(N, L, rule, mode):
    g = Grid(N, L, "double", rule, True, mode)
    u = solve(g, g.taylor_green(), 0.1, DT, NU)
    k, w = tg_regression_values(g, u)
    assert round(k - TG_K, 7) == 0, k
    assert round(w - TG_W, 7) == 0, w


def test_taylor_green_pins_single():
    g = Grid((32, 32, 32), (TWO_PI,) * 3, "single", "2/3-rule", True, 0)
    u = solve(g, g.taylor_green(), 0.1, DT, NU)
    assert u.dtype == np.complex64
    k, w = tg_regression_values(g, u)
    assert round(k - TG_K, 5) == 0 and round(w - TG_W, 5) == 0


# ------------------------------------------------------------------ transforms
@pytest.mark.parametrize("N", [(4, 6, 8), (6, 4, 4), (8, 8, 6)])
def test_forward_brute_force_dft(N):
    rng = np.random.default_rng(1)
    L = (TWO_PI, 4.0, 3.0)
    g = Grid(N, L, "double", "None")
    u = rng.standard_normal(N)
    uh = g.forward(u)
    x = [np.arange(n) for n in N]
    for (i, j, k) in [(0, 0, 0), (1, 2, 1), (N[0] - 1, 1, N[2] // 2), (N[0] // 2, N[1] - 1, 2)]:
        # mode numbers from the storage index (fftfreq on axes 0, 1; rfftfreq on axis 2)
        m = [g.kint[0][i], g.kint[1][j], g.kint[2][k]]
        ph = np.exp(-2j * np.pi * (m[0] * x[0][:, None, None] / N[0] + m[1] * x[1][None, :, None] / N[1]
                                   + m[2] * x[2][None, None, :] / N[2]))
        ref = np.sum(u * ph) / np.prod(N)
        assert abs(uh[i, j, k] - ref) < 1e-14
    assert rel(g.backward(uh), u) < 1e-14


def test_wavenumbers_scaled_by_domain():
    g = Grid((8, 8, 8), (6 * np.pi, 4 * np.pi, 2 * np.pi))
    assert np.allclose(g.k1d[0][:4], np.arange(4) / 3.0)
    assert g.k1d[0][4] == -4 / 3.0 and g.k1d[2][4] == 4.0  # Nyquist -N/2 on C2C, +N/2 on R2C


# ------------------------------------------------------------------ closed forms
def _shear(g, m):
    X = np.meshgrid(*g.X, indexing="ij")
    U = np.zeros((3,) + g.N)
    U[0] = np.sin(m * X[1])
    return g.forward(U), X


@pytest.mark.parametrize("rule", ["2/3-rule", "3/2-rule"])
def test_shear_mode_curl_and_rhs(rule):
    g = Grid((16, 16, 16), (TWO_PI,) * 3, "double", rule)
    m = 3
    uh, X = _shear(g, m)
    curl = g.backward(g.cross2(uh))
    assert np.max(abs(curl[0])) < 1e-13 and np.max(abs(curl[1])) < 1e-13
    assert np.max(abs(curl[2] + m * np.cos(m * X[1]))) < 1e-13  # w_z = dv/dx - du/dy
    rhs = g.compute_rhs(uh, 0.01)
    # u x w is the gradient (0, m sin cos, 0): removed by the projection
    assert rel(rhs, -0.01 * m * m * uh) < 1e-13


def _abc(g, A=1.0, B=0.7, C=0.4):
    X = np.meshgrid(*g.X, indexing="ij")
    U = np.empty((3,) + g.N)
    U[0] = A * np.sin(X[2]) + C * np.cos(X[1])
    U[1] = B * np.sin(X[0]) + A * np.cos(X[2])
    U[2] = C * np.sin(X[1]) + B * np.cos(X[0])
    return g.forward(U), U


@pytest.mark.parametrize("rule", ["2/3-rule", "3/2-rule", "None"])
def test_abc_beltrami(rule):
    g = Grid((16, 16, 16), (TWO_PI,) * 3, "double", rule)
    uh, U = _abc(g)
    assert rel(g.backward(g.cross2(uh)), U) < 1e-14          # curl u = u
    assert np.max(abs(g.conv(uh))) < 1e-15                    # u x curl u = 0
    nu, dt = 0.05, 0.1
    assert rel(g.compute_rhs(uh, nu), -nu * uh) < 5e-14       # |k|^2 = 1 on every mode
    z = nu * dt
    amp = 1 - z + z ** 2 / 2 - z ** 3 / 6 + z ** 4 / 24       # classical RK4 on u' = -nu u
    assert rel(g.rk4_step(uh, dt, nu), amp * uh) < 1e-14


def test_projection_removes_divergence():
    rng = np.random.default_rng(3)
    g = Grid((12, 16, 8), (TWO_PI, 3.0, 5.0), "double", "2/3-rule")
    uh = g.forward(rng.standard_normal((3,) + g.N))
    rhs = g.compute_rhs(uh, 0.0)
    div = g.K[0] * rhs[0] + g.K[1] * rhs[1] + g.K[2] * rhs[2]
    assert np.max(abs(div)) < 1e-13 * np.max(abs(rhs))


def test_nyquist_mask():
    rng = np.random.default_rng(4)
    g = Grid((8, 8, 8), (TWO_PI,) * 3, "double", "None", True)
    uh = g.forward(rng.standard_normal((3, 8, 8, 8)))
    rhs = g.compute_rhs(uh, 0.0)
    assert np.all(rhs[:, 4] == 0) and np.all(rhs[:, :, 4] == 0) and np.all(rhs[..., 4] == 0)
    g2 = Grid((8, 8, 8), (TWO_PI,) * 3, "double", "None", False)
    assert np.max(abs(g2.compute_rhs(uh, 0.0)[:, 4])) > 1e-3


# ------------------------------------------------------------------ dealiasing
def _full_spectrum(g, a_hat):
    """(3, N0, N1, N2) complex spectrum from the R2C half via Hermitian symmetry."""
    N = g.N
    full = np.zeros(a_hat.shape[:-1] + (N[2],), dtype=complex)
    full[..., :g.Nh] = a_hat
    for kz in range(1, (N[2] + 1) // 2):
        sub = a_hat[..., kz]
        full[..., N[2] - kz] = np.conj(np.roll(np.flip(sub, axis=(-2, -1)), shift=(1, 1), axis=(-2, -1)))
    return full


def _triads(g, uh, exact):
    """sum over p + q = k of u_hat(p) x w_hat(q) for kept input modes p, q.
    exact=True: integer sum (no wrap-around); False: modulo the transform size."""
    N, M = g.N, g.M
    keep_in = [k[0] for k in g.keep]
    U = _full_spectrum(g, uh)
    W = _full_spectrum(g, g.cross2(uh))
    modes = [np.fft.fftfreq(N[i], 1.0 / N[i]).astype(int) for i in range(3)]
    kept = []
    for i in range(3):
        neg, pos = keep_in[i] if i < 2 else (keep_in[2][1], keep_in[2][1])
        kept.append([j for j in range(N[i]) if -neg <= modes[i][j] <= pos])
    out = np.zeros((3,) + N, dtype=complex)
    for a in kept[0]:
        for b in kept[1]:
            for c in kept[2]:
                p = np.array([modes[0][a], modes[1][b], modes[2][c]])
                up = U[:, a, b, c]
                for a2 in kept[0]:
                    for b2 in kept[1]:
                        for c2 in kept[2]:
                            q = np.array([modes[0][a2], modes[1][b2], modes[2][c2]])
                            k = p + q
                            if exact:
                                if np.any(k < -(np.array(N) // 2)) or np.any(k >= np.array(N) // 2):
                                    continue
                            else:
                                k = ((k + np.array(M) // 2) % np.array(M)) - np.array(M) // 2
                                if np.any(k < -(np.array(N) // 2)) or np.any(k >= np.array(N) // 2):
                                    continue
                            wq = W[:, a2, b2, c2]
                            out[:, k[0] % N[0], k[1] % N[1], k[2] % N[2]] += np.cross(up, wq)
    return out[..., :g.Nh]


@pytest.mark.parametrize("rule,mode,exact", [("2/3-rule", 0, True), ("3/2-rule", 0, True),
                                             ("2/3-rule", 1, False), ("None", 0, False)])
def test_dealias_triad_sums(rule, mode, exact):
    rng = np.random.default_rng(5)
    g = Grid((6, 6, 6), (TWO_PI,) * 3, "double", rule, False, mode)
    uh = g.forward(rng.standard_normal((3, 6, 6, 6)))
    uh[:, 3] = 0
    uh[:, :, 3] = 0
    uh[..., 3] = 0
    conv = g.conv(uh)
    ref = _triads(g, uh, exact)
    neg, pos = g.keep[0][1]
    sel = [m for m in range(6) if -neg <= np.fft.fftfreq(6, 1 / 6)[m] <= pos and m != 3]
    kz = [m for m in range(g.Nh) if m <= g.keep[2][1][1] and m != 3]
    assert np.max(abs(conv[np.ix_(range(3), sel, sel, kz)] - ref[np.ix_(range(3), sel, sel, kz)])) < 1e-13


def test_keep_ranges():
    assert keep_ranges(32, "2/3-rule", 0, False) == ((10, 10), (16, 15))
    assert keep_ranges(32, "2/3-rule", 1, False) == ((11, 11), (11, 11))
    assert keep_ranges(256, "2/3-rule", 0, True) == ((0, 85), (0, 128))
    assert keep_ranges(256, "2/3-rule", 1, True) == ((0, 85), (0, 85))
    assert keep_ranges(256, "3/2-rule", 0, False) == ((127, 127), (128, 127))


# ------------------------------------------------------------------ diagnostics
def test_energy_parseval_and_tg_closed_forms():
    rng = np.random.default_rng(6)
    g = Grid((8, 12, 10), (TWO_PI,) * 3, "double", "None")
    U = rng.standard_normal((3,) + g.N)
    assert abs(g.kinetic_energy(g.forward(U)) - 0.5 * np.mean(np.sum(U * U, 0))) < 1e-14
    g = Grid((16, 16, 16), (TWO_PI,) * 3)
    uh = g.taylor_green()
    assert abs(g.kinetic_energy(uh) - 1 / 8) < 1e-15
    assert abs(0.5 * g.energy_fourier(g.cross2(uh)) - 3 / 8) < 1e-15
    assert abs(g.dissipation(uh, NU) - 2 * NU * 3 / 8) < 1e-18


def test_forcing_holds_energy():
    rng = np.random.default_rng(7)
    g = Grid((16, 16, 16), (TWO_PI,) * 3)
    uh = g.forward(rng.standard_normal((3,) + g.N))
    uh[:, 0, 0, 0] = 0
    low = g.energy_fourier(uh * (g.K2 <= 4))
    target = g.energy_fourier(uh) - 0.5 * low   # alpha^2 = 1/2 on |k| <= 2
    f = g.force(uh, target, 2)
    assert abs(g.energy_fourier(f) - target) < 1e-12
    hi = g.K2 > 4
    assert np.array_equal(f[:, hi], uh[:, hi])


def test_spectrum_single_shell():
    g = Grid((16, 16, 16), (TWO_PI,) * 3)
    uh = g.taylor_green()  # every mode has |k|^2 = 3: shell (1.5, 2.5]
    Ek, bins = g.spectrum(uh)
    assert np.count_nonzero(Ek > 1e-12 * Ek.max()) == 1 and Ek[1] == Ek.max()


def brute_force_spectrum(N, L, uh):
    """Shell spectrum by an explicit loop over every stored mode, written from
    the definition (demo/Isotropic.py:88-115), with no digitize and no
    vectorised weights: mode (i, j, l) has wavenumber
    k = (2 pi/L_a) * (index, or index - N_a above N_a/2) on axes 0 and 1 and
    (2 pi/L_2) * l on the R2C axis; its weight is 1 on l = 0 and l = N2/2 and 2
    elsewhere; shell s = 1..Nb-1 holds s - 1/2 < |k| <= s + 1/2, and
    E(s) = (4 pi/3) ((s + 1/2)^3 - (s - 1/2)^3) * (sum of w sum_c |u_c|^2) / count."""
    import math
    N0, N1, N2 = N
    Nh = N2 // 2 + 1
    Nb = int(math.sqrt(sum((n / 2.0) ** 2 for n in N) / 3))
    sums = [0.0] * Nb
    counts = [0] * Nb

    def wn(i, n, length):
        return 2 * math.pi / length * (i if 2 * i < n else i - n)

    for i in range(N0):
        kx = wn(i, N0, L[0])
        for j in range(N1):
            ky = wn(j, N1, L[1])
            for m in range(Nh):
                kz = 2 * math.pi / L[2] * m
                k = math.sqrt(kx * kx + ky * ky + kz * kz)
                w = 1.0 if m == 0 or m == N2 // 2 else 2.0
                e = w * sum(abs(complex(uh[c, i, j, m])) ** 2 for c in range(3))
                for s in range(1, Nb):
                    if s - 0.5 < k <= s + 0.5:
                        sums[s - 1] += e
                        counts[s - 1] += 1
    Ek = [0.0] * Nb
    for s in range(1, Nb):
        if counts[s - 1]:
            Ek[s - 1] = 4 * math.pi / 3 * ((s + 0.5) ** 3 - (s - 0.5) ** 3) * sums[s - 1] / counts[s - 1]
    return np.array(Ek), np.array(counts)


@pytest.mark.parametrize("N,L", [((8, 8, 8), (TWO_PI,) * 3),
                                 ((12, 8, 16), (TWO_PI, TWO_PI, TWO_PI)),
                                 ((8, 12, 10), (TWO_PI, 2 * TWO_PI, 0.5 * TWO_PI))])
def test_spectrum_brute_force(N, L):
    """Every shell of a random multi-shell field against the explicit loop:
    the 4 pi/3 factor, the (k^3 - k0^3)/count normalisation, the kz weights
    1/2/1 and the scaled wavenumbers of a non-2pi box all enter."""
    rng = np.random.default_rng(3)
    g = Grid(N, L)
    uh = g.forward(rng.standard_normal((3,) + N))
    Ek, bins = g.spectrum(uh)
    ref, counts = brute_force_spectrum(N, L, uh)
    assert np.count_nonzero(ref) >= 3
    assert np.allclose(bins, np.arange(len(bins)) + 0.5)
    assert np.max(np.abs(Ek - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_spectrum_closed_form():
    """A single Fourier mode (kx, ky, kz) = (2, 1, 3) of amplitude a in one
    component, with its conjugate partner on the kz = 0 plane absent (kz > 0):
    |k| = sqrt(14) lies in shell 4 (3.5, 4.5], whose modes number exactly the
    lattice points counted by hand below; E(4) = 4pi/3 (4.5^3 - 3.5^3) 2|a|^2 / n4."""
    N = (16, 16, 16)
    g = Grid(N)
    uh = np.zeros((3, 16, 16, 9), dtype=np.complex128)
    a = 0.3 - 0.4j
    uh[1, 2, 1, 3] = a
    Ek, _ = g.spectrum(uh)
    n4 = 0
    for i in range(-8, 8):
        for j in range(-8, 8):
            for m in range(0, 9):
                if 3.5 ** 2 < i * i + j * j + m * m <= 4.5 ** 2:
                    n4 += 1
    expect = 4 * np.pi / 3 * (4.5 ** 3 - 3.5 ** 3) * 2 * abs(a) ** 2 / n4
    assert abs(Ek[3] - expect) <= 1e-15 * expect
    assert np.count_nonzero(Ek) == 1


# ------------------------------------------------------------------ time loop
def test_dt_schedule_last_step():
    d = dt_schedule(0.1, 0.01)
    assert len(d) == 10
    t = 0.0
    for x in d[:-1]:
        t += x
    assert d[-1] == 0.1 - t and all(x == 0.01 for x in d[:-1])
    assert abs(sum(d) - 0.1) < 1e-15
    # single precision: float32 t accumulates 0.01f ten times and misses T, so
    # end_of_tstep (NS.py:118-120) schedules an 11th step of dt = T - t
    s = dt_schedule(0.1, 0.01, precision="single")
    assert all(isinstance(x, np.float32) for x in s)
    assert len(s) == 11 and 0 < s[-1] < 1e-7
    assert len(dt_schedule(0.05, 0.02)) == 3 and abs(dt_schedule(0.05, 0.02)[-1] - 0.01) < 1e-15
