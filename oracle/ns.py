"""NumPy/SciPy restatement of the spectralDNS triply periodic NS solver.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.

What it restates (reference file:line):
  * get_context arrays, wavenumbers, K2, K/K^2 ....... spectralDNS/solvers/NS.py:12-48
  * Vortex convection  u x curl(u) ..................... NS.py:124-136, 191-198
  * cross1 / cross2 .................................... spectralDNS/maths/cross.py:13-32
  * Nyquist mask + add_pressure_diffusion + Source .... NS.py:203-217, 253-259
  * RK4 stage arithmetic ............................... spectralDNS/maths/integrators.py:149-159, 183-191
  * time loop and last-step dt adjustment ............. spectralDNS/__init__.py:94-112, NS.py:112-122
  * nu/dt cast to the working precision ............... spectralDNS/config.py:123-127
  * energy_fourier weights, spectrum, forcing ......... demo/Isotropic.py:88-147, 154-184
  * Taylor-Green initial field ......................... demo/TG.py:14-28

Conventions (identical in the HIP kernels):
  * spectral layout (3, N0, N1, N2//2+1), kz contiguous; physical (3, N0, N1, N2).
  * u_hat = (1/(N0 N1 N2)) sum u exp(-i k.x); backward is the unscaled inverse.
  * wavenumbers k_i = (2 pi / L_i) * fftfreq(N_i)*N_i on axes 0, 1 (Nyquist -N/2),
    rfftfreq on axis 2 (Nyquist +N/2) (spectralDNS3D_short.py:38-41).
  * dealiasing is a tensor-product keep set per axis for the inputs of the
    backward transforms inside the convection term and for the outputs of its
    forward transform (``keep_ranges``).  The 3/2 rule zero-pads to M = 3N/2 and
    drops the (masked) Nyquist mode on the way in.
  * operation order of the pointwise work follows NS.py:208-215 and
    integrators.py:155-157 so that FP64 results agree to rounding.

How each part is pinned (tests/test_oracle_*.py):
  * transforms: brute-force DFT sums on N <= 8 and round trips.
  * curl / cross / projection: ABC (Beltrami) field with curl u = u, so the
    convection term is a pure gradient and rhs = -nu k^2 u_hat exactly; the
    RK4 amplification polynomial per step; a shear mode; k . rhs = 0.
  * dealiasing keep sets: brute-force triad sums (p + q = k mod N) on N = 8, 12.
  * energy: Parseval against the physical mean and the Taylor-Green closed
    form E(0) = 1/8, Omega(0) = 3/8.
  * end to end: the reference's Taylor-Green pins k(0.1) = 0.124953117517 and
    w(0.1) = 0.375249930801 to 7 decimals (tests/TG.py:125-126) on 16^3, 32^3,
    64^3 and 64x32x16 on [0,6pi)x[0,4pi)x[0,2pi) (tests/test_NSVV.py:27-30).
  * shenfun's exact 2/3-rule index (dealias_direct) cannot be imported here:
    the cutoff |k_i| < N_i/3 used by dealias_mode 0 is this oracle's stated
    definition, pinned by the triad sums, and "parity unpinned" against shenfun.
"""
import os

import numpy as np
import scipy.fft as sfft

_WORKERS = max(1, len(os.sched_getaffinity(0)))


def datatypes(precision):
    """(real, complex) numpy types (spectralDNS/solvers/spectralinit.py:25-29)."""
    return {"single": (np.float32, np.complex64),
            "double": (np.float64, np.complex128)}[precision]


def keep_ranges(N, dealias, dealias_mode, r2c):
    """Inclusive wavenumber bounds kept on one axis.

    Returns ((in_neg, in_pos), (out_neg, out_pos)): the modes -neg <= k <= pos
    that enter the dealiased backward transform and that are kept from the
    dealiased forward transform.  For the R2C axis neg is always 0.
    """
    half = N // 2
    full = (0, half) if r2c else (half, half - 1)
    if dealias in (None, "None"):
        return full, full
    if dealias == "2/3-rule":
        if dealias_mode == 0:
            # keep 3|k| < N on the inputs (shenfun dealias_direct semantics as
            # documented; the forward output is not truncated)
            K = -(-N // 3) - 1
            inn = (0, K) if r2c else (K, K)
            return inn, full
        # output mask |k| < 2/3 (N/2+1) (cpp/spectralDNS.cpp:122-129,
        # spectralDNS3D_short.py:44-46); applied to inputs too, which is a
        # no-op whenever u_hat satisfies the mask
        K = (2 * (half + 1) - 1) // 3
        rng = (0, min(K, half)) if r2c else (min(K, half), min(K, half - 1))
        return rng, rng
    if dealias == "3/2-rule":
        inn = (0, half - 1) if r2c else (half - 1, half - 1)
        return inn, full
    raise ValueError(dealias)


def _blocks(N, M, neg, pos, r2c):
    """(stored-index slice, padded-bin slice) pairs covering -neg..pos."""
    out = [(slice(0, pos + 1), slice(0, pos + 1))]
    if not r2c and neg > 0:
        out.append((slice(N - neg, N), slice(M - neg, M)))
    return out


class Grid:
    """Spaces, wavenumbers and transforms of one triply periodic box (NS.py:12-48)."""

    def __init__(self, N, L=(2 * np.pi,) * 3, precision="double",
                 dealias="2/3-rule", mask_nyquist=True, dealias_mode=0):
        self.N = tuple(int(n) for n in N)
        assert all(n % 2 == 0 for n in self.N), "even N only"
        self.L = tuple(float(x) for x in L)
        self.precision = precision
        self.float, self.complex = datatypes(precision)
        self.dealias = dealias
        self.dealias_mode = dealias_mode
        self.mask_nyquist_on = bool(mask_nyquist)
        pad = dealias == "3/2-rule"
        self.M = tuple(3 * n // 2 if pad else n for n in self.N)
        self.Nh = self.N[2] // 2 + 1
        self.keep = [keep_ranges(self.N[i], dealias, dealias_mode, i == 2) for i in range(3)]
        # wavenumbers: fftfreq on axes 0,1, rfftfreq on axis 2, scaled by 2 pi/L
        # (NS.py:37-41), computed in float64 and cast like K[i].astype(float)
        ks = [np.fft.fftfreq(self.N[0], 1.0 / self.N[0]),
              np.fft.fftfreq(self.N[1], 1.0 / self.N[1]),
              np.arange(self.Nh, dtype=np.float64)]
        self.kint = [k.astype(np.int64) for k in ks]
        fac = [2 * np.pi / self.L[i] for i in range(3)]
        k1d = [(ks[i] * fac[i]).astype(self.float) for i in range(3)]
        self.k1d = k1d
        self.K = [k1d[0][:, None, None], k1d[1][None, :, None], k1d[2][None, None, :]]
        K2 = np.zeros((self.N[0], self.N[1], self.Nh), dtype=self.float)
        for i in range(3):
            K2 += self.K[i] * self.K[i]
        self.K2 = K2
        self.K_over_K2 = np.zeros((3,) + K2.shape, dtype=self.float)
        for i in range(3):
            self.K_over_K2[i] = self.K[i] / np.where(K2 == 0, 1, K2)
        # physical mesh of the unpadded space (shenfun local_mesh)
        self.X = [(np.arange(self.N[i]) * self.L[i] / self.N[i]).astype(self.float)
                  for i in range(3)]

    # ------------------------------------------------------------------ transforms
    def backward(self, fhat, dealiased=False):
        """Spectral -> physical.  dealiased=True is VTp.backward (NS.py:29-32, 194-195)."""
        if not dealiased:
            return sfft.irfftn(fhat, s=self.N, axes=(-3, -2, -1), norm="forward",
                               workers=_WORKERS).astype(self.float, copy=False)
        M = self.M
        pad = np.zeros(fhat.shape[:-3] + (M[0], M[1], M[2] // 2 + 1), dtype=self.complex)
        (n0, p0), (n1, p1), (_, p2) = [k[0] for k in self.keep]
        for s0, d0 in _blocks(self.N[0], M[0], n0, p0, False):
            for s1, d1 in _blocks(self.N[1], M[1], n1, p1, False):
                pad[..., d0, d1, 0:p2 + 1] = fhat[..., s0, s1, 0:p2 + 1]
        return sfft.irfftn(pad, s=M, axes=(-3, -2, -1), norm="forward",
                           workers=_WORKERS).astype(self.float, copy=False)

    def forward(self, u, dealiased=False):
        """Physical -> spectral, scaled by 1/(M0 M1 M2).  dealiased=True is VTp.forward."""
        if not dealiased:
            return sfft.rfftn(u, axes=(-3, -2, -1), norm="forward",
                              workers=_WORKERS).astype(self.complex, copy=False)
        M = self.M
        F = sfft.rfftn(u, axes=(-3, -2, -1), norm="forward", workers=_WORKERS)
        out = np.zeros(u.shape[:-3] + (self.N[0], self.N[1], self.Nh), dtype=self.complex)
        (n0, p0), (n1, p1), (_, p2) = [k[1] for k in self.keep]
        for s0, d0 in _blocks(self.N[0], M[0], n0, p0, False):
            for s1, d1 in _blocks(self.N[1], M[1], n1, p1, False):
                out[..., s0, s1, 0:p2 + 1] = F[..., d0, d1, 0:p2 + 1]
        return out

    # ------------------------------------------------------------------ kernels
    @staticmethod
    def cross1(a, b):
        """c = a x b (maths/cross.py:13-25)."""
        c = np.empty_like(a)
        c[0] = a[1] * b[2] - a[2] * b[1]
        c[1] = a[2] * b[0] - a[0] * b[2]
        c[2] = a[0] * b[1] - a[1] * b[0]
        return c

    def cross2(self, a_hat):
        """c = 1j*(K x a_hat) (maths/cross.py:27-32 with a = K)."""
        K = self.K
        c = np.empty_like(a_hat)
        c[0] = K[1] * a_hat[2] - K[2] * a_hat[1]
        c[1] = K[2] * a_hat[0] - K[0] * a_hat[2]
        c[2] = K[0] * a_hat[1] - K[1] * a_hat[0]
        c *= 1j
        return c

    def conv(self, u_hat):
        """Vortex form: rhs = F(u x curl u) on the dealiased space (NS.py:191-198)."""
        u_d = self.backward(u_hat, dealiased=True)
        w_d = self.backward(self.cross2(u_hat), dealiased=True)
        return self.forward(self.cross1(u_d, w_d), dealiased=True)

    def mask_nyquist(self, rhs):
        """Zero the Nyquist planes (NS.py:34, 253-254)."""
        rhs[..., self.N[0] // 2, :, :] = 0
        rhs[..., :, self.N[1] // 2, :] = 0
        rhs[..., :, :, self.N[2] // 2] = 0
        return rhs

    def add_pressure_diffusion(self, rhs, u_hat, nu):
        """NS.py:203-217 (operation order kept)."""
        P_hat = np.sum(rhs * self.K_over_K2, 0)
        for i in range(3):
            rhs[i] -= P_hat * self.K[i]
        rhs -= nu * self.K2 * u_hat
        return rhs, P_hat

    def compute_rhs(self, u_hat, nu, source=None):
        """ComputeRHS (NS.py:219-261)."""
        nu = self.float(nu)
        rhs = self.conv(u_hat)
        if self.mask_nyquist_on:
            rhs = self.mask_nyquist(rhs)
        rhs, P_hat = self.add_pressure_diffusion(rhs, u_hat, nu)
        if source is not None:
            rhs += source
        return rhs

    def rk4_step(self, u0, dt, nu, source=None):
        """One RK4 step (integrators.py:149-159, coefficients :185-186); every
        stage's ComputeRHS adds `source` (NS.py:256-258) when given."""
        a = np.array([1. / 6., 1. / 3., 1. / 3., 1. / 6.], dtype=self.float)
        b = np.array([0.5, 0.5, 1.], dtype=self.float)
        dt = self.float(dt)
        u0 = u0.copy()
        u1 = u0.copy()
        u2 = u0.copy()
        for rk in range(4):
            rhs = self.compute_rhs(u0, nu, source)
            if rk < 3:
                u0[:] = u1 + b[rk] * dt * rhs
            u2 += a[rk] * dt * rhs
        u0[:] = u2
        return u0

    # ------------------------------------------------------------------ diagnostics
    def energy_fourier(self, a_hat):
        """sum over components and modes of w |a_hat|^2 with w = 1 on kz = 0 and kz = N/2,
        2 elsewhere (demo/Isotropic.py:90-93), accumulated in float64."""
        w = np.full(self.Nh, 2.0)
        w[0] = 1.0
        w[-1] = 1.0
        a = a_hat.astype(np.complex128)
        return float(np.sum(w * (a.real ** 2 + a.imag ** 2)))

    def kinetic_energy(self, u_hat):
        return 0.5 * self.energy_fourier(u_hat)

    def dissipation(self, u_hat, nu):
        """eps = nu * <|curl u|^2> (demo/Isotropic.py:241-249)."""
        return float(nu) * self.energy_fourier(self.cross2(u_hat))

    def spectrum(self, u_hat):
        """Shell-binned energy spectrum (demo/Isotropic.py:88-115)."""
        a = u_hat.astype(np.complex128)
        m2 = (a.real ** 2 + a.imag ** 2)
        uiui = np.zeros(u_hat.shape[1:])
        uiui[..., 1:-1] = 2 * np.sum(m2[..., 1:-1], axis=0)
        uiui[..., 0] = np.sum(m2[..., 0], axis=0)
        uiui[..., -1] = np.sum(m2[..., -1], axis=0)
        uiui *= (4. / 3. * np.pi)
        Nb = int(np.sqrt(np.sum((np.array(self.N) / 2) ** 2) / 3))
        bins = np.array(range(0, Nb)) + 0.5
        z = np.digitize(np.sqrt(self.K2.astype(np.float64)), bins, right=True)
        Ek = np.zeros(Nb)
        ll = np.zeros(Nb)
        for i, k in enumerate(bins[1:]):
            k0 = bins[i]
            ii = np.where((z > k0) & (z <= k))
            ll[i] = len(ii[0])
            Ek[i] = (k ** 3 - k0 ** 3) * np.sum(uiui[ii])
        for i in range(Nb):
            if ll[i] != 0:
                Ek[i] = Ek[i] / ll[i]
        return Ek, bins

    def force(self, u_hat, target_energy, kf):
        """Large-scale forcing that holds energy_fourier at target (demo/Isotropic.py:154-184)."""
        k2_mask = np.where(self.K2 <= kf ** 2, 1, 0)
        u_hat = u_hat.copy()
        u_hat[:, 0, 0, 0] = 0
        energy_new = self.energy_fourier(u_hat)
        energy_lower = self.energy_fourier(u_hat * k2_mask)
        energy_upper = energy_new - energy_lower
        alpha = np.sqrt((target_energy - energy_upper) / energy_lower)
        u_hat *= (alpha * k2_mask + (1 - k2_mask))
        return u_hat

    # ------------------------------------------------------------------ initial fields
    def taylor_green(self):
        """demo/TG.py:24-28: physical TG field, then set_velocity (VT.forward)."""
        X = np.meshgrid(*self.X, indexing="ij")
        U = np.empty((3,) + self.N, dtype=self.float)
        U[0] = np.sin(X[0]) * np.cos(X[1]) * np.cos(X[2])
        U[1] = -np.cos(X[0]) * np.sin(X[1]) * np.cos(X[2])
        U[2] = 0
        return self.forward(U)


def dt_schedule(T, dt, t=0.0, precision="double"):
    """The dt sequence of the reference time loop.

    Mirrors spectralDNS/__init__.py:94-112 (loop condition, t += dt_took) with
    NS.end_of_tstep (NS.py:112-122), including the float32 cast of dt on every
    read of params.dt in single precision (config.py:123-127).
    """
    fl = np.float32 if precision == "single" else np.float64
    stored = dt
    dts = []
    while t + fl(stored) <= T + 1e-12:
        took = fl(stored)
        dts.append(took)
        t += took
        if abs(t - T) < 1e-12:
            break
        if abs(t + fl(stored) - T) < 1e-12 or t + fl(stored) >= T + 1e-12:
            stored = T - t
    return dts


def solve(grid, u_hat, T, dt, nu, t=0.0, update=None):
    """Integrate with RK4 from t to T; update(grid, u_hat, t, tstep) after each step."""
    tstep = 0
    for d in dt_schedule(T, dt, t, grid.precision):
        u_hat = grid.rk4_step(u_hat, d, nu)
        t += d
        tstep += 1
        if update is not None:
            update(grid, u_hat, t, tstep)
    return u_hat


def tg_regression_values(grid, u_hat):
    """k and w as computed by the TG regression hook (demo/TG.py:116-122)."""
    U = grid.backward(u_hat)
    curl = grid.backward(grid.cross2(u_hat))
    n = np.prod(grid.N)
    w = np.sum(curl.astype(np.float64) * curl.astype(np.float64)) / n / 2
    k = np.sum(U.astype(np.float64) * U.astype(np.float64)) / n / 2
    return float(k), float(w)
