"""NumPy restatement of the MHD solver in Elsasser form (spectralDNS/solvers/MHD.py).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

What it restates (reference file:line), on oracle.ns.Grid's transforms:
  * set_Elsasser ....................................... solvers/MHD.py:104-113
  * divergenceConvection: ZZ_ij = F(z0_i z1_j), z0 = u + b,
    z1 = u - b on the dealiased space .................. MHD.py:115-130
  * Conv (Divergence) .................................. MHD.py:132-148
  * add_pressure_diffusion with nu on u and eta on b .. MHD.py:150-165
  * ComputeRHS (conv, Nyquist mask, pressure+diffusion) MHD.py:167-176
  * RK4 on ub_hat (6 components) ....................... maths/integrators.py:149-159
  * the Taylor-Green start of tests/TGMHD.py:4-14

How it is pinned (tests/test_oracle_mhd.py): the reference's own regression
values k(0.1) = 0.124565408177 and b(0.1) = 0.124637762143 at 7 decimals
(tests/TGMHD.py:16-25) on the uniform and the non-uniform box of
tests/test_MHD.py:15-30 (nu = 0.000625, eta = 0.01, dt = 0.01, 2/3 rule); and
with b = 0 the velocity follows the NS oracle (pinned on its own) on the
modes the 2/3 rule keeps: the Elsasser products reduce to u_i u_j, and the
divergence form differs from the rotational one by a gradient the
projection removes and by the aliasing left on the dropped modes.
"""
import numpy as np


def set_elsasser(g, ZZ):
    K = g.K
    c = np.empty((6,) + ZZ.shape[2:], dtype=ZZ.dtype)
    c[:3] = -1j * (K[0] * (ZZ[:, 0] + ZZ[0, :])
                   + K[1] * (ZZ[:, 1] + ZZ[1, :])
                   + K[2] * (ZZ[:, 2] + ZZ[2, :])) / 2.0
    c[3:] = 1j * (K[0] * (ZZ[0, :] - ZZ[:, 0])
                  + K[1] * (ZZ[1, :] - ZZ[:, 1])
                  + K[2] * (ZZ[2, :] - ZZ[:, 2])) / 2.0
    return c


def conv(g, ub_hat):
    ub = g.backward(ub_hat, dealiased=True)
    u, b = ub[:3], ub[3:]
    z0, z1 = u + b, u - b
    ZZ = np.empty((3, 3) + ub_hat.shape[1:], dtype=ub_hat.dtype)
    for i in range(3):
        for j in range(3):
            ZZ[i, j] = g.forward(z0[i] * z1[j], dealiased=True)
    return set_elsasser(g, ZZ)


def compute_rhs(g, ub_hat, nu, eta):
    nu, eta = g.float(nu), g.float(eta)
    rhs = conv(g, ub_hat)
    if g.mask_nyquist_on:
        rhs = g.mask_nyquist(rhs)
    P_hat = np.sum(rhs[:3] * g.K_over_K2, 0)
    for i in range(3):
        rhs[i] -= P_hat * g.K[i]
    rhs[:3] -= nu * g.K2 * ub_hat[:3]
    rhs[3:] -= eta * g.K2 * ub_hat[3:]
    return rhs


def rk4_step(g, u0, dt, nu, eta):
    a = np.array([1. / 6., 1. / 3., 1. / 3., 1. / 6.], dtype=g.float)
    b = np.array([0.5, 0.5, 1.], dtype=g.float)
    dt = g.float(dt)
    u0 = u0.copy()
    u1 = u0.copy()
    u2 = u0.copy()
    for rk in range(4):
        rhs = compute_rhs(g, u0, nu, eta)
        if rk < 3:
            u0[:] = u1 + b[rk] * dt * rhs
        u2 += a[rk] * dt * rhs
    u0[:] = u2
    return u0


def taylor_green(g):
    """U and B of tests/TGMHD.py:4-14, transformed."""
    X = np.meshgrid(*g.X, indexing="ij")
    UB = np.zeros((6,) + g.N)
    UB[0] = np.sin(X[0]) * np.cos(X[1]) * np.cos(X[2])
    UB[1] = -np.cos(X[0]) * np.sin(X[1]) * np.cos(X[2])
    UB[3] = np.sin(X[0]) * np.sin(X[1]) * np.cos(X[2])
    UB[4] = np.cos(X[0]) * np.cos(X[1]) * np.cos(X[2])
    return g.forward(UB.astype(g.float))


def regression_values(g, ub_hat):
    """k and b of tests/TGMHD.py:16-25 (mean of the squares over the box, halved)."""
    UB = g.backward(ub_hat).astype(np.float64)
    n = np.prod(g.N)
    return float(np.sum(UB[:3] ** 2) / n / 2), float(np.sum(UB[3:] ** 2) / n / 2)
