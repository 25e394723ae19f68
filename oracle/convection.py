"""NumPy restatement of the NS solver's Standard, Divergence and Skewed
convection forms (spectralDNS/solvers/NS.py:138-199) on oracle.ns.Grid.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

What it restates (reference file:line):
  * standard_convection: rhs_i = F(u_j d_j u_i), gradients from
    Tp.backward(1j*K_j*U_hat_i) ........................ solvers/NS.py:138-145
  * divergence_convection: rhs_i = 1j sum_j K_j F(u_i u_j), with the
    reference's grouping of the sums ..................... NS.py:147-165
  * Conv for "Standard" (times -1), "Divergence" (times -1), "Skewed"
    (standard + divergence, times -0.5) ................. NS.py:167-190

How it is pinned (tests/test_oracle_convection.py): after the Nyquist mask
and the pressure projection every form gives the Vortex form's RHS (the NS
oracle, pinned on its own) on the modes the 2/3 rule keeps -- the forms
differ by a gradient and by the aliasing left on the dropped modes -- and
for a field whose products do not alias (no dealiasing, modes |k| < N/4)
on every mode; and the reference's Taylor-Green pins k(0.1), w(0.1) hold at
7 decimals for each form (tests/TG.py:116-126).
"""
import numpy as np


def standard_convection(g, u_dealias, u_hat):
    rhs = np.empty_like(u_hat)
    for i in range(3):
        grad = np.stack([g.backward(1j * g.K[j] * u_hat[i], dealiased=True) for j in range(3)])
        rhs[i] = g.forward(np.sum(u_dealias * grad, 0), dealiased=True)
    return rhs


def divergence_convection(g, rhs, u_dealias, add=False):
    K = g.K
    if not add:
        rhs = np.zeros_like(rhs)
    fw = lambda a: g.forward(a, dealiased=True)  # noqa: E731
    UU = [fw(u_dealias[0] * u_dealias[i]) for i in range(3)]
    rhs[0] += 1j * (K[0] * UU[0] + K[1] * UU[1] + K[2] * UU[2])
    rhs[1] += 1j * K[0] * UU[1]
    rhs[2] += 1j * K[0] * UU[2]
    UU = [fw(u_dealias[1] * u_dealias[1]), fw(u_dealias[1] * u_dealias[2]), fw(u_dealias[2] * u_dealias[2])]
    rhs[1] += (1j * K[1] * UU[0] + 1j * K[2] * UU[1])
    rhs[2] += (1j * K[1] * UU[1] + 1j * K[2] * UU[2])
    return rhs


def conv(g, u_hat, form):
    u_d = g.backward(u_hat, dealiased=True)
    if form == "Standard":
        return -standard_convection(g, u_d, u_hat)
    if form == "Divergence":
        return -divergence_convection(g, np.zeros_like(u_hat), u_d)
    if form == "Skewed":
        rhs = standard_convection(g, u_d, u_hat)
        return -0.5 * divergence_convection(g, rhs, u_d, add=True)
    if form == "Vortex":
        return g.conv(u_hat)
    raise ValueError(form)


def compute_rhs(g, u_hat, nu, form):
    """ComputeRHS (NS.py:219-261) with the given convection form."""
    nu = g.float(nu)
    rhs = conv(g, u_hat, form)
    if g.mask_nyquist_on:
        rhs = g.mask_nyquist(rhs)
    rhs, _ = g.add_pressure_diffusion(rhs, u_hat, nu)
    return rhs


def rk4_step(g, u0, dt, nu, form):
    a = np.array([1. / 6., 1. / 3., 1. / 3., 1. / 6.], dtype=g.float)
    b = np.array([0.5, 0.5, 1.], dtype=g.float)
    dt = g.float(dt)
    u0 = u0.copy()
    u1 = u0.copy()
    u2 = u0.copy()
    for rk in range(4):
        rhs = compute_rhs(g, u0, nu, form)
        if rk < 3:
            u0[:] = u1 + b[rk] * dt * rhs
        u2 += a[rk] * dt * rhs
    u0[:] = u2
    return u0
