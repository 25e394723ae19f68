"""Forced homogeneous isotropic turbulence on the GPU through the reference's
host API (demo/Isotropic.py of spectralDNS).

Initial field: seeded random-phase solenoidal field with E(k) ~ k^4
exp(-2 (k/k0)^2) (spectraldns_amd.isotropic).  After every step the ``update``
hook holds energy_fourier(u_hat) at its initial value by rescaling the modes
|k| <= Kf2 (Isotropic.py:154-184, one fused GPU reduction + rescale) and
prints energy, dissipation and Re_lambda every ``compute_energy`` steps; the
shell spectrum (Isotropic.py:88-147) is computed on the GPU every
``compute_spectrum`` steps.

    python demo/Isotropic.py --N 64 64 64 --Kf2 2 --T 0.05 --dt 0.001 NS
"""
import os
import sys

import numpy as np
import torch

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from spectraldns_amd import config, get_solver, solve  # noqa: E402
from spectraldns_amd.fourier import energy_fourier, spectrum  # noqa: E402
from spectraldns_amd.isotropic import random_solenoidal_field  # noqa: E402


def initialize(solver, context):
    c = context
    u = random_solenoidal_field(config.params.N, config.params.L, seed=config.params.seed,
                                k0=config.params.k0, energy=config.params.energy)
    c.U_hat.copy_(torch.from_numpy(u).to(c.U_hat.dtype))
    if "VV" in config.params.solver:
        # the state of the velocity-vorticity solver is the curl (Isotropic.py:71-72)
        c.W_hat = solver.cross2(c.W_hat, c.K, c.U_hat)
    c.target_energy = energy_fourier(c.U_hat, c.T)
    config.params.t = 0.0
    config.params.tstep = 0


def init_from_file(filename, solver, context):
    """Restart from a checkpoint (Isotropic.py:267-282): U_hat from its
    ``U/3D/0`` data, the mean mode zeroed, W_hat for the VV solver, and the
    target energy of the forcing.  Also returns the checkpoint's t and tstep."""
    from spectraldns_amd.h5io import read_checkpoint
    f = read_checkpoint(filename)
    assert "U/3D/0" in f
    U_hat = context.U_hat
    U_hat.copy_(torch.from_numpy(f["U/3D/0"]).to(U_hat.dtype))
    U_hat[:, 0, 0, 0] = 0.0
    if "VV" in config.params.solver:
        context.W_hat = solver.cross2(context.W_hat, context.K, context.U_hat)
    context.target_energy = energy_fourier(U_hat, context.T)
    return f["t"], f["tstep"]


def update(context):
    c = context
    params = config.params
    solver = config.solver
    vv = "VV" in params.solver
    if vv:
        # velocity from the vorticity state before the forcing (Isotropic.py:160-161)
        c.U_hat = solver.cross2(c.U_hat, c.K_over_K2, c.W_hat)
    # the compensated single-precision state (params.compensated) is forced as a pair
    c.T.plan.force(c.U_hat, c.target_energy, params.Kf2, u_lo=c.get("U_hat_lo"))
    if vv:
        # and the vorticity from the forced velocity (Isotropic.py:186-187)
        c.W_hat = solver.cross2(c.W_hat, c.K, c.U_hat)
    if params.tstep % params.compute_spectrum == 0:
        Ek, bins = spectrum(c.U_hat, c.T)
        c.spectra[params.tstep] = Ek
    if params.tstep % params.compute_energy == 0:
        out = torch.empty(2, dtype=torch.float64, device=c.U_hat.device)
        c.T.plan.energy(c.U_hat, params.nu, out)
        kk, eps = out.cpu().numpy()
        Re_lam = np.sqrt(20 * kk ** 2 / (3 * params.nu * eps))
        print("%2.4f %2.6e %2.6e %2.3f" % (params.t, kk, eps, Re_lam))


if __name__ == "__main__":
    config.update({"nu": 0.0008, "dt": 0.001, "T": 0.05, "L": [2 * np.pi] * 3, "dealias": "3/2-rule"},
                  "triplyperiodic")
    config.triplyperiodic.add_argument("--N", default=[64, 64, 64], nargs=3, help="Mesh size. Trumps M.")
    config.triplyperiodic.add_argument("--compute_energy", type=int, default=10)
    config.triplyperiodic.add_argument("--compute_spectrum", type=int, default=50)
    config.triplyperiodic.add_argument("--Kf2", type=float, default=2.0)
    config.triplyperiodic.add_argument("--k0", type=float, default=4.0)
    config.triplyperiodic.add_argument("--energy", type=float, default=0.5)
    config.triplyperiodic.add_argument("--seed", type=int, default=1234)
    config.triplyperiodic.add_argument("--restart", default="", help="checkpoint (<name>_c.npz) to start from")
    sol = get_solver(update=update, mesh="triplyperiodic")
    context = sol.get_context()
    context.spectra = {}
    initialize(sol, context)
    if config.params.restart:
        config.params.t, config.params.tstep = init_from_file(config.params.restart, sol, context)
    solve(sol, context)
