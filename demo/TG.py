"""Taylor-Green vortex on the GPU through the reference's host API
(demo/TG.py of spectralDNS): initialize, the per-step ``update`` hook that
prints kinetic energy / enstrophy every ``compute_energy`` steps, and the
regression test k(0.1) = 0.124953117517, w(0.1) = 0.375249930801 (tests/TG.py:116-126).

    python demo/TG.py --M 5 5 5 NS
    python demo/TG.py --dealias 3/2-rule --precision single --ntol 5 NS
    python demo/TG.py --integrator BS5_adaptive VV

For the VV solver the state is the vorticity W_hat = 1j*(K x U_hat)
(tests/TG.py:29-35).
"""
import os
import sys

import numpy as np
import torch

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from spectraldns_amd import config, get_solver, solve  # noqa: E402
from spectraldns_amd.fourier import energy_fourier  # noqa: E402


def initialize(solver, context):
    U, X = context.U, context.X
    U[0] = torch.sin(X[0]) * torch.cos(X[1]) * torch.cos(X[2])
    U[1] = -torch.cos(X[0]) * torch.sin(X[1]) * torch.cos(X[2])
    U[2] = 0
    solver.set_velocity(**context)
    if "VV" in config.params.solver:
        solver.cross2(context.W_hat, context.K, context.U_hat)
    config.params.t = 0.0
    config.params.tstep = 0


def _mean_sq(a):
    return float(torch.sum(a.double() * a.double()).item()) / np.prod(config.params.N) / 2


def update(context):
    params = config.params
    solver = config.solver
    if params.tstep % params.compute_energy == 0:
        U = solver.get_velocity(**context)
        curl = solver.get_curl(**context)
        if "VV" in params.solver:  # U_hat from the velocity the vorticity gives
            solver.set_velocity(**context)
        kk = _mean_sq(U)
        ww = _mean_sq(curl)
        ww2 = energy_fourier(context.U_hat, context.T) / 2
        divu = _mean_sq(solver.get_divergence(**context))
        print("%2.2f %2.8f %2.8f %2.8e %2.8e" % (params.t, kk, ww, ww2 - kk, divu))


def regression_test(context):
    params = config.params
    solver = config.solver
    U = solver.get_velocity(**context)
    curl = solver.get_curl(**context)
    w = _mean_sq(curl)
    k = _mean_sq(U)
    assert round(w - 0.375249930801, params.ntol) == 0, w
    assert round(k - 0.124953117517, params.ntol) == 0, k
    print("regression ok: k = %.12f, w = %.12f" % (k, w))


if __name__ == "__main__":
    config.update({"nu": 0.000625, "dt": 0.01, "T": 0.1, "L": [2 * np.pi, 2 * np.pi, 2 * np.pi],
                   "M": [5, 5, 5]}, "triplyperiodic")
    config.triplyperiodic.add_argument("--compute_energy", type=int, default=2)
    config.triplyperiodic.add_argument("--plot_step", type=int, default=2)
    sol = get_solver(update=update, regression_test=regression_test, mesh="triplyperiodic")
    context = sol.get_context()
    initialize(sol, context)
    solve(sol, context)
