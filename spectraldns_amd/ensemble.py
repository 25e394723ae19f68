"""Uneven-size ensembles of independent triply periodic boxes.

Each box is its own NS problem on [0, 2 pi)^3 (or the L it is given) with its
own N, viscosity, time step and initial data; the whole ensemble advances with
one launch per pass of the RK4 step (``_abi.Batch``).  The default ensemble is
96 boxes: 32 at 32^3, 32 at 48^3 and 32 at 64^3, 2/3-rule dealiasing,
viscosities log-uniform in [1e-3, 1e-2], initial data alternating between the
Taylor-Green vortex (demo/TG.py:14-28) and a seeded random-phase solenoidal
field with E(k) ~ k^4 exp(-2 (k/k0)^2), k0 = 4, energy 0.5
(demo/Isotropic.py:29-76 construction, ``isotropic.random_solenoidal_field``).

Initial data are generated on the host in FP64 (NumPy), so a checker can start
from the same arrays; the time stepping runs only on the GPU.
"""
import numpy as np

from .isotropic import random_solenoidal_field

__all__ = ["default_sizes", "taylor_green_hat", "Ensemble"]


def default_sizes():
    return [32] * 32 + [48] * 32 + [64] * 32


def taylor_green_hat(N, L=(2 * np.pi,) * 3):
    """u_hat of u = sin x cos y cos z, v = -cos x sin y cos z, w = 0 (TG.py:14-28),
    forward-scaled by 1/(N0 N1 N2) as the solver's transforms, complex128."""
    N = tuple(int(n) for n in N)
    x = [np.arange(n) * (l / n) for n, l in zip(N, L)]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    u = np.empty((3,) + N)
    u[0] = np.sin(X) * np.cos(Y) * np.cos(Z)
    u[1] = -np.cos(X) * np.sin(Y) * np.cos(Z)
    u[2] = 0.0
    return np.fft.rfftn(u, axes=(1, 2, 3)) / float(np.prod(N))


class Ensemble:
    """Boxes of ``sizes`` (cubic N per box), their initial data, parameters,
    and the GPU batch that advances them.

    init: "mixed" (even boxes Taylor-Green, odd boxes random phase), "tg" or
    "random".  Box i's random field uses seed + i; viscosities come from a
    generator seeded with ``seed``.  tuned: apply the ensemble autotuner's
    stored schedules for these sizes.
    """

    def __init__(self, sizes=None, precision="single", dealias="2/3-rule", dealias_mode=0, seed=2024,
                 dt=0.005, nu_range=(1e-3, 1e-2), init="mixed", k0=4.0, energy=0.5, tuned=True):
        import torch
        from ._abi import Batch
        self.sizes = list(default_sizes() if sizes is None else sizes)
        self.n = len(self.sizes)
        rng = np.random.Generator(np.random.PCG64(seed))
        lo, hi = np.log(nu_range[0]), np.log(nu_range[1])
        self.nu = [float(np.exp(v)) for v in rng.uniform(lo, hi, self.n)]
        self.dt = [float(dt)] * self.n if np.isscalar(dt) else [float(d) for d in dt]
        self.kind = []
        self.u0 = []
        for i, N in enumerate(self.sizes):
            tg = init == "tg" or (init == "mixed" and i % 2 == 0)
            self.kind.append("tg" if tg else "random")
            if tg:
                self.u0.append(taylor_green_hat((N, N, N)))
            else:
                self.u0.append(random_solenoidal_field((N, N, N), seed=seed + i, k0=k0, energy=energy))
        cd = torch.complex64 if precision == "single" else torch.complex128
        self.u_hat = [torch.from_numpy(u).to(cd).cuda().contiguous() for u in self.u0]
        self.batch = Batch([((N, N, N), (2 * np.pi,) * 3) for N in self.sizes], precision, dealias, True,
                           dealias_mode)
        self.batch.set_state(self.u_hat)
        self.batch.set_params(self.nu, self.dt)
        # the batched passes' schedules stored by the ensemble autotuner, if any
        from .autotune import apply_batch_schedule
        self.schedule = apply_batch_schedule(self.batch, self.sizes, precision, dealias, dealias_mode) \
            if tuned else None

    def reset(self):
        """Copy the initial data back into the state tensors."""
        import torch
        for t, u in zip(self.u_hat, self.u0):
            t.copy_(torch.from_numpy(u).to(t.dtype))

    def steps(self, nsteps, diag=None):
        self.batch.rk4_steps(nsteps, diag)

    def points(self):
        """Physical grid points of the whole ensemble."""
        return int(sum(N ** 3 for N in self.sizes))
