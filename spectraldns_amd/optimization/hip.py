"""HIP implementations of the reference's plug-in kernels, by the names the
``@optimizer`` layer resolves (optimization/cython_maths.in:13-86,
cython_solvers.in:40-80).  Every call is one libsdns_hip launch on the
current torch stream."""
import torch

from .._abi import Plan, SdnsError

__all__ = ["cross1", "cross2", "add_pressure_diffusion_NS", "add_linear"]

_util_plans = {}


def _plan_for(t):
    """A small plan used only for its precision and stream (cross1 needs no grid)."""
    prec = "single" if t.dtype in (torch.float32, torch.complex64) else "double"
    if prec not in _util_plans:
        _util_plans[prec] = Plan((4, 4, 4), (1.0, 1.0, 1.0), prec, "None")
    return _util_plans[prec]


def cross1(c, a, b):
    """c = a x b on real (3, ...) arrays (maths/cross.py:13-25)."""
    return _plan_for(a).cross1(c, a, b)


def cross2(c, a, b):
    """c = 1j*(a x b) with a the wavenumber list K of a plan, or the K_over_K2
    handle of a velocity-vorticity context (maths/cross.py:27-32)."""
    plan = getattr(a, "plan", None)
    if plan is None:
        raise SdnsError("cross2 needs the wavenumber list K (or K_over_K2) of a context (it carries the plan)")
    if getattr(a, "k_over_k2", False):
        return plan.cross2_k_over_k2(c, b)
    return plan.cross2(c, b)


def add_linear(rhs, w_hat, nu, K2, Source):
    """solvers/VV.py:110-115: rhs -= nu*K2*w_hat; rhs += Source."""
    plan = getattr(K2, "plan", None)
    if plan is None:
        raise SdnsError("add_linear needs the K2 handle of a context (it carries the plan)")
    return plan.add_linear(rhs, w_hat, nu, Source)


def add_pressure_diffusion_NS(rhs, u_hat, nu, K2, K, P_hat, K_over_K2):
    """NS.py:203-217: P_hat = sum(rhs*K/K^2); rhs -= P_hat*K; rhs -= nu*K2*u_hat."""
    plan = getattr(K, "plan", None)
    if plan is None:
        raise SdnsError("add_pressure_diffusion needs the wavenumber list K of a context")
    return plan.add_pressure_diffusion(rhs, u_hat, nu, P_hat)
