"""Time integrators (spectralDNS/maths/integrators.py:177-239).

RK4 (integrators.py:149-159) runs as one fused launch sequence per step: the
stage RHS (dealiased convection, Nyquist mask, pressure, diffusion) and the
stage updates u0 = u1 + b dt rhs, u2 += a dt rhs happen inside the HIP pass
kernels, with a = [1/6, 1/3, 1/3, 1/6] and b = [1/2, 1/2, 1] in the working
precision.  The other integrators of the reference (ForwardEuler, AB2, BS5)
are not built.
"""
import numpy as np

from ..optimization import wraps

__all__ = ["getintegrator", "RK4"]


def RK4(u0, u1, u2, rhs, a, b, dt, solver, context):
    """One classical RK4 step on u0 in place; returns (u0, dt, dt).

    With params.compensated in single precision the state is carried as
    u0 + context.U_hat_lo (a compensated pair, see sdns_rk4_steps)."""
    plan = context.VT.plan
    lo = context.get("U_hat_lo") if solver.params.get("compensated", False) else None
    plan.rk4_steps(u0, [dt], solver.params.nu, p_hat=context.P_hat, u_lo=lo)
    return u0, dt, dt


def getintegrator(rhs, u0, solver, context):
    """Return the integrator chosen by params.integrator (only RK4 is built)."""
    params = solver.params
    if params.integrator != "RK4":
        raise NotImplementedError("integrator %r is not built; the HIP path implements RK4"
                                  % params.integrator)
    if not getattr(solver.conv, "fused_vortex", False):
        raise NotImplementedError("the fused RK4 step serves the Vortex convection of NS")
    if solver.ComputeRHS is not getattr(solver, "_ComputeRHS_fused", None):
        raise NotImplementedError("a replaced ComputeRHS cannot run inside the fused RK4 step")
    if params.get("compensated", False) and params.precision == "single" and "U_hat_lo" not in context:
        import torch
        context["U_hat_lo"] = torch.zeros_like(u0)
    a = np.array([1. / 6., 1. / 3., 1. / 3., 1. / 6.], dtype=context.float)
    b = np.array([0.5, 0.5, 1.], dtype=context.float)

    @wraps(RK4)
    def func():
        return RK4(u0, None, None, rhs, a, b, params.dt, solver, context)
    return func
