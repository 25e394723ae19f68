"""Time integrators (spectralDNS/maths/integrators.py:15-239).

RK4 (integrators.py:149-159) runs as one fused launch sequence per step: the
stage RHS (dealiased convection, Nyquist mask, pressure, diffusion) and the
stage updates u0 = u1 + b dt rhs, u2 += a dt rhs happen inside the HIP pass
kernels, with a = [1/6, 1/3, 1/3, 1/6] and b = [1/2, 1/2, 1] in the working
precision.

ForwardEuler, AB2 (integrators.py:161-175) and the embedded Runge-Kutta driver
adaptiveRK with the Bogacki-Shampine 5(4) tableau (BS5_fixed, BS5_adaptive;
integrators.py:15-147, 193-224) call solver.ComputeRHS (the fused HIP RHS)
per stage; their array updates are the HIP kernels sdns_lincomb (each product
rounded, terms added left to right, as NumPy evaluates ``u = x0; u += c*x``)
and sdns_error_norm (FP64 sums of |err/sc|^2 per component).  The step-size
controller runs on the host, as in the reference.
"""
import numpy as np

from ..optimization import wraps

__all__ = ["getintegrator", "RK4", "RK4_stages", "ForwardEuler", "AB2", "adaptiveRK"]

# Bogacki-Shampine 5(4): the published coefficients (integrators.py:200-209)
_BS5_A = [[0, 0, 0, 0, 0, 0, 0, 0],
          [1 / 6, 0, 0, 0, 0, 0, 0, 0],
          [2 / 27, 4 / 27, 0, 0, 0, 0, 0, 0],
          [183 / 1372, -162 / 343, 1053 / 1372, 0, 0, 0, 0, 0],
          [68 / 297, -4 / 11, 42 / 143, 1960 / 3861, 0, 0, 0, 0],
          [597 / 22528, 81 / 352, 63099 / 585728, 58653 / 366080, 4617 / 20480, 0, 0, 0],
          [174197 / 959244, -30942 / 79937, 8152137 / 19744439, 666106 / 1039181, -29421 / 29068,
           482048 / 414219, 0, 0],
          [587 / 8064, 0, 4440339 / 15491840, 24353 / 124800, 387 / 44800, 2152 / 5985, 7267 / 94080, 0]]
_BS5_B = [587 / 8064, 0, 4440339 / 15491840, 24353 / 124800, 387 / 44800, 2152 / 5985, 7267 / 94080, 0]
_BS5_BHAT = [2479 / 34992, 0, 123 / 416, 612941 / 3411720, 43 / 1440, 2272 / 6561, 79937 / 1113912, 3293 / 556956]


def RK4(u0, u1, u2, rhs, a, b, dt, solver, context):
    """One classical RK4 step on u0 in place; returns (u0, dt, dt).

    Every stage adds context.Source, as ComputeRHS does (NS.py:256-258).
    With params.compensated in single precision the state is carried as
    u0 + context.U_hat_lo (a compensated pair, see sdns_rk4_steps)."""
    plan = context.VT.plan
    lo = context.get("U_hat_lo") if solver.params.get("compensated", False) else None
    plan.rk4_steps(u0, [dt], solver.params.nu, p_hat=context.P_hat, u_lo=lo, source=context.Source)
    return u0, dt, dt


def RK4_stages(u0, u1, u2, rhs, a, b, dt, solver, context):
    """RK4 stage by stage (integrators.py:149-159) for a ComputeRHS other than
    the fused NS one (the VV and MHD solvers, the non-Vortex convection forms,
    or a replaced ComputeRHS): four solver.ComputeRHS calls and the updates
    as sdns_lincomb launches.  The stage values go to u1 (the reference's u0
    keeps the step's start value there instead) and u2 starts as
    u0 + (a dt) rhs at the first stage, so the reference's copies
    u2 = u1 = u0 and u0 = u2 are not needed; every value is the same,
    bit for bit."""
    plan = context.VT.plan
    y = u0
    for rk in range(4):
        rhs = solver.ComputeRHS(rhs, y, solver, **context)
        if rk == 0:
            plan.lincomb(u2, [u0, rhs], [1.0, a[rk] * dt])
        elif rk < 3:
            plan.lincomb(u2, [u2, rhs], [1.0, a[rk] * dt])
        else:
            plan.lincomb(u0, [u2, rhs], [1.0, a[rk] * dt])
        if rk < 3:
            plan.lincomb(u1, [u0, rhs], [1.0, b[rk] * dt])
            y = u1
    return u0, dt, dt


def ForwardEuler(u0, rhs, dt, solver, context):
    """u0 += rhs*dt (integrators.py:161-165)."""
    rhs = solver.ComputeRHS(rhs, u0, solver, **context)
    context.VT.plan.lincomb(u0, [u0, rhs], [1.0, dt])
    return u0, dt, dt


def AB2(u0, u1, rhs, dt, tstep, solver, context):
    """Adams-Bashforth 2, ForwardEuler on the first step; u1 keeps rhs*dt
    (integrators.py:167-175)."""
    plan = context.VT.plan
    rhs = solver.ComputeRHS(rhs, u0, solver, **context)
    if tstep == 0:
        plan.lincomb(u0, [u0, rhs], [1.0, dt])
    else:
        plan.lincomb(u0, [u0, rhs, u1], [1.0, 1.5 * dt, -0.5])
    plan.lincomb(u1, [rhs], [dt])
    return u0, dt, dt


def adaptiveRK(A, b, bhat, err_order, fY_hat, u0_new, sc, err, fsal, offset, aTOL, rTOL, adaptive, errnorm,
               rhs, u0, solver, dt, tstep, context, additional_callback, params, predictivecontroller=False):
    """One step of an embedded Runge-Kutta method (integrators.py:15-147).

    Same arguments as the reference; ``sc`` is the stage-value work array (the
    scale array itself is formed inside the error-norm kernel).  Only errnorm
    "2" and the plain controller are built, as in the reference."""
    if errnorm != "2":
        raise NotImplementedError("errnorm %r (the reference asserts on 'inf')" % errnorm)
    if predictivecontroller:
        raise NotImplementedError("the predictive controller is disabled in the reference")
    plan = context.VT.plan
    s = A.shape[0]
    fl = context.float
    nspec = np.sqrt(np.array(plan.spectral_shape).prod())
    facmax, fac, facmin = 2, 0.8, 0.01
    while True:
        dt_prev = dt
        if fsal:
            offset[0] = (offset[0] - 1) % s
        for i in range(s):
            if not fsal or (tstep == 0 or i != 0):
                terms = [u0] + [fY_hat[(j + offset[0]) % s] for j in range(i)]
                coefs = [1.0] + [fl(dt * A[i, j]) for j in range(i)]
                plan.lincomb(sc, terms, coefs)
                solver.ComputeRHS(fY_hat[(i + offset[0]) % s], sc, solver, **context)
            if i == 0:
                context.fu0 = fY_hat[(0 + offset[0]) % s]
                additional_callback(context)
        stages = [fY_hat[(j + offset[0]) % s] for j in range(s)]
        plan.lincomb(u0_new, [u0] + stages, [1.0] + [fl(dt * b[j]) for j in range(s)])
        plan.lincomb(err, stages, [fl(dt * (b[j] - bhat[j])) for j in range(s)])
        nsq = plan.error_norm(err, u0, u0_new, aTOL, rTOL).cpu().numpy()
        est = float(np.max(np.sqrt(nsq))) / nspec
        exponent = 1.0 / (err_order + 1)
        factor = min(facmax, max(facmin, fac * pow((1 / est), exponent)))
        if adaptive:
            dt = dt * factor
            if est > 1.0:
                facmax = 1
                context.is_step_rejected_callback = True
                context.dt_rejected = dt_prev
                additional_callback(context)
                if fsal:
                    offset[0] += 1
                continue
        break
    u0.copy_(u0_new)
    return u0, dt, dt_prev


def getintegrator(rhs, u0, solver, context):
    """Return the integrator chosen by params.integrator (integrators.py:177-239)."""
    params = solver.params
    if params.integrator not in ("RK4", "ForwardEuler", "AB2", "BS5_adaptive", "BS5_fixed"):
        raise NotImplementedError("integrator %r" % params.integrator)
    if params.get("compensated", False) and params.integrator != "RK4":
        raise NotImplementedError("the compensated state is carried by the fused NS RK4 step only "
                                  "(integrator %r)" % params.integrator)
    if params.integrator == "RK4" and not (getattr(solver.conv, "fused_vortex", False) and
                                           solver.ComputeRHS is getattr(solver, "_ComputeRHS_fused", None)):
        if params.get("compensated", False):
            raise NotImplementedError("the compensated state is carried by the fused NS RK4 step only")
        a = np.array([1. / 6., 1. / 3., 1. / 3., 1. / 6.], dtype=context.float)
        b = np.array([0.5, 0.5, 1.], dtype=context.float)
        u1, u2 = u0.clone(), u0.clone()

        @wraps(RK4_stages)
        def func():
            return RK4_stages(u0, u1, u2, rhs, a, b, params.dt, solver, context)
        return func

    if params.integrator == "RK4":
        if params.get("compensated", False) and params.precision == "single" and "U_hat_lo" not in context:
            import torch
            context["U_hat_lo"] = torch.zeros_like(u0)
        a = np.array([1. / 6., 1. / 3., 1. / 3., 1. / 6.], dtype=context.float)
        b = np.array([0.5, 0.5, 1.], dtype=context.float)

        @wraps(RK4)
        def func():
            return RK4(u0, None, None, rhs, a, b, params.dt, solver, context)
        return func

    u1 = u0.clone()
    if params.integrator in ("BS5_adaptive", "BS5_fixed"):
        import torch
        A = np.array(_BS5_A, dtype=context.float)
        b = np.array(_BS5_B, dtype=context.float)
        bhat = np.array(_BS5_BHAT, dtype=context.float)
        err_order, errnorm, fsal = 4, "2", True
        adaptive = params.integrator == "BS5_adaptive"
        offset = [0]
        s = A.shape[0]
        fY_hat = torch.zeros((s,) + tuple(u0.shape), dtype=u0.dtype, device=u0.device)
        sc = torch.zeros_like(u0)
        err = torch.zeros_like(u0)

        @wraps(adaptiveRK)
        def func():
            return adaptiveRK(A, b, bhat, err_order, fY_hat, u1, sc, err, fsal, offset, params.TOL, params.TOL,
                              adaptive, errnorm, rhs, u0, solver, params.dt, params.tstep, context,
                              solver.additional_callback, params)
        return func

    if params.integrator == "ForwardEuler":
        @wraps(ForwardEuler)
        def func():
            return ForwardEuler(u0, rhs, params.dt, solver, context)
        return func

    if params.integrator == "AB2":
        @wraps(AB2)
        def func():
            return AB2(u0, u1, rhs, params.dt, params.tstep, solver, context)
        return func

    raise NotImplementedError("integrator %r" % params.integrator)
