"""MI355X-native triply periodic Navier-Stokes RK4 hot path of spectralDNS.

Host API of the reference (spectralDNS/__init__.py): ``config``,
``get_solver`` and the generic ``solve`` loop.  The solver module
``spectraldns_amd.solvers.NS`` keeps the reference's context keys and hooks;
its transforms, curl, cross product, dealiasing, pressure projection and RK4
stages run as hand-written gfx950 HIP kernels in ``libsdns_hip.so``.
"""
import importlib

from . import config

__version__ = "0.1.0"
__all__ = ["config", "get_solver", "solve"]


def get_solver(update=None, regression_test=None, additional_callback=None,
               mesh="triplyperiodic", parse_args=None):
    """Return the solver module chosen by the parsed configuration.

    Follows spectralDNS/__init__.py:19-66: parse argv (or ``parse_args``) into
    ``config.params``, import ``spectraldns_amd.solvers.<solver>``, install the
    optional hooks and link the module as ``config.solver``.
    """
    assert parse_args is None or isinstance(parse_args, list)
    if mesh != "triplyperiodic":
        raise NotImplementedError("only the triply periodic mesh is built")
    args = getattr(config, mesh).parse_args(parse_args)
    config.params.update(vars(args))
    name = config.params.solver
    if name not in ("NS", "VV", "MHD"):
        raise NotImplementedError("solver %r is not built; the HIP path serves NS, VV and MHD" % name)
    solver = importlib.import_module("spectraldns_amd.solvers." + name)
    if update:
        solver.update = update
    if regression_test:
        solver.regression_test = regression_test
    if additional_callback:
        solver.additional_callback = additional_callback
    config.solver = solver
    return solver


def solve(solver, context):
    """Generic time loop (spectralDNS/__init__.py:69-123), same hook order:
    integrate, t += dt, tstep += 1, update, hdf5file.update, timer, profiler,
    end_of_tstep; then timer.final, profile, regression_test, file close."""
    solver.timer = solver.Timer()
    params = solver.params
    solver.conv = solver.getConvection(params.convection)
    integrate = solver.getintegrator(context.dU, context.u, solver, context)
    dt_in = params.dt

    while params.t + params.dt <= params.T + 1e-12:
        u, params.dt, dt_took = integrate()
        params.t += dt_took
        params.tstep += 1
        solver.update(context)
        context.hdf5file.update(params, **context)
        solver.timer()
        if not solver.profiler.getstats() and params.make_profile:
            solver.profiler.enable()
        if solver.end_of_tstep(context):
            break

    params.dt = dt_in
    solver.timer.final(params.verbose)
    if params.make_profile:
        solver.results = solver.create_profile(solver.profiler)
    solver.regression_test(context)
    context.hdf5file.close()
