"""Parameters of the triply periodic NS solver.

Same interface as spectralDNS/config.py for the path this package serves:
``params`` (an attribute dictionary), the argparse tree ``parser`` ->
``triplyperiodic`` -> solver sub-command, and ``update(new, mesh)``.

Semantics kept from the reference:
  * ``params.N`` is ``2**params.M`` unless N was given (config.py:115-118).
  * attribute reads of nu, dt, Ri, Pr, eta are cast to float32 in single
    precision and float64 otherwise (config.py:123-127); item reads are not.
  * M and N are stored as read-only int arrays, L as a read-only float array
    whose entries may be expressions of ``pi`` (config.py:136-146).
  * ``update`` sets parser defaults and merges ``planner_effort``
    (config.py:307-314).

Additions: ``--optimization hip`` (the only kernel backend; the reference's
choices are accepted and also run the HIP kernels) and ``--dealias_mode``
(0: shenfun ``dealias_direct`` semantics, 1: output mask |k| < 2/3 (N/2+1)).
"""
import argparse
import ast
import json
import operator
from collections import defaultdict

import numpy as np

__all__ = ["AttributeDict", "Params", "params", "parser", "triplyperiodic", "update"]

_CAST = ("nu", "dt", "Ri", "Pr", "eta")


class AttributeDict(dict):
    """dict whose items are also attributes (config.py:54-93)."""

    def __getattr__(self, key):
        try:
            return self[key]
        except KeyError:
            raise AttributeError(key) from None

    def __setattr__(self, key, val):
        self[key] = val

    def __delattr__(self, key):
        try:
            del self[key]
        except KeyError:
            raise AttributeError(key) from None


_OPS = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul,
        ast.Div: operator.truediv, ast.Pow: operator.pow, ast.USub: operator.neg,
        ast.UAdd: operator.pos}


def _eval_length(expr):
    """Evaluate a domain length such as '2*pi' or 6.283 (numbers, pi, + - * / **)."""
    if isinstance(expr, (int, float, np.floating, np.integer)):
        return float(expr)

    def ev(node):
        if isinstance(node, ast.Expression):
            return ev(node.body)
        if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
            return float(node.value)
        if isinstance(node, ast.Name) and node.id == "pi":
            return np.pi
        if isinstance(node, ast.BinOp) and type(node.op) in _OPS:
            return _OPS[type(node.op)](ev(node.left), ev(node.right))
        if isinstance(node, ast.UnaryOp) and type(node.op) in _OPS:
            return _OPS[type(node.op)](ev(node.operand))
        raise ValueError("unsupported length expression %r" % expr)

    return float(ev(ast.parse(str(expr), mode="eval")))


class Params(AttributeDict):
    """Solver parameters with the reference's derived values and casts."""

    def __getattr__(self, key):
        if dict.__contains__(self, key):
            val = dict.__getitem__(self, key)
            if key in _CAST:
                single = dict.get(self, "precision") == "single"
                return (np.float32 if single else np.float64)(val)
            return val
        if key == "N":
            if not dict.__contains__(self, "M"):
                raise AttributeError("N")
            return 2 ** dict.__getitem__(self, "M")
        if key == "dx":
            return self.L / self.N
        raise AttributeError(key)

    def __setitem__(self, key, val):
        if key in ("M", "N"):
            val = np.array([int(str(v)) for v in val], dtype=int)
            val.flags.writeable = False
        elif key == "L":
            val = np.array([_eval_length(v) for v in val], dtype=float)
            val.flags.writeable = False
        dict.__setitem__(self, key, val)

    def update(self, other=(), **kw):
        for k, v in dict(other, **kw).items():
            self[k] = v


fft_plans = defaultdict(lambda: "FFTW_MEASURE", {"dct": "FFTW_MEASURE"})


class PlanAction(argparse.Action):
    """--planner_effort JSON (kept for command-line compatibility; the HIP
    transforms take their schedule from the plan, not from FFTW flags)."""

    def __call__(self, parser_, namespace, values, option_string=None):
        fft_plans.update(json.loads(values))
        setattr(namespace, self.dest, fft_plans)


params = Params()

parser = argparse.ArgumentParser(prog="spectraldns_amd", add_help=False)
parser.add_argument("--precision", default="double", choices=("single", "double"))
parser.add_argument("--optimization", default="hip",
                    choices=("hip", "", "cython", "weave", "numba", "pythran"),
                    help="Kernel backend (every choice runs the HIP kernels)")
parser.add_argument("--make_profile", default=0, type=int, help="1: profile the run with cProfile")
parser.add_argument("--dt", default=0.01, type=float, help="dt of the time loop (the last step may be shortened to end at T)")
parser.add_argument("--T", default=0.1, type=float, help="simulated time at which the loop stops")
parser.add_argument("--write_result", default=1e8, metavar=("tstep"), type=int,
                    help="steps between writes of the physical results")
parser.add_argument("--checkpoint", default=1e8, type=int, help="steps between checkpoints of the spectral state")
parser.add_argument("--nu", default=0.000625, type=float, help="kinematic viscosity nu")
parser.add_argument("--t", default=0.0, type=float, help="simulated time at the start")
parser.add_argument("--tstep", default=0, type=int, help="step counter at the start")
parser.add_argument("--filemode", default="w", choices=("w", "r", "a"), help="open mode of the results files")
parser.add_argument("--dealias", default="2/3-rule", choices=("2/3-rule", "3/2-rule", "None"),
                    help="dealiasing of the nonlinear term: 2/3 truncation, 3/2 padding or none")
parser.add_argument("--dealias_mode", default=0, type=int, choices=(0, 1),
                    help="2/3-rule variant: 0 truncate inputs |k| < N/3, 1 output mask |k| < 2/3(N/2+1)")
parser.add_argument("--compensated", dest="compensated", action="store_true",
                    help="single precision: carry u_hat as a compensated (hi, lo) pair in RK4")
parser.set_defaults(compensated=False)
parser.add_argument("--decomposition", default="slab", choices=("slab", "pencil"),
                    help="Kept for compatibility (one GPU holds the whole box)")
parser.add_argument("--ntol", default=7, type=int, help="decimal places the regression checks compare")
parser.add_argument("--threads", default=1, type=int, help="Kept for compatibility")
parser.add_argument("--planner_effort", action=PlanAction, default=fft_plans,
                    help="Kept for compatibility")
parser.add_argument("--h5filename", default="results", type=str, help="base name of the results files")
parser.add_argument("--verbose", dest="verbose", action="store_true", help="report the step timings after the run")
parser.add_argument("--no-verbose", dest="verbose", action="store_false", help="no timing report after the run")
parser.set_defaults(verbose=True)
parser.add_argument("--mask_nyquist", dest="mask_nyquist", action="store_true",
                    help="zero the Nyquist planes of the right-hand side")
parser.add_argument("--no-mask_nyquist", dest="mask_nyquist", action="store_false",
                    help="keep the Nyquist planes of the right-hand side")
parser.set_defaults(mask_nyquist=True)

triplyperiodic = argparse.ArgumentParser(parents=[parser])
triplyperiodic.add_argument("--convection", default="Vortex",
                            choices=("Standard", "Divergence", "Skewed", "Vortex"),
                            help="Nonlinear term (the HIP path implements Vortex)")
triplyperiodic.add_argument("--L", default=[2 * np.pi, 2 * np.pi, 2 * np.pi], metavar=("Lx", "Ly", "Lz"),
                            nargs=3, help="box lengths (expressions in pi allowed, e.g. 2*pi)")
triplyperiodic.add_argument("--M", default=[6, 6, 6], metavar=("Mx", "My", "Mz"), nargs=3,
                            help="log2 of the grid points per axis (N_i = 2**M_i) when N is not given")
triplyperiodic.add_argument("--TOL", type=float, default=1e-6, help="error tolerance of the adaptive step-size control")
triplyperiodic.add_argument("--integrator", default="RK4",
                            choices=("RK4", "ForwardEuler", "AB2", "BS5_adaptive", "BS5_fixed"),
                            help="time integration scheme")
_sub = triplyperiodic.add_subparsers(dest="solver")
_sub.add_parser("NS", help="Navier-Stokes in velocity form")
_sub.add_parser("VV", help="Navier-Stokes in velocity-vorticity form")
_mhd = _sub.add_parser("MHD", help="Magnetohydrodynamics solver (2/3 rule or no dealiasing)")
_mhd.add_argument("--eta", default=0.01, type=float, help="magnetic diffusivity eta")
_bq = _sub.add_parser("Bq", help="Boussinesq (not built)")
_bq.add_argument("--Ri", default=0.1, type=float, help="Richardson number Ri")
_bq.add_argument("--Pr", default=1.0, type=float, help="Prandtl number Pr")


def update(new, mesh="triplyperiodic"):
    """Set parser defaults of ``mesh`` (config.py:307-314)."""
    assert isinstance(new, dict)
    if mesh != "triplyperiodic":
        raise NotImplementedError("only the triply periodic mesh is built")
    if "planner_effort" in new:
        fft_plans.update(new["planner_effort"])
        new["planner_effort"] = fft_plans
    globals()[mesh].set_defaults(**new)
