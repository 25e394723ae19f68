"""Shared prelude of the solver modules (spectralDNS/solvers/spectralinit.py:19-62):
parameters, data types, default hooks and the (single-process) communicator."""
import cProfile

import numpy as np

from .. import config
from ..utilities import create_profile, MemoryUsage, Timer, reset_profile
from ..h5io import HDF5File
from ..optimization import optimizer
from ..maths import cross1, cross2, project, getintegrator


class SerialComm:
    """The one-process communicator: reductions return their argument
    (one GPU holds the whole box)."""

    def Get_rank(self):
        return 0

    def Get_size(self):
        return 1

    def reduce(self, x, op=None, root=0):
        return x

    def allreduce(self, x, op=None):
        return x

    def bcast(self, x, root=0):
        return x

    def Barrier(self):
        pass


comm = SerialComm()
num_processes = 1
rank = 0
params = config.params
profiler = cProfile.Profile()


def datatypes(precision):
    """(float, complex, mpitype) of the working precision (spectralinit.py:25-29)."""
    assert precision in ("single", "double")
    return {"single": (np.float32, np.complex64, None),
            "double": (np.float64, np.complex128, None)}[precision]


def regression_test(context):
    """Hook run once after the last time step (a no-op unless the user sets one)."""


def update(context):
    """Hook run after each time step (a no-op unless the user sets one)."""


def additional_callback(context):
    """Hook the adaptive integrator calls after a stage or a rejected step."""


def solve_linear(context):
    """Linear-solve hook of implicit schemes (none of the built solvers needs it)."""


def conv(*args):
    """Placeholder for the nonlinear term; solve() installs getConvection's result."""
    raise NotImplementedError


def set_source(Source, **context):
    Source[:] = 0
    return Source


def end_of_tstep(context):
    return False
