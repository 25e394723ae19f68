"""Step timer, memory report and profile helpers (spectralDNS/utilities/__init__.py:18-68).

The GPU works asynchronously, so the timer synchronises the device before it
reads the clock: a step's time is the wall time until its kernels finished.
"""
import pstats
from time import time

__all__ = ["Timer", "MemoryUsage", "create_profile", "reset_profile"]


def _sync():
    try:
        import torch
        if torch.cuda.is_available() and torch.cuda.is_initialized():
            torch.cuda.synchronize()
    except ImportError:
        pass


class Timer:
    """Fastest and slowest complete time step (utilities/__init__.py:18-68)."""

    def __init__(self):
        self.fastest_timestep = 1e8
        self.slowest_timestep = 0
        _sync()
        self.t0 = time()
        self.tic = self.t0

    def __call__(self):
        _sync()
        t1 = time()
        dt = t1 - self.t0
        self.fastest_timestep = min(dt, self.fastest_timestep)
        self.slowest_timestep = max(dt, self.slowest_timestep)
        self.t0 = t1

    def final(self, verbose=True):
        """Print total time and the (fastest, slowest) pairs; one process, so the
        reference's min/max over ranks reduce to the local values."""
        toc = time() - self.tic
        fast = (self.fastest_timestep, self.slowest_timestep)
        if verbose:
            print("Time = {}".format(toc))
            print("Fastest = {}".format(fast))
            print("Slowest = {}".format(fast))


def MemoryUsage(label=""):
    """Report device memory in use (the reference reports process RSS)."""
    try:
        import torch
        if torch.cuda.is_available():
            return torch.cuda.memory_allocated() / 2 ** 20
    except ImportError:
        pass
    return 0.0


_PROFILE_KEYS = ("compute_rhs", "rk4_steps", "forward", "backward", "ComputeRHS", "RK4",
                 "cross1", "cross2", "add_pressure_diffusion", "Conv")


def create_profile(profiler):
    """Cumulative times of the path's entry points from a cProfile profile
    (utilities/create_profile.py:10-74 keeps a fixed list of names)."""
    profiler.disable()
    st = pstats.Stats(profiler)
    out = {}
    for (fname, line, func), (cc, nc, tt, ct, callers) in st.stats.items():
        if func in _PROFILE_KEYS:
            out[func] = (nc, ct)
    return out


def reset_profile(profiler):
    profiler.disable()
    profiler.clear() if hasattr(profiler, "clear") else None
