"""Step timer, memory report and profile helpers (spectralDNS/utilities/__init__.py:18-68).

The GPU works asynchronously.  On a GPU the timer records a HIP event on the
current stream at every step and never synchronises inside the time loop: a
step's time is the GPU time between consecutive events, folded into the
fastest/slowest values as the events complete (a non-blocking query), and
``final`` waits for the last ones.  Without a GPU it reads the wall clock.
"""
import pstats
from time import time

__all__ = ["Timer", "MemoryUsage", "create_profile", "reset_profile"]


def _cuda():
    try:
        import torch
        if torch.cuda.is_available() and torch.cuda.is_initialized():
            return torch
    except ImportError:
        pass
    return None


class Timer:
    """Fastest and slowest complete time step (utilities/__init__.py:18-68)."""

    def __init__(self):
        self.fastest_timestep = 1e8
        self.slowest_timestep = 0
        self.t0 = time()
        self.tic = self.t0
        self._torch = _cuda()
        self._pending = []
        self._last = None
        if self._torch is not None:
            self._last = self._torch.cuda.Event(enable_timing=True)
            self._last.record()

    def _sample(self, dt):
        self.fastest_timestep = min(dt, self.fastest_timestep)
        self.slowest_timestep = max(dt, self.slowest_timestep)

    def _drain(self, block):
        while self._pending and (block or self._pending[0][1].query()):
            a, b = self._pending.pop(0)
            b.synchronize()
            self._sample(a.elapsed_time(b) / 1000.0)

    def __call__(self):
        if self._torch is None:
            t1 = time()
            self._sample(t1 - self.t0)
            self.t0 = t1
            return
        e = self._torch.cuda.Event(enable_timing=True)
        e.record()
        self._pending.append((self._last, e))
        self._last = e
        self._drain(block=False)

    def final(self, verbose=True):
        """Print total time and the (fastest, slowest) pairs; one process, so the
        reference's min/max over ranks reduce to the local values."""
        self._drain(block=True)
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
