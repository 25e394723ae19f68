"""Kernel plug-in layer (spectralDNS/optimization/__init__.py:12-55).

The reference's ``@optimizer`` swaps a NumPy function for a compiled one
named ``name``, ``name_<solver>`` or ``name_<mesh>`` in the module
``{optimization}_{precision}`` and silently falls back to NumPy.  Here the
only backend is the HIP library: the decorator binds ``name`` (or
``name_NS``) from ``spectraldns_amd.optimization.hip`` and there is no
fallback -- a function without a HIP implementation raises when called.
"""
from functools import wraps

from . import hip as _hip

__all__ = ["optimizer", "wraps"]


def optimizer(func):
    name = func.__name__
    impl = getattr(_hip, name, None) or getattr(_hip, name + "_NS", None)
    if impl is None:
        @wraps(func)
        def missing(*args, **kwargs):
            raise NotImplementedError("%s has no HIP implementation" % name)
        return missing

    @wraps(func)
    def wrapped(*args, **kwargs):
        return impl(*args, **kwargs)
    return wrapped
