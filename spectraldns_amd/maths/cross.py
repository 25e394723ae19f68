"""Cross products (spectralDNS/maths/cross.py:13-32), bound to the HIP kernels."""
from ..optimization import optimizer

__all__ = ["cross1", "cross2"]


@optimizer
def cross1(c, a, b):
    """Regular c = a x b, where a and b are real."""


@optimizer
def cross2(c, a, b):
    """c = 1j*(a x b), where a is the wavenumber list and b is complex."""
