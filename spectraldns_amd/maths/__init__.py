"""cross products, projection and time integrators (spectralDNS/maths)."""
from .cross import cross1, cross2
from .maths import project
from .integrators import getintegrator

__all__ = ["cross1", "cross2", "project", "getintegrator"]
