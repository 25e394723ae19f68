"""Divergence-free projection (spectralDNS/maths/maths.py:8-11)."""

__all__ = ["project"]


def project(u, K, K_over_K2=None):
    """u -= sum(K_over_K2*u, 0)*K, in place on the GPU.

    The HIP pressure kernel with nu = 0 computes exactly this projection
    (P = sum u K/K^2, u -= P K)."""
    plan = getattr(K, "plan", None)
    if plan is None:
        raise ValueError("project needs the wavenumber list K of a context")
    return plan.add_pressure_diffusion(u, u, 0.0, None)
