"""Velocity-vorticity formulation, triply periodic box (spectralDNS/solvers/VV.py).

The primary variable is the vorticity W_hat.  The module keeps the
reference's protocol on top of the NS module: ``get_context`` (the NS context
without P and P_hat, with W_hat and u = W_hat), ``compute_velocity``,
``get_velocity``, ``get_curl``, ``get_divergence``, ``getConvection``,
``add_linear`` and ``ComputeRHS``.

The Vortex convection F(u x w) runs on the fused convection passes with the
vorticity given: the velocity u_hat = 1j*(K/K^2) x W_hat (one kernel) and
W_hat itself are transformed back together (sdns_convection_uw), so w is the
reference's w_dealias = VTp.backward(w_hat) for any W_hat, including one with
a mean or a gradient part (a random start, a Source with a mean).  K2 and
K_over_K2 are handles that carry the plan; the kernels form K^2 on the fly,
and indexing a handle materialises the array.
"""
from .NS import *  # noqa: F401,F403
from .NS import config, params, HDF5File, optimizer, cross2, end_of_tstep  # noqa: F401
from .NS import get_context as NS_context
from .NS import set_velocity  # noqa: F401
from ..fourier import Function


class _WavenumberField:
    """K2 or K_over_K2 of a context: carries the plan for the kernels, which
    form the values themselves; indexing or .tensor() builds the array."""

    def __init__(self, context, key):
        self._context, self._key = context, key
        self.plan = context.VT.plan
        self.k_over_k2 = key == "K_over_K2"
        self._val = None

    def tensor(self):
        if self._val is None:
            import torch
            K = self._context["K"]
            k2 = K[0] * K[0]
            k2 = k2 + K[1] * K[1]
            k2 = k2 + K[2] * K[2]
            if self._key == "K2":
                self._val = k2
            else:
                den = torch.where(k2 == 0, torch.ones_like(k2), k2)
                self._val = torch.stack([(K[i] / den).expand_as(k2) for i in range(3)])
        return self._val

    def __getitem__(self, i):
        return self.tensor()[i]

    @property
    def shape(self):
        return self.tensor().shape


def get_context():
    """Context of the VV solver (VV.py:23-39)."""
    c = NS_context()
    del c["P"], c["P_hat"]
    c["W_hat"] = Function(c.VT)
    c["u"] = c["W_hat"]
    c["K2"] = _WavenumberField(c, "K2")
    c["K_over_K2"] = _WavenumberField(c, "K_over_K2")
    c["hdf5file"] = VVFile(config.params.solver,
                           checkpoint={"space": c.VT, "data": {"0": {"curl": [c.W_hat]}}},
                           results={"space": c.VT, "data": {"U": [c.U], "curl": [c.curl]}})
    return c


class VVFile(HDF5File):
    """Results file that transforms U and the curl before storing (VV.py:41-53)."""

    def update_components(self, **context):
        get_velocity(**context)
        get_curl(**context)


def compute_velocity(U, w_hat, work, VT, K_over_K2):
    """u = VT.backward(1j*(K/K^2) x w_hat) (VV.py:57-73)."""
    v_hat = work[(w_hat, 1, True)]
    v_hat = cross2(v_hat, K_over_K2, w_hat)
    return VT.backward(v_hat, U)


def get_velocity(W_hat, U, work, VT, K_over_K2, **context):
    """Velocity from the vorticity (VV.py:75-78)."""
    return compute_velocity(U, W_hat, work, VT, K_over_K2)


def get_divergence(T, K, U_hat, W_hat, **context):
    """div of 1j*(K x W_hat) (VV.py:80-84)."""
    import torch
    U_hat = cross2(U_hat, K, W_hat)
    div_hat = torch.empty(T.shape(True), dtype=T.plan.complex_dtype, device=T.plan.device)
    T.plan.divergence(div_hat, U_hat)
    return T.backward(div_hat)


def get_curl(curl, W_hat, VT, **context):
    """curl = VT.backward(W_hat) (VV.py:86-89)."""
    return VT.backward(W_hat, curl)


def getConvection(convection):
    """Return Conv(rhs, w_hat, work, Tp, VTp, K, K_over_K2, u_dealias) (VV.py:91-108)."""
    if convection in ("Standard", "Divergence", "Skewed"):
        raise NotImplementedError
    if convection != "Vortex":
        raise NotImplementedError("convection %r" % convection)

    def Conv(rhs, w_hat, work, Tp, VTp, K, K_over_K2, u_dealias=None):
        v_hat = _cross_convection(w_hat, work, VTp, K_over_K2)
        rhs = cross2(rhs, K, v_hat)  # rhs = 1j*(K x v_hat)
        return rhs

    Conv.convection = convection
    Conv.vv_fused_tail = True
    return Conv


def _cross_convection(w_hat, work, VTp, K_over_K2):
    """v_hat = F_k(u_dealias x w_dealias) with u_hat = 1j*(K/K^2) x w_hat and
    w_dealias = VTp.backward(w_hat) (VV.py:95-97), in the dealiased
    convection passes.  Both work arrays are written whole by the next
    launch, so they are not zero-filled."""
    u_hat = work[(w_hat, 1, False)]
    u_hat = cross2(u_hat, K_over_K2, w_hat)
    v_hat = work[(w_hat, 0, False)]
    return VTp.plan.convection_uw(u_hat, w_hat, v_hat)


@optimizer
def add_linear(rhs, w_hat, nu, K2, Source):
    """Add contributions from source and diffusion to the rhs (VV.py:110-115)."""


def ComputeRHS(rhs, w_hat, solver, **context):
    """rhs = conv(w_hat), Nyquist mask, -nu K^2 w_hat, + Source (VV.py:117-146).

    The Source is always added, as the reference does."""
    c = context
    src = c["Source"]
    if getattr(solver.conv, "vv_fused_tail", False) and solver.add_linear is add_linear:
        # the module's own Conv and add_linear: cross2, mask and add_linear in one launch
        v_hat = _cross_convection(w_hat, c["work"], c["VTp"], c["K_over_K2"])
        return c["VT"].plan.vv_rhs_tail(rhs, v_hat, w_hat, params.nu, src)
    rhs = solver.conv(rhs, w_hat, c["work"], c["Tp"], c["VTp"], c["K"], c["K_over_K2"], None)
    if c["mask"] is not None:
        c["VT"].plan.mask_nyquist(rhs)
    rhs = solver.add_linear(rhs, w_hat, params.nu, c["K2"], src)
    return rhs
