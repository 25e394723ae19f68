"""Classical pseudo-spectral Navier-Stokes solver, triply periodic box
(spectralDNS/solvers/NS.py).

Same module protocol as the reference: ``get_context``, ``getConvection``,
``ComputeRHS``, ``add_pressure_diffusion``, the accessors ``get_velocity``,
``set_velocity``, ``get_pressure``, ``get_curl``, ``get_divergence`` and
``end_of_tstep``; the context keeps the reference's keys.  Arrays are torch
tensors on the GPU and every numerical step is a libsdns_hip launch.

K2, K_over_K2 and the padded u_dealias are never needed by the kernels (the
wavenumbers are generated inside them); the context builds them on first
access only.
"""
from .spectralinit import *  # noqa: F401,F403
from .spectralinit import config, params, datatypes, HDF5File, optimizer, getintegrator  # noqa: F401
from .spectralinit import cross1, cross2, Timer, MemoryUsage, create_profile, profiler, comm, rank  # noqa: F401
from .._abi import Plan
from ..fourier import FourierSpace, Array, Function, WorkArrays


class Context(config.AttributeDict):
    """Attribute dictionary whose K2, K_over_K2 and u_dealias entries are built
    on first access."""

    def __missing__(self, key):
        import torch
        VT = dict.__getitem__(self, "VT")
        K = dict.__getitem__(self, "K")
        if key == "K2":
            val = K[0] * K[0]
            val = val + K[1] * K[1]
            val = val + K[2] * K[2]
        elif key == "K_over_K2":
            K2 = self["K2"]
            den = torch.where(K2 == 0, torch.ones_like(K2), K2)
            val = torch.stack([(K[i] / den).expand_as(K2) for i in range(3)])
        elif key == "u_dealias":
            val = Array(dict.__getitem__(self, "VTp"))
        else:
            raise KeyError(key)
        dict.__setitem__(self, key, val)
        return val


def get_context():
    """Set up the context of the classical NS solver (NS.py:12-72)."""
    float, complex, mpitype = datatypes(params.precision)
    plan = Plan(params.N, params.L, params.precision, params.dealias, params.mask_nyquist,
                int(params.get("dealias_mode", 0)))
    T = FourierSpace(plan, 0)
    VT = FourierSpace(plan, 1)
    Tp = FourierSpace(plan, 0, dealiased=True)
    VTp = FourierSpace(plan, 1, dealiased=True)
    mask = T.get_mask_nyquist() if params.mask_nyquist else None
    X = T.local_mesh(True)
    K = T.local_wavenumbers(scaled=True)
    U = Array(VT)
    U_hat = Function(VT)
    P = Array(T)
    P_hat = Function(T)
    u = U_hat
    dU = Function(VT)
    curl = Array(VT)
    Source = Function(VT)
    work = WorkArrays()
    hdf5file = NSFile(config.params.solver,
                      checkpoint={"space": VT, "data": {"0": {"U": [U_hat]}}},
                      results={"space": VT, "data": {"U": [U], "P": [P]}})
    ctx = dict(locals())
    return Context(ctx)


class NSFile(HDF5File):
    """Results file that transforms U and P to physical space before storing."""

    def update_components(self, **context):
        get_velocity(**context)
        get_pressure(**context)


def get_curl(curl, U_hat, work, VT, K, **context):
    """curl = VT.backward(1j*(K x U_hat)) (NS.py:86-89)."""
    return compute_curl(curl, U_hat, work, VT, K)


def get_velocity(U, U_hat, VT, **context):
    """U = VT.backward(U_hat) (NS.py:91-94)."""
    return VT.backward(U_hat, U)


def get_pressure(P, P_hat, T, **context):
    """P = T.backward(-1j*P_hat) (NS.py:96-99)."""
    tmp = T.plan.mul_minus_i(P_hat.clone(), P_hat)
    return T.backward(tmp, P)


def set_velocity(U, U_hat, VT, **context):
    """U_hat = VT.forward(U) (NS.py:101-104)."""
    return VT.forward(U, U_hat)


def get_divergence(T, K, U_hat, mask, **context):
    """div u = T.backward(1j*(K . U_hat)) (NS.py:106-110)."""
    import torch
    div_hat = torch.empty(T.shape(True), dtype=T.plan.complex_dtype, device=T.plan.device)
    T.plan.divergence(div_hat, U_hat)
    return T.backward(div_hat)


def end_of_tstep(context):
    """Make the last step hit T exactly (NS.py:112-122)."""
    if abs(params.t - params.T) < 1e-12:
        return True
    if (abs(params.t + params.dt - params.T) < 1e-12 or
            params.t + params.dt >= params.T + 1e-12):
        params.dt = params.T - params.t
    return False


def compute_curl(c, a, work, T, K):
    """c = F^-1(1j*K x a) (NS.py:124-129)."""
    curl_hat = work[(a, 0, False)]
    curl_hat = cross2(curl_hat, K, a)
    return T.backward(curl_hat, c)


def Cross(c, a, b, work, T):
    """c_k = F_k(a x b) (NS.py:131-136)."""
    d = work[(a, 1, False)]
    d = cross1(d, a, b)
    return T.forward(d, c)


def _dealiased_backward(plan, u_hat, u):
    """VTp.backward(u_hat): the inverse transform of the kept modes (2/3 rule or
    no dealiasing, truncation inside the transform passes) or onto the padded
    3N/2 grid (3/2 rule)."""
    return plan.backward_dealiased(u_hat, u)


def standard_convection(rhs, u_dealias, U_hat, work, Tp, K):
    """rhs_i = Tp.forward(u_j du_i/dx_j) (NS.py:138-145): per component, the
    truncated gradient 1j*K_j*U_hat_i, its dealiased inverse transform, the dot
    product with u_dealias and the dealiased forward transform."""
    plan = K.plan
    gradUi = work[(u_dealias, 1, False)]
    g_hat = work[(U_hat, 2, False)]
    s = work[(u_dealias[0], 0, False)]
    for i in range(3):
        plan.gradient_truncated(g_hat, U_hat[i])
        plan.backward_dealiased(g_hat, gradUi)
        plan.dot3(s, u_dealias, gradUi)
        plan.forward_dealiased(s, rhs[i])
    return rhs


def divergence_convection(rhs, u_dealias, work, Tp, K, add=False, scale=1.0):
    """rhs_i (+)= scale * 1j*sum_j K_j F(u_i u_j) (NS.py:147-165): one kernel for
    the six products, their forward transforms, one kernel for the sums."""
    import torch
    plan = K.plan
    bufs = getattr(plan, "_uu_buffers", None)
    if bufs is None:
        bufs = plan._uu_buffers = (
            torch.empty((6,) + tuple(plan.dealiased_shape), dtype=plan.real_dtype, device=plan.device),
            torch.empty((6,) + tuple(plan.spectral_shape), dtype=plan.complex_dtype, device=plan.device))
    uu, uu_hat = bufs
    plan.uu_products(uu, u_dealias)
    plan.forward_dealiased(uu[0:3], uu_hat[0:3])
    plan.forward_dealiased(uu[3:6], uu_hat[3:6])
    if add:
        d = work[(rhs, 3, False)]
        plan.div_uu(d, uu_hat, scale)
        plan.lincomb(rhs, [rhs, d], [1.0, 1.0])
    else:
        plan.div_uu(rhs, uu_hat, scale)
    return rhs


def getConvection(convection):
    """Return Conv(rhs, u_hat, work, Tp, VTp, K, u_dealias) (NS.py:167-201).

    The Vortex form F(u x curl u) on the dealiased space is one fused launch
    sequence (2/3 and 3/2 rule).  The Standard, Divergence and Skewed forms
    run their transforms and products as separate launches: the dealiased
    transforms of the plan (padded 3N/2 grid for the 3/2 rule, input
    truncation for the 2/3 rule) and the pointwise product kernels."""
    if convection == "Vortex":
        def Conv(rhs, u_hat, work, Tp, VTp, K, u_dealias=None):
            return VTp.plan.convection(u_hat, rhs)

        Conv.convection = convection
        Conv.fused_vortex = True
        return Conv
    if convection not in ("Standard", "Divergence", "Skewed"):
        raise NotImplementedError("convection %r" % convection)

    def Conv(rhs, u_hat, work, Tp, VTp, K, u_dealias):
        plan = VTp.plan
        u_dealias = _dealiased_backward(plan, u_hat, u_dealias)
        if convection == "Standard":
            rhs = standard_convection(rhs, u_dealias, u_hat, work, Tp, K)
            plan.lincomb(rhs, [rhs], [-1.0])
        elif convection == "Divergence":
            rhs = divergence_convection(rhs, u_dealias, work, Tp, K, False, scale=-1.0)
        else:
            rhs = standard_convection(rhs, u_dealias, u_hat, work, Tp, K)
            rhs = divergence_convection(rhs, u_dealias, work, Tp, K, True)
            plan.lincomb(rhs, [rhs], [-0.5])
        return rhs

    Conv.convection = convection
    return Conv


@optimizer
def add_pressure_diffusion(rhs, u_hat, nu, K2, K, P_hat, K_over_K2):
    """Add contributions from pressure and diffusion to the rhs (NS.py:203-217)."""


def ComputeRHS(rhs, u_hat, solver, **context):
    """rhs = conv(u_hat), Nyquist mask, pressure and diffusion, + Source (NS.py:219-261).

    With the module's Vortex convection this is one fused launch sequence.
    The Source is always added, as the reference does (it is zero unless the
    user fills it, by any means).  A replaced ``solver.conv`` runs step by
    step with the same kernels."""
    c = context
    VT = c["VT"]
    src = c["Source"]
    if getattr(solver.conv, "fused_vortex", False):
        return VT.plan.compute_rhs(u_hat, rhs, params.nu, source=src, p_hat=c["P_hat"])
    u_dealias = c.get("u_dealias")
    if u_dealias is None:  # (the context builds it on first access; keyword unpacking skips it)
        VTp = c["VTp"]
        u_dealias = getattr(VTp, "_u_dealias", None)
        if u_dealias is None:
            u_dealias = VTp._u_dealias = Array(VTp)
    rhs = solver.conv(rhs, u_hat, c["work"], c["Tp"], c["VTp"], c["K"], u_dealias)
    if c["mask"] is not None:
        VT.plan.mask_nyquist(rhs)
    rhs = solver.add_pressure_diffusion(rhs, u_hat, params.nu, None, c["K"], c["P_hat"], None)
    rhs += src
    return rhs


_ComputeRHS_fused = ComputeRHS
