"""Magnetohydrodynamics in Elsasser form, triply periodic box
(spectralDNS/solvers/MHD.py).

The primary variable is UB_hat = (U_hat, B_hat), six spectral components.
The module keeps the reference's protocol (``get_context`` with its keys,
``set_Elsasser``, ``divergenceConvection``, ``getConvection('Divergence')``,
``add_pressure_diffusion`` with nu on U and eta on B, ``ComputeRHS``,
``get_divergence``) and RK4 runs stage by stage over ComputeRHS.  Each step
of the RHS is a libsdns_hip launch: the inverse transforms of the six
fields over their kept modes (the 2/3-rule truncation inside the transform
passes), one kernel for the nine Elsasser
products (u_i + b_i)(u_j - b_j), their forward transforms, one kernel for
set_Elsasser, the Nyquist mask, the NS pressure/diffusion kernel on U and
the diffusion kernel on B.

Built for the 2/3 rule (shenfun's dealias_direct, dealias_mode 0), the 3/2
rule (the six fields transformed onto the padded 3N/2 grid, the nine
products transformed back and truncated, MHD.py:30-32) and for no
dealiasing.
"""
from .spectralinit import *  # noqa: F401,F403
from .spectralinit import config, params, datatypes, HDF5File  # noqa: F401
from .NS import end_of_tstep  # noqa: F401
from .VV import _WavenumberField
from .._abi import Plan
from ..fourier import FourierSpace, Array, Function


class CompositeSpace:
    """ncomp scalar fields of one plan (shenfun CompositeSpace([T]*ncomp)).
    Transforms run three components per launch sequence; the dealiased
    space's backward transform reads only the kept modes."""

    def __init__(self, plan, ncomp, dealiased=False):
        self.plan, self.ncomp, self.dealiased = plan, ncomp, dealiased

    def shape(self, forward_output=False):
        if forward_output:
            return (self.ncomp,) + tuple(self.plan.spectral_shape)
        return (self.ncomp,) + tuple(self.plan.dealiased_shape if self.dealiased else self.plan.N)

    def forward(self, u, u_hat):
        for c in range(0, u.shape[0], 3):
            if self.dealiased:
                self.plan.forward_dealiased(u[c:c + 3], u_hat[c:c + 3])
            else:
                self.plan.forward(u[c:c + 3], u_hat[c:c + 3])
        return u_hat

    def backward(self, u_hat, u):
        for c in range(0, u_hat.shape[0], 3):
            if self.dealiased:
                self.plan.backward_dealiased(u_hat[c:c + 3], u[c:c + 3])
            else:
                self.plan.backward(u_hat[c:c + 3], u[c:c + 3])
        return u


def get_context():
    """Context of the MHD solver (MHD.py:12-79)."""
    import torch
    float, complex, mpitype = datatypes(params.precision)
    if int(params.get("dealias_mode", 0)) != 0:
        raise NotImplementedError("MHD runs shenfun's 2/3 rule (dealias_mode 0)")
    plan = Plan(params.N, params.L, params.precision, params.dealias, params.mask_nyquist, 0)
    T = FourierSpace(plan, 0)
    VT = FourierSpace(plan, 1)
    VM = CompositeSpace(plan, 6)
    Tp = FourierSpace(plan, 0, dealiased=True)
    VTp = FourierSpace(plan, 1, dealiased=True)
    VMp = CompositeSpace(plan, 6, dealiased=True)
    mask = T.get_mask_nyquist() if params.mask_nyquist else None
    X = T.local_mesh(True)
    K = T.local_wavenumbers(scaled=True)
    dev = plan.device
    UB = torch.zeros(VM.shape(False), dtype=plan.real_dtype, device=dev)
    P = Array(T)
    curl = Array(VT)
    UB_hat = torch.zeros(VM.shape(True), dtype=plan.complex_dtype, device=dev)
    P_hat = Function(T)
    dU = torch.zeros_like(UB_hat)
    Source = torch.zeros_like(UB)
    ub_dealias = torch.zeros(VMp.shape(False), dtype=plan.real_dtype, device=dev)
    ZZ_hat = torch.zeros((3, 3) + plan.spectral_shape, dtype=plan.complex_dtype, device=dev)
    U = UB[:3]
    U_hat = UB_hat[:3]
    B = UB[3:]
    B_hat = UB_hat[3:]
    u = UB_hat
    hdf5file = MHDFile(config.params.solver,
                       checkpoint={"space": VM, "data": {"0": {"UB": [UB_hat]}}},
                       results={"space": VM, "data": {"UB": [UB]}})
    c = config.AttributeDict(locals())
    c["K2"] = _WavenumberField(c, "K2")
    c["K_over_K2"] = _WavenumberField(c, "K_over_K2")
    return c


class MHDFile(HDF5File):
    def update_components(self, UB, UB_hat, VM, **kw):
        """Transform to real data when storing the solution (MHD.py:81-84)."""
        VM.backward(UB_hat, UB)


def get_divergence(T, K, U_hat, **context):
    """div u = T.backward(1j*(K . U_hat)) (MHD.py:86-89)."""
    import torch
    div_hat = torch.empty(T.shape(True), dtype=T.plan.complex_dtype, device=T.plan.device)
    T.plan.divergence(div_hat, U_hat)
    return T.backward(div_hat)


def set_Elsasser(c, ZZ, K):
    """c[:3] = -1j sum_j K_j (ZZ[:, j] + ZZ[j, :])/2, c[3:] = 1j sum_j K_j (ZZ[j, :] - ZZ[:, j])/2
    (MHD.py:91-100); K carries the plan."""
    return K.plan.set_elsasser(c, ZZ.reshape((9,) + tuple(ZZ.shape[2:])))


def divergenceConvection(c, ub_dealias, Tp, K, ZZ_hat):
    """ZZ_ij = Tp.forward((u_i + b_i)(u_j - b_j)), then set_Elsasser (MHD.py:102-114).
    The reference passes z0 = u + b and z1 = u - b; here one kernel forms the
    nine products from the dealiased (u, b)."""
    import torch
    plan = K.plan
    zz = getattr(Tp, "_zz", None)
    if zz is None:
        zz = Tp._zz = torch.empty((9,) + tuple(plan.dealiased_shape), dtype=plan.real_dtype, device=plan.device)
    plan.elsasser_products(zz, ub_dealias)
    flat = ZZ_hat.view((9,) + tuple(ZZ_hat.shape[2:]))
    for k in range(0, 9, 3):
        plan.forward_dealiased(zz[k:k + 3], flat[k:k + 3])
    return set_Elsasser(c, ZZ_hat, K)


def getConvection(convection):
    """Return Conv(rhs, ub_hat, Tp, VMp, K, ub_dealias, ZZ_hat) (MHD.py:116-133)."""
    if convection in ("Standard", "Vortex", "Skewed"):
        raise NotImplementedError
    if convection != "Divergence":
        raise NotImplementedError("convection %r" % convection)

    def Conv(rhs, ub_hat, Tp, VMp, K, ub_dealias, ZZ_hat):
        ub_dealias = VMp.backward(ub_hat, ub_dealias)
        rhs = divergenceConvection(rhs, ub_dealias, Tp, K, ZZ_hat)
        return rhs

    Conv.convection = convection
    return Conv


def add_pressure_diffusion(rhs, ub_hat, nu, eta, K2, K, P_hat, K_over_K2):
    """P_hat = sum(rhs[:3] K/K^2); rhs[:3] -= P_hat K; rhs[:3] -= nu K2 u_hat;
    rhs[3:] -= eta K2 b_hat (MHD.py:135-148)."""
    plan = K.plan
    plan.add_pressure_diffusion(rhs[:3], ub_hat[:3], nu, P_hat)
    plan.add_linear(rhs[3:], ub_hat[3:], eta, None)
    return rhs


def ComputeRHS(rhs, ub_hat, solver, **context):
    """conv, Nyquist mask, pressure and diffusion (MHD.py:150-176)."""
    c = context
    rhs = solver.conv(rhs, ub_hat, c["Tp"], c["VMp"], c["K"], c["ub_dealias"], c["ZZ_hat"])
    if c["mask"] is not None:
        c["T"].plan.mask_nyquist(rhs[:3])
        c["T"].plan.mask_nyquist(rhs[3:])
    rhs = solver.add_pressure_diffusion(rhs, ub_hat, params.nu, params.eta, c["K2"], c["K"], c["P_hat"],
                                        c["K_over_K2"])
    return rhs
