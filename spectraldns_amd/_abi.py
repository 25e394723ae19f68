"""ctypes binding of libsdns_hip.so (C-ABI declared in include/sdns.h).

There is no CPU fallback: if the library is missing or fails to load, every
entry point raises.  Tensors are torch tensors on the GPU; only their device
pointers cross the ABI.  Shapes, dtypes and contiguity are checked here on the
host before any kernel is launched.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsdns_hip.so")

SDNS_SINGLE, SDNS_DOUBLE = 0, 1
DEALIAS = {"None": 0, None: 0, "2/3-rule": 1, "3/2-rule": 2}

# every exported symbol of include/sdns.h with (restype, argtypes)
_P = ctypes.c_void_p
_SIGS = {
    "sdns_plan_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.POINTER(ctypes.c_int32),
                                        ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, _P]),
    "sdns_plan_destroy": (ctypes.c_int, [_P]),
    "sdns_plan_set_stream": (ctypes.c_int, [_P, _P]),
    "sdns_plan_set_variant": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int]),
    "sdns_workspace_bytes": (ctypes.c_size_t, [_P]),
    "sdns_set_workspace": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    "sdns_forward": (ctypes.c_int, [_P, _P, _P, ctypes.c_int]),
    "sdns_backward": (ctypes.c_int, [_P, _P, _P, ctypes.c_int]),
    "sdns_cross1": (ctypes.c_int, [_P, _P, _P, _P, ctypes.c_size_t]),
    "sdns_cross2": (ctypes.c_int, [_P, _P, _P]),
    "sdns_add_pressure_diffusion": (ctypes.c_int, [_P, _P, _P, ctypes.c_double, _P]),
    "sdns_divergence": (ctypes.c_int, [_P, _P, _P]),
    "sdns_mul_minus_i": (ctypes.c_int, [_P, _P, _P, ctypes.c_size_t]),
    "sdns_cross2_k_over_k2": (ctypes.c_int, [_P, _P, _P]),
    "sdns_add_linear": (ctypes.c_int, [_P, _P, _P, ctypes.c_double, _P]),
    "sdns_vv_rhs_tail": (ctypes.c_int, [_P, _P, _P, _P, ctypes.c_double, _P]),
    "sdns_dealias_truncate": (ctypes.c_int, [_P, _P, _P]),
    "sdns_backward_dealiased": (ctypes.c_int, [_P, _P, _P, ctypes.c_int]),
    "sdns_forward_dealiased": (ctypes.c_int, [_P, _P, _P, ctypes.c_int]),
    "sdns_elsasser_products": (ctypes.c_int, [_P, _P, _P, ctypes.c_size_t]),
    "sdns_set_elsasser": (ctypes.c_int, [_P, _P, _P]),
    "sdns_gradient_truncated": (ctypes.c_int, [_P, _P, _P]),
    "sdns_dot3": (ctypes.c_int, [_P, _P, _P, _P, ctypes.c_size_t]),
    "sdns_uu_products": (ctypes.c_int, [_P, _P, _P, ctypes.c_size_t]),
    "sdns_div_uu": (ctypes.c_int, [_P, _P, _P, ctypes.c_double]),
    "sdns_mask_nyquist": (ctypes.c_int, [_P, _P]),
    "sdns_convection": (ctypes.c_int, [_P, _P, _P]),
    "sdns_compute_rhs": (ctypes.c_int, [_P, _P, _P, ctypes.c_double, _P, _P]),
    "sdns_rk4_steps": (ctypes.c_int, [_P, _P, _P, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                      ctypes.c_double, _P, _P]),
    "sdns_rk4_steps_source": (ctypes.c_int, [_P, _P, _P, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                             ctypes.c_double, _P, _P, _P]),
    "sdns_convection_uw": (ctypes.c_int, [_P, _P, _P, _P]),
    "sdns_energy": (ctypes.c_int, [_P, _P, ctypes.c_double, _P]),
    "sdns_lincomb": (ctypes.c_int, [_P, _P, ctypes.c_int, ctypes.POINTER(_P), ctypes.POINTER(ctypes.c_double)]),
    "sdns_error_norm": (ctypes.c_int, [_P, _P, _P, _P, ctypes.c_double, ctypes.c_double, _P]),
    "sdns_force": (ctypes.c_int, [_P, _P, ctypes.c_double, ctypes.c_double, _P]),
    "sdns_force_compensated": (ctypes.c_int, [_P, _P, _P, ctypes.c_double, ctypes.c_double, _P]),
    "sdns_spectrum": (ctypes.c_int, [_P, _P, ctypes.c_int, _P]),
    "sdns_spectrum_noalloc": (ctypes.c_int, [_P, _P, ctypes.c_int, _P]),
    "sdns_timing_enable": (ctypes.c_int, [_P, ctypes.c_int]),
    "sdns_timing_read": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]),
    "sdns_timing_passes": (ctypes.c_int, [_P, ctypes.c_int]),
    "sdns_batch_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int, ctypes.POINTER(ctypes.c_int32),
                                         ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, _P]),
    "sdns_batch_destroy": (ctypes.c_int, [_P]),
    "sdns_batch_set_state": (ctypes.c_int, [_P, ctypes.POINTER(_P)]),
    "sdns_batch_set_params": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "sdns_batch_use_graph": (ctypes.c_int, [_P, ctypes.c_int]),
    "sdns_batch_set_stream": (ctypes.c_int, [_P, _P]),
    "sdns_batch_rk4_steps": (ctypes.c_int, [_P, ctypes.c_int, _P]),
    "sdns_batch_box_plan": (ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(_P)]),
    "sdns_batch_set_variant": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "sdns_closure_nparams": (ctypes.c_size_t, []),
    "sdns_closure_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.POINTER(ctypes.c_int32), _P]),
    "sdns_closure_destroy": (ctypes.c_int, [_P]),
    "sdns_closure_set_stream": (ctypes.c_int, [_P, _P]),
    "sdns_closure_set_params": (ctypes.c_int, [_P, _P]),
    "sdns_closure_forward": (ctypes.c_int, [_P, _P, ctypes.c_float, ctypes.c_float, _P]),
    "sdns_closure_train_step": (ctypes.c_int, [_P, _P, _P, ctypes.c_float, ctypes.c_float, _P, _P]),
    "sdns_adam_step": (ctypes.c_int, [_P, _P, _P, _P, ctypes.c_size_t, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_float, ctypes.c_float, ctypes.c_int, _P]),
    "sdns_box_filter": (ctypes.c_int, [_P, _P, ctypes.c_int, _P, _P]),
    "sdns_closure_rk4_steps": (ctypes.c_int, [_P, _P, _P, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                              ctypes.c_double, ctypes.c_float, ctypes.c_float, ctypes.c_int, _P]),
    "sdns_last_error": (ctypes.c_char_p, []),
}

_lib = None


class SdnsError(RuntimeError):
    """A libsdns_hip call returned an error code."""


def load():
    """Load libsdns_hip.so once; raise if it is missing (no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SdnsError("libsdns_hip.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        raise SdnsError("libsdns_hip error %d: %s" % (rc, _lib.sdns_last_error().decode()))


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class Plan:
    """One box: sizes, dealiasing, precision, twiddles, workspace (NS.py:12-48)."""

    def __init__(self, N, L, precision="double", dealias="2/3-rule", mask_nyquist=True, dealias_mode=0,
                 tuned=True):
        import torch
        lib = load()
        if not torch.cuda.is_available():
            raise SdnsError("no GPU visible: the HIP path has no CPU fallback")
        self.N = tuple(int(n) for n in N)
        self.L = tuple(float(x) for x in L)
        self.precision = precision
        self.dealias = dealias
        self.dealias_mode = int(dealias_mode)
        self.nyquist_masked = bool(mask_nyquist)
        self.real_dtype = torch.float32 if precision == "single" else torch.float64
        self.complex_dtype = torch.complex64 if precision == "single" else torch.complex128
        self.device = torch.device("cuda", torch.cuda.current_device())
        h = _P()
        Na = (ctypes.c_int32 * 3)(*self.N)
        La = (ctypes.c_double * 3)(*self.L)
        stream = torch.cuda.current_stream().cuda_stream
        _check(lib.sdns_plan_create(ctypes.byref(h), Na, La,
                                    SDNS_SINGLE if precision == "single" else SDNS_DOUBLE,
                                    DEALIAS[dealias], int(self.nyquist_masked), self.dealias_mode,
                                    _P(stream)))
        self._h = h
        self._lib = lib
        nbytes = lib.sdns_workspace_bytes(h)
        # workspace from torch's caching allocator (256-byte aligned)
        self._ws = torch.empty(nbytes + 256, dtype=torch.uint8, device=self.device)
        base = self._ws.data_ptr()
        off = (-base) % 256
        _check(lib.sdns_set_workspace(h, _P(base + off), nbytes))
        self.workspace_bytes = nbytes
        self.spectral_shape = (self.N[0], self.N[1], self.N[2] // 2 + 1)
        self.physical_shape = self.N
        # physical grid of the dealiased space: padded (3N/2) for the 3/2 rule
        self.dealiased_shape = tuple(3 * n // 2 for n in self.N) if dealias == "3/2-rule" else self.N
        from .autotune import apply_schedule
        self.schedule = apply_schedule(self) if tuned else None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.sdns_plan_destroy(h)
            self._h = None

    # ---------------------------------------------------------------- checks
    def _sync_stream(self):
        import torch
        _check(self._lib.sdns_plan_set_stream(self._h, _P(torch.cuda.current_stream().cuda_stream)))

    def _spec(self, t, ncomp=None, name="array"):
        if t.device.type != "cuda" or t.dtype != self.complex_dtype or not t.is_contiguous():
            raise SdnsError("%s must be a contiguous %s tensor on the GPU" % (name, self.complex_dtype))
        if tuple(t.shape[-3:]) != self.spectral_shape:
            raise SdnsError("%s has shape %s, expected (..., %d, %d, %d)" % ((name, tuple(t.shape)) + self.spectral_shape))
        nc = 1 if t.dim() == 3 else t.shape[0]
        if t.dim() not in (3, 4) or (ncomp is not None and nc != ncomp):
            raise SdnsError("%s must have %s components" % (name, ncomp))
        return nc

    def _phys(self, t, name="array", shape=None):
        shape = self.physical_shape if shape is None else shape
        if t.device.type != "cuda" or t.dtype != self.real_dtype or not t.is_contiguous():
            raise SdnsError("%s must be a contiguous %s tensor on the GPU" % (name, self.real_dtype))
        if tuple(t.shape[-3:]) != shape or t.dim() not in (3, 4):
            raise SdnsError("%s has shape %s, expected (..., %d, %d, %d)" % ((name, tuple(t.shape)) + shape))
        return 1 if t.dim() == 3 else t.shape[0]

    # ---------------------------------------------------------------- entry points
    def forward(self, u, u_hat):
        nc = self._phys(u, "u")
        if self._spec(u_hat, nc, "u_hat") != nc:
            raise SdnsError("component mismatch")
        self._sync_stream()
        _check(self._lib.sdns_forward(self._h, _ptr(u), _ptr(u_hat), nc))
        return u_hat

    def backward(self, u_hat, u):
        nc = self._spec(u_hat, None, "u_hat")
        if self._phys(u, "u") != nc:
            raise SdnsError("component mismatch")
        self._sync_stream()
        _check(self._lib.sdns_backward(self._h, _ptr(u_hat), _ptr(u), nc))
        return u

    def backward_dealiased(self, u_hat, u):
        """VTp.backward: the inverse transform of the kept input modes only (2/3
        rule or no dealiasing, the truncation done inside the passes), or the
        padded inverse transform onto the 3N/2 grid ``dealiased_shape`` (3/2 rule)."""
        nc = self._spec(u_hat, None, "u_hat")
        if self._phys(u, "u", self.dealiased_shape) != nc:
            raise SdnsError("component mismatch")
        self._sync_stream()
        _check(self._lib.sdns_backward_dealiased(self._h, _ptr(u_hat), _ptr(u), nc))
        return u

    def forward_dealiased(self, u, u_hat):
        """Tp.forward: sdns_forward (2/3 rule, no dealiasing), or the forward
        transform of the padded grid truncated to the plan's modes (3/2 rule)."""
        nc = self._phys(u, "u", self.dealiased_shape)
        if self._spec(u_hat, nc, "u_hat") != nc:
            raise SdnsError("component mismatch")
        self._sync_stream()
        _check(self._lib.sdns_forward_dealiased(self._h, _ptr(u), _ptr(u_hat), nc))
        return u_hat

    def cross1(self, c, a, b):
        for t, nm in ((c, "c"), (a, "a"), (b, "b")):
            if t.device.type != "cuda" or t.dtype != self.real_dtype or not t.is_contiguous() or t.shape[0] != 3:
                raise SdnsError("cross1: %s must be a contiguous (3, ...) %s GPU tensor" % (nm, self.real_dtype))
        if not (c.shape == a.shape == b.shape):
            raise SdnsError("cross1: shape mismatch")
        self._sync_stream()
        _check(self._lib.sdns_cross1(self._h, _ptr(c), _ptr(a), _ptr(b), a[0].numel()))
        return c

    def cross2(self, c_hat, a_hat):
        self._spec(c_hat, 3, "c")
        self._spec(a_hat, 3, "a")
        self._sync_stream()
        _check(self._lib.sdns_cross2(self._h, _ptr(c_hat), _ptr(a_hat)))
        return c_hat

    def cross2_k_over_k2(self, u_hat, w_hat):
        """u_hat = 1j*((K/K^2) x w_hat), K^2 = 0 taken as 1 (solvers/VV.py:57-73)."""
        self._spec(u_hat, 3, "u_hat")
        self._spec(w_hat, 3, "w_hat")
        self._sync_stream()
        _check(self._lib.sdns_cross2_k_over_k2(self._h, _ptr(u_hat), _ptr(w_hat)))
        return u_hat

    def dealias_truncate(self, out, a_hat):
        """out = a_hat with the modes outside the kept input range zeroed (3 components)."""
        self._spec(out, 3, "out")
        self._spec(a_hat, 3, "a_hat")
        self._sync_stream()
        _check(self._lib.sdns_dealias_truncate(self._h, _ptr(out), _ptr(a_hat)))
        return out

    def elsasser_products(self, zz, ub):
        """zz[3i+j] = (u_i + b_i)(u_j - b_j) on physical (6, ...) -> (9, ...) arrays."""
        for t, nm, nc in ((zz, "zz", 9), (ub, "ub", 6)):
            if t.device.type != "cuda" or t.dtype != self.real_dtype or not t.is_contiguous() \
                    or tuple(t.shape) not in ((nc,) + self.physical_shape, (nc,) + self.dealiased_shape) \
                    or tuple(t.shape[1:]) != tuple(ub.shape[1:]):
                raise SdnsError("%s must be a contiguous (%d,) + %s %s GPU tensor"
                                % (nm, nc, self.physical_shape, self.real_dtype))
        self._sync_stream()
        _check(self._lib.sdns_elsasser_products(self._h, _ptr(zz), _ptr(ub), ub[0].numel()))
        return zz

    def _physn(self, t, nc, name):
        # pointwise kernels: the plan's grid or the padded grid of the 3/2 rule
        lead = (nc,) if nc > 1 else ()
        if t.device.type != "cuda" or t.dtype != self.real_dtype or not t.is_contiguous() \
                or tuple(t.shape) not in (lead + self.physical_shape, lead + self.dealiased_shape):
            raise SdnsError("%s must be a contiguous %s %s GPU tensor of %d component(s)"
                            % (name, self.physical_shape, self.real_dtype, nc))

    def _specn(self, t, nc, name):
        if t.device.type != "cuda" or t.dtype != self.complex_dtype or not t.is_contiguous() \
                or tuple(t.shape) != ((nc,) if nc > 1 else ()) + self.spectral_shape:
            raise SdnsError("%s must be a contiguous %s %s GPU tensor of %d component(s)"
                            % (name, self.spectral_shape, self.complex_dtype, nc))

    def gradient_truncated(self, out, a_hat):
        """out[j] = 1j*K_j*a_hat on the kept modes, 0 elsewhere (one component in, three out)."""
        self._specn(out, 3, "out")
        self._specn(a_hat, 1, "a_hat")
        self._sync_stream()
        _check(self._lib.sdns_gradient_truncated(self._h, _ptr(out), _ptr(a_hat)))
        return out

    def dot3(self, out, u, g):
        """out = u0*g0 + u1*g1 + u2*g2 on physical arrays."""
        self._physn(out, 1, "out")
        self._physn(u, 3, "u")
        self._physn(g, 3, "g")
        if not (tuple(out.shape[-3:]) == tuple(u.shape[-3:]) == tuple(g.shape[-3:])):
            raise SdnsError("dot3: grid mismatch")
        self._sync_stream()
        _check(self._lib.sdns_dot3(self._h, _ptr(out), _ptr(u), _ptr(g), out.numel()))
        return out

    def uu_products(self, uu, u):
        """uu = (u0u0, u0u1, u0u2, u1u1, u1u2, u2u2) on physical arrays."""
        self._physn(uu, 6, "uu")
        self._physn(u, 3, "u")
        if tuple(uu.shape[-3:]) != tuple(u.shape[-3:]):
            raise SdnsError("uu_products: grid mismatch")
        self._sync_stream()
        _check(self._lib.sdns_uu_products(self._h, _ptr(uu), _ptr(u), u[0].numel()))
        return uu

    def div_uu(self, out, uu_hat, scale=1.0):
        """out = scale * 1j*sum_j K_j F(u_i u_j) from the six spectral products."""
        self._specn(out, 3, "out")
        self._specn(uu_hat, 6, "uu_hat")
        self._sync_stream()
        _check(self._lib.sdns_div_uu(self._h, _ptr(out), _ptr(uu_hat), float(scale)))
        return out

    def set_elsasser(self, c_hat, zz_hat):
        """c_hat (6 components) from the 9 spectral products zz_hat (MHD.py:104-113)."""
        if tuple(c_hat.shape) != (6,) + self.spectral_shape or tuple(zz_hat.shape) != (9,) + self.spectral_shape:
            raise SdnsError("set_elsasser: c_hat (6, ...) and zz_hat (9, ...) spectral arrays required")
        for t in (c_hat, zz_hat):
            if t.device.type != "cuda" or t.dtype != self.complex_dtype or not t.is_contiguous():
                raise SdnsError("set_elsasser: contiguous %s GPU tensors required" % self.complex_dtype)
        self._sync_stream()
        _check(self._lib.sdns_set_elsasser(self._h, _ptr(c_hat), _ptr(zz_hat)))
        return c_hat

    def vv_rhs_tail(self, rhs, v_hat, w_hat, nu, source=None):
        """rhs = mask(1j*(K x v_hat)) - nu*K2*w_hat (+ source): the VV RHS after
        the convection passes, one launch."""
        for t, nm in ((rhs, "rhs"), (v_hat, "v_hat"), (w_hat, "w_hat")):
            self._spec(t, 3, nm)
        if source is not None:
            self._spec(source, 3, "Source")
        self._sync_stream()
        _check(self._lib.sdns_vv_rhs_tail(self._h, _ptr(rhs), _ptr(v_hat), _ptr(w_hat), float(nu),
                                          _ptr(source) if source is not None else None))
        return rhs

    def add_linear(self, rhs, w_hat, nu, source=None):
        """rhs -= nu*K2*w_hat; rhs += source (solvers/VV.py:110-115), in place."""
        self._spec(rhs, 3, "rhs")
        self._spec(w_hat, 3, "w_hat")
        if source is not None:
            self._spec(source, 3, "Source")
        self._sync_stream()
        _check(self._lib.sdns_add_linear(self._h, _ptr(rhs), _ptr(w_hat), float(nu),
                                         _ptr(source) if source is not None else None))
        return rhs

    def add_pressure_diffusion(self, rhs, u_hat, nu, p_hat=None):
        self._spec(rhs, 3, "rhs")
        self._spec(u_hat, 3, "u_hat")
        if p_hat is not None:
            self._spec(p_hat, 1, "P_hat")
        self._sync_stream()
        _check(self._lib.sdns_add_pressure_diffusion(self._h, _ptr(rhs), _ptr(u_hat), float(nu),
                                                      _ptr(p_hat) if p_hat is not None else None))
        return rhs

    def divergence(self, div_hat, u_hat):
        self._spec(div_hat, 1, "div_hat")
        self._spec(u_hat, 3, "u_hat")
        self._sync_stream()
        _check(self._lib.sdns_divergence(self._h, _ptr(div_hat), _ptr(u_hat)))
        return div_hat

    def mul_minus_i(self, out, a):
        if out.shape != a.shape or out.dtype != self.complex_dtype or a.dtype != self.complex_dtype \
                or not (out.is_contiguous() and a.is_contiguous()):
            raise SdnsError("mul_minus_i: matching contiguous complex tensors required")
        self._sync_stream()
        _check(self._lib.sdns_mul_minus_i(self._h, _ptr(out), _ptr(a), a.numel()))
        return out

    def mask_nyquist(self, a_hat):
        self._spec(a_hat, 3, "a_hat")
        self._sync_stream()
        _check(self._lib.sdns_mask_nyquist(self._h, _ptr(a_hat)))
        return a_hat

    def convection(self, u_hat, rhs):
        self._spec(u_hat, 3, "u_hat")
        self._spec(rhs, 3, "rhs")
        self._sync_stream()
        _check(self._lib.sdns_convection(self._h, _ptr(u_hat), _ptr(rhs)))
        return rhs

    def compute_rhs(self, u_hat, rhs, nu, source=None, p_hat=None):
        self._spec(u_hat, 3, "u_hat")
        self._spec(rhs, 3, "rhs")
        if source is not None:
            self._spec(source, 3, "Source")
        if p_hat is not None:
            self._spec(p_hat, 1, "P_hat")
        self._sync_stream()
        _check(self._lib.sdns_compute_rhs(self._h, _ptr(u_hat), _ptr(rhs), float(nu),
                                          _ptr(source) if source is not None else None,
                                          _ptr(p_hat) if p_hat is not None else None))
        return rhs

    def convection_uw(self, u_hat, w_hat, out):
        """out = F(u_dealias x w_dealias) with the vorticity w_hat given."""
        self._spec(u_hat, 3, "u_hat")
        self._spec(w_hat, 3, "w_hat")
        self._spec(out, 3, "out")
        self._sync_stream()
        _check(self._lib.sdns_convection_uw(self._h, _ptr(u_hat), _ptr(w_hat), _ptr(out)))
        return out

    def rk4_steps(self, u_hat, dts, nu, diag=None, u_lo=None, p_hat=None, source=None):
        import torch
        self._spec(u_hat, 3, "u_hat")
        n = len(dts)
        if diag is not None:
            if diag.dtype != torch.float64 or diag.device.type != "cuda" or diag.numel() < 2 * n \
                    or not diag.is_contiguous():
                raise SdnsError("diag must be a contiguous float64 GPU tensor of >= 2*nsteps entries")
        if u_lo is not None:
            self._spec(u_lo, 3, "u_lo")
        if p_hat is not None:
            self._spec(p_hat, 1, "P_hat")
        if source is not None:
            self._spec(source, 3, "Source")
        arr = (ctypes.c_double * max(n, 1))(*[float(d) for d in dts])
        self._sync_stream()
        _check(self._lib.sdns_rk4_steps_source(self._h, _ptr(u_hat), _ptr(u_lo) if u_lo is not None else None,
                                               arr, n, float(nu), _ptr(source) if source is not None else None,
                                               _ptr(diag) if diag is not None else None,
                                               _ptr(p_hat) if p_hat is not None else None))
        return u_hat

    def lincomb(self, out, xs, cs):
        """out = cs[0]*xs[0] + cs[1]*xs[1] + ... over state-shaped arrays (in place
        when out is one of xs); coefficients rounded to the working precision.
        Arrays of 3k components run as k launches of 3."""
        if out.dim() == 4 and out.shape[0] > 3 and out.shape[0] % 3 == 0:
            for c in range(0, out.shape[0], 3):
                self.lincomb(out[c:c + 3], [x[c:c + 3] for x in xs], cs)
            return out
        self._spec(out, 3, "out")
        for x in xs:
            self._spec(x, 3, "term")
        n = len(xs)
        if n != len(cs) or not 1 <= n <= 10:
            raise SdnsError("lincomb takes 1..10 terms with one coefficient each")
        ptrs = (_P * n)(*[x.data_ptr() for x in xs])
        coef = (ctypes.c_double * n)(*[float(c) for c in cs])
        self._sync_stream()
        _check(self._lib.sdns_lincomb(self._h, _ptr(out), n, ptrs, coef))
        return out

    def error_norm(self, err, u0, u1, atol, rtol, out=None):
        """Per component sum of |err/sc|^2, sc = atol + max(|u0|, |u1|)*rtol, as a
        float64 GPU tensor of 3 entries (3k entries for arrays of 3k components,
        e.g. the six of the MHD state, as k launches of 3)."""
        import torch
        if err.dim() == 4 and err.shape[0] > 3 and err.shape[0] % 3 == 0 and out is None:
            return torch.cat([self.error_norm(err[c:c + 3], u0[c:c + 3], u1[c:c + 3], atol, rtol)
                              for c in range(0, err.shape[0], 3)])
        for t, nm in ((err, "err"), (u0, "u0"), (u1, "u1")):
            self._spec(t, 3, nm)
        if out is None:
            out = torch.empty(3, dtype=torch.float64, device=err.device)
        elif out.dtype != torch.float64 or out.numel() < 3 or out.device.type != "cuda":
            raise SdnsError("out must be a float64 GPU tensor of 3 entries")
        self._sync_stream()
        _check(self._lib.sdns_error_norm(self._h, _ptr(err), _ptr(u0), _ptr(u1), float(atol), float(rtol),
                                         _ptr(out)))
        return out

    def energy(self, u_hat, nu, out):
        import torch
        self._spec(u_hat, 3, "u_hat")
        if out.dtype != torch.float64 or out.numel() < 2 or out.device.type != "cuda":
            raise SdnsError("out must be a float64 GPU tensor of 2 entries")
        self._sync_stream()
        _check(self._lib.sdns_energy(self._h, _ptr(u_hat), float(nu), _ptr(out)))
        return out

    def force(self, u_hat, target, kf, sums=None, u_lo=None):
        """Isotropic.py:154-184 forcing in place (energy_fourier held at target).
        With u_lo (single precision) the state is the compensated pair
        u_hat + u_lo: energies of the pair, the pair scaled as one value."""
        import torch
        self._spec(u_hat, 3, "u_hat")
        if sums is not None and (sums.dtype != torch.float64 or sums.numel() < 2 or sums.device.type != "cuda"):
            raise SdnsError("sums must be a float64 GPU tensor of 2 entries")
        sp = _ptr(sums) if sums is not None else None
        self._sync_stream()
        if u_lo is not None:
            self._spec(u_lo, 3, "u_lo")
            _check(self._lib.sdns_force_compensated(self._h, _ptr(u_hat), _ptr(u_lo), float(target), float(kf), sp))
        else:
            _check(self._lib.sdns_force(self._h, _ptr(u_hat), float(target), float(kf), sp))
        return u_hat

    def spectrum_sums(self, u_hat, nbins, out=None, noalloc=False):
        """(uiui sums, mode counts) per shell z = 1..nbins, as float64 GPU tensors
        (views of ``out``, 2*nbins doubles, when given).  ``noalloc`` calls
        sdns_spectrum_noalloc, whose partial sums live in the library's static
        scratch instead of the plan's workspace."""
        import torch
        self._spec(u_hat, 3, "u_hat")
        if out is None:
            out = torch.empty(2 * nbins, dtype=torch.float64, device=self.device)
        elif out.dtype != torch.float64 or out.numel() < 2 * nbins or out.device.type != "cuda":
            raise SdnsError("out must be a float64 GPU tensor of 2*nbins entries")
        self._sync_stream()
        fn = self._lib.sdns_spectrum_noalloc if noalloc else self._lib.sdns_spectrum
        _check(fn(self._h, _ptr(u_hat), int(nbins), _ptr(out)))
        return out[:nbins], out[nbins:2 * nbins]

    def box_filter(self, u, r):
        """(ubar (3, N/r...), tau (6, N/r...)) of the physical velocity u (block means, FP64 sums)."""
        import torch
        self._phys(u, "u")
        if u.shape[0] != 3 or any(n % r for n in self.N):
            raise SdnsError("box_filter: u must have 3 components and r must divide N")
        shp = tuple(n // r for n in self.N)
        ub = torch.empty((3,) + shp, dtype=self.real_dtype, device=self.device)
        tau = torch.empty((6,) + shp, dtype=self.real_dtype, device=self.device)
        self._sync_stream()
        _check(self._lib.sdns_box_filter(self._h, _ptr(u), int(r), _ptr(ub), _ptr(tau)))
        return ub, tau

    def set_variant(self, pass_id, variant):
        """Schedule variant (0 default, 1 alternative) of pass 0..4, 5 = plain transforms;
        the z pass (2) has variants 0..5 and the x-forward + RK pass (4) 0..3."""
        _check(self._lib.sdns_plan_set_variant(self._h, int(pass_id), int(variant)))

    # ---------------------------------------------------------------- timing
    PASS_NAMES = ("x_inverse", "y_inverse_curl", "z_cross", "y_forward", "x_forward_rk", "other")

    def timing_enable(self, on=True):
        _check(self._lib.sdns_timing_enable(self._h, int(bool(on))))

    def timing_passes(self, names=None):
        """Record only these passes (names from PASS_NAMES; None: all)."""
        mask = 63 if names is None else sum(1 << self.PASS_NAMES.index(n) for n in names)
        _check(self._lib.sdns_timing_passes(self._h, mask))

    def timing_read(self):
        """{pass name: (milliseconds, launches)} since the last read."""
        ms = (ctypes.c_double * 6)()
        cnt = (ctypes.c_int * 6)()
        _check(self._lib.sdns_timing_read(self._h, ms, cnt))
        return {n: (ms[i], cnt[i]) for i, n in enumerate(self.PASS_NAMES)}


class Batch:
    """Ensemble of independent boxes (own N, L, nu, dt) advanced together:
    every pass of the RK4 step is one launch over all boxes and the step is
    replayed from a captured hipGraph (sdns_batch_* in include/sdns.h)."""

    def __init__(self, boxes, precision="single", dealias="2/3-rule", mask_nyquist=True, dealias_mode=0):
        import torch
        lib = load()
        if not torch.cuda.is_available():
            raise SdnsError("no GPU visible: the HIP path has no CPU fallback")
        self.boxes = [(tuple(int(n) for n in N), tuple(float(x) for x in L)) for N, L in boxes]
        if not self.boxes:
            raise SdnsError("an ensemble needs at least one box")
        self.nbox = len(self.boxes)
        self.precision = precision
        self.dealias = dealias
        self.complex_dtype = torch.complex64 if precision == "single" else torch.complex128
        self.device = torch.device("cuda", torch.cuda.current_device())
        Na = (ctypes.c_int32 * (3 * self.nbox))(*[n for N, _ in self.boxes for n in N])
        La = (ctypes.c_double * (3 * self.nbox))(*[x for _, L in self.boxes for x in L])
        h = _P()
        _check(lib.sdns_batch_create(ctypes.byref(h), self.nbox, Na, La,
                                     SDNS_SINGLE if precision == "single" else SDNS_DOUBLE,
                                     DEALIAS[dealias], int(bool(mask_nyquist)), int(dealias_mode),
                                     _P(torch.cuda.current_stream().cuda_stream)))
        self._h = h
        self._lib = lib
        self._state = None
        self.spectral_shapes = [(3, N[0], N[1], N[2] // 2 + 1) for N, _ in self.boxes]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.sdns_batch_destroy(h)
            self._h = None

    def set_state(self, u_hats):
        """Box i advances u_hats[i] in place (contiguous (3, N0, N1, N2/2+1) GPU tensors)."""
        u_hats = list(u_hats)
        if len(u_hats) != self.nbox:
            raise SdnsError("%d states for %d boxes" % (len(u_hats), self.nbox))
        for i, (t, shp) in enumerate(zip(u_hats, self.spectral_shapes)):
            if t.device.type != "cuda" or t.dtype != self.complex_dtype or not t.is_contiguous() \
                    or tuple(t.shape) != shp:
                raise SdnsError("box %d: state must be a contiguous %s GPU tensor of shape %s"
                                % (i, self.complex_dtype, shp))
        arr = (_P * self.nbox)(*[t.data_ptr() for t in u_hats])
        _check(self._lib.sdns_batch_set_state(self._h, arr))
        self._state = u_hats  # keep the tensors alive while the batch holds their pointers

    def set_params(self, nu, dt):
        nu = [float(x) for x in nu]
        dt = [float(x) for x in dt]
        if len(nu) != self.nbox or len(dt) != self.nbox:
            raise SdnsError("nu and dt need one value per box")
        _check(self._lib.sdns_batch_set_params(self._h, (ctypes.c_double * self.nbox)(*nu),
                                               (ctypes.c_double * self.nbox)(*dt)))

    def use_graph(self, on=True):
        _check(self._lib.sdns_batch_use_graph(self._h, int(bool(on))))

    def line_lengths(self):
        """{pass id: sorted line lengths of that batched pass over the boxes}:
        x and y lines of the (padded) grid, half z lines for the z pass."""
        pad = (lambda n: 3 * n // 2) if self.dealias == "3/2-rule" else (lambda n: n)
        out = {}
        for N, _ in self.boxes:
            M = [pad(n) for n in N]
            for k, l in enumerate((M[0], M[1], M[2] // 2, M[1], M[0])):
                out.setdefault(k, set()).add(l)
        return {k: sorted(v) for k, v in out.items()}

    def set_variant(self, pass_id, length, variant):
        """Schedule variant (0 or 1) of one batched pass for one line length."""
        _check(self._lib.sdns_batch_set_variant(self._h, int(pass_id), int(length), int(variant)))

    def rk4_steps(self, nsteps, diag=None):
        """nsteps RK4 steps of every box; diag (float64 GPU tensor of 2*nbox
        entries) receives (energy, dissipation) per box after the last step."""
        import torch
        if self._state is None:
            raise SdnsError("set_state first")
        if diag is not None and (diag.dtype != torch.float64 or diag.device.type != "cuda"
                                 or diag.numel() < 2 * self.nbox or not diag.is_contiguous()):
            raise SdnsError("diag must be a contiguous float64 GPU tensor of >= 2*nbox entries")
        _check(self._lib.sdns_batch_set_stream(self._h, _P(torch.cuda.current_stream().cuda_stream)))
        _check(self._lib.sdns_batch_rk4_steps(self._h, int(nsteps), _ptr(diag) if diag is not None else None))
