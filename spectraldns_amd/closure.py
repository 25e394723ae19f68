"""Learned subgrid-stress closure for coarse-grid runs (sdns_closure_* in include/sdns.h).

The network maps the box-filtered velocity u (3 fields) to the subgrid stress
tau_ij = bar(u_i u_j) - bar(u_i) bar(u_j) (6 fields: xx, yy, zz, xy, xz, yz):
three periodic 3x3x3 conv layers 3 -> 32 -> 32 -> 6 with bias and SiLU after
layers 1 and 2, bf16 weights/activations and FP32 accumulation on the matrix
cores.  Inputs are scaled by ``in_scale`` and outputs by ``out_scale``
(normalisation chosen from the training data).

Training pairs come from a resolved run: the physical DNS velocity is box
filtered by r on the GPU (``Plan.box_filter``) into (bar u, tau).  Training
(forward, loss, backward, Adam) runs on the GPU kernels; the host only loops.
The coarse run advances u_hat with the solver's RK4 where every stage adds the
Source P(-i k_j tau_ij) (projected and dealiased).
"""
import ctypes

import numpy as np

from ._abi import SdnsError, _P, _check, _ptr, load

__all__ = ["Closure", "init_params", "NPARAM"]

LAYERS = ((3, 32), (32, 32), (32, 6))
NPARAM = sum(27 * ci * co + co for ci, co in LAYERS)


def init_params(seed=0, scale=1.0):
    """Canonical FP32 parameter vector: uniform(-a, a) weights with
    a = scale * sqrt(3 / (27 cin)) (unit-variance preserving for unit inputs),
    zero biases, from a seeded PCG64 stream (layer by layer, C order)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    parts = []
    for cin, cout in LAYERS:
        lim = scale * np.sqrt(3.0 / (27 * cin))
        parts.append(rng.uniform(-lim, lim, 27 * cout * cin))
        parts.append(np.zeros(cout))
    return np.concatenate(parts).astype(np.float32)


class Closure:
    """Network parameters, Adam state and the GPU workspace of one coarse grid."""

    def __init__(self, N, params=None, seed=0):
        import torch
        lib = load()
        if not torch.cuda.is_available():
            raise SdnsError("no GPU visible: the HIP path has no CPU fallback")
        if lib.sdns_closure_nparams() != NPARAM:
            raise SdnsError("parameter count mismatch with libsdns_hip.so")
        self.N = tuple(int(n) for n in N)
        self.P = int(np.prod(self.N))
        self.device = torch.device("cuda", torch.cuda.current_device())
        h = _P()
        _check(lib.sdns_closure_create(ctypes.byref(h), (ctypes.c_int32 * 3)(*self.N),
                                       _P(torch.cuda.current_stream().cuda_stream)))
        self._h = h
        self._lib = lib
        p = init_params(seed) if params is None else np.asarray(params, dtype=np.float32)
        self.params = torch.from_numpy(p.copy()).to(self.device)
        self.grad = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.t = 0
        self.loss = torch.zeros(1, dtype=torch.float64, device=self.device)
        self.in_scale = 1.0
        self.out_scale = 1.0
        self.set_params(self.params)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.sdns_closure_destroy(h)
            self._h = None

    def _stream(self):
        import torch
        _check(self._lib.sdns_closure_set_stream(self._h, _P(torch.cuda.current_stream().cuda_stream)))

    def _phys(self, t, nc, name):
        import torch
        if t.device.type != "cuda" or t.dtype != torch.float32 or not t.is_contiguous() \
                or tuple(t.shape) != (nc,) + self.N:
            raise SdnsError("%s must be a contiguous float32 GPU tensor of shape %s" % (name, (nc,) + self.N))

    def set_params(self, params):
        import torch
        if params.dtype != torch.float32 or params.numel() != NPARAM or params.device.type != "cuda":
            raise SdnsError("params must be a float32 GPU tensor of %d entries" % NPARAM)
        if params.data_ptr() != self.params.data_ptr():
            self.params.copy_(params)
        self._stream()
        _check(self._lib.sdns_closure_set_params(self._h, _ptr(self.params)))

    def forward(self, u, tau=None):
        """tau = out_scale * net(in_scale * u) (planar FP32 fields)."""
        import torch
        self._phys(u, 3, "u")
        if tau is None:
            tau = torch.empty((6,) + self.N, dtype=torch.float32, device=self.device)
        self._phys(tau, 6, "tau")
        self._stream()
        _check(self._lib.sdns_closure_forward(self._h, _ptr(u), float(self.in_scale), float(self.out_scale),
                                              _ptr(tau)))
        return tau

    def train_step(self, u, tau):
        """Loss (device float64 tensor) and gradient (self.grad) of one training pair."""
        self._phys(u, 3, "u")
        self._phys(tau, 6, "tau")
        self._stream()
        _check(self._lib.sdns_closure_train_step(self._h, _ptr(u), _ptr(tau), float(self.in_scale),
                                                 float(self.out_scale), _ptr(self.grad), _ptr(self.loss)))
        return self.loss

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, grad=None):
        """One Adam step on the parameters with self.grad (or grad), then repack."""
        import torch
        g = self.grad if grad is None else grad
        self.t += 1
        _check(self._lib.sdns_adam_step(_ptr(self.params), _ptr(g), _ptr(self.m), _ptr(self.v), NPARAM,
                                        float(lr), float(beta1), float(beta2), float(eps), self.t,
                                        _P(torch.cuda.current_stream().cuda_stream)))
        _check(self._lib.sdns_closure_set_params(self._h, _ptr(self.params)))

    def rk4_steps(self, plan, u_hat, dts, nu, diag=None):
        """RK4 steps of the coarse plan with the learned Source at every stage."""
        import torch
        if plan.N != self.N or plan.precision != "single":
            raise SdnsError("the closure runs on a single-precision plan of its own grid")
        plan._spec(u_hat, 3, "u_hat")
        n = len(dts)
        if diag is not None and (diag.dtype != torch.float64 or diag.numel() < 2 * n or diag.device.type != "cuda"):
            raise SdnsError("diag must be a float64 GPU tensor of >= 2*nsteps entries")
        arr = (ctypes.c_double * max(n, 1))(*[float(d) for d in dts])
        plan._sync_stream()
        self._stream()
        _check(self._lib.sdns_closure_rk4_steps(self._h, plan._h, _ptr(u_hat), arr, n, float(nu),
                                                float(self.in_scale), float(self.out_scale),
                                                _ptr(diag) if diag is not None else None))
        return u_hat
