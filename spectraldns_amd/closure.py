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

    def rk4_steps(self, plan, u_hat, dts, nu, diag=None, per_stage=False):
        """RK4 steps of the coarse plan with the learned Source, evaluated once
        per step (held through the stages) or, with per_stage, at every stage."""
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
                                                float(self.in_scale), float(self.out_scale), int(bool(per_stage)),
                                                _ptr(diag) if diag is not None else None))
        return u_hat


# ---------------------------------------------------------------------- experiment pipeline
def shell_energy(plan, u_hat, nb):
    """E(k) = 0.5 sum over the shell k - 1/2 < |K| <= k + 1/2 of w |u_hat|^2, k = 1..nb
    (GPU shell sums, FP64), as a NumPy array."""
    sums, _ = plan.spectrum_sums(u_hat.contiguous(), nb)
    return 0.5 * sums.cpu().numpy()


class ForcedRun:
    """Forced isotropic run on one box: RK4 steps, each followed by the
    large-scale forcing of demo/Isotropic.py:154-184 that holds
    energy_fourier(u_hat) at its initial value on |K| <= kf.  With a closure
    the RK4 stages add its learned Source."""

    def __init__(self, N, nu, dt, u_hat=None, seed=1234, kf=2.0, precision="single", dealias="2/3-rule",
                 k0=4.0, energy=0.5, closure=None, target=None, per_stage=False, dealias_mode=0):
        import torch
        from ._abi import Plan
        from .isotropic import random_solenoidal_field
        self.N = tuple(int(n) for n in N)
        self.nu, self.dt, self.kf = float(nu), float(dt), float(kf)
        self.plan = Plan(self.N, (2 * np.pi,) * 3, precision, dealias, True, dealias_mode)
        cd = torch.complex64 if precision == "single" else torch.complex128
        if u_hat is None:
            u_hat = torch.from_numpy(random_solenoidal_field(self.N, seed=seed, k0=k0, energy=energy)).to(cd).cuda()
        self.u_hat = u_hat.contiguous()
        self.plan.mask_nyquist(self.u_hat)
        self.closure = closure
        self.per_stage = per_stage
        self._e = torch.zeros(2, dtype=torch.float64, device="cuda")
        self.target = self.energy_fourier() if target is None else float(target)
        self.steps_done = 0

    def energy_fourier(self):
        self.plan.energy(self.u_hat, self.nu, self._e)
        return 2.0 * float(self._e[0].item())

    def diagnostics(self):
        """(kinetic energy, resolved dissipation) in FP64."""
        self.plan.energy(self.u_hat, self.nu, self._e)
        e = self._e.cpu().numpy()
        return float(e[0]), float(e[1])

    def step(self, n=1):
        for _ in range(n):
            if self.closure is None:
                self.plan.rk4_steps(self.u_hat, [self.dt], self.nu)
            else:
                self.closure.rk4_steps(self.plan, self.u_hat, [self.dt], self.nu, per_stage=self.per_stage)
            self.plan.force(self.u_hat, self.target, self.kf)
        self.steps_done += n

    def physical(self):
        import torch
        u = torch.empty((3,) + self.N, dtype=self.plan.real_dtype, device="cuda")
        return self.plan.backward(self.u_hat, u)

    def filtered(self, r):
        """Box-filtered velocity and subgrid stress (physical, FP32) of the current state."""
        return self.plan.box_filter(self.physical(), r)


def keep_mask(plan, kc):
    """Boolean spectral mask |k_i| <= kc on every axis (tensor product, as the
    solver's 2/3-rule keep sets), Nyquist planes excluded."""
    import torch
    N = plan.N
    kx = torch.fft.fftfreq(N[0], 1.0 / N[0], device="cuda").abs()
    ky = torch.fft.fftfreq(N[1], 1.0 / N[1], device="cuda").abs()
    kz = torch.arange(N[2] // 2 + 1, device="cuda", dtype=kx.dtype)
    m = (kx[:, None, None] <= kc) & (ky[None, :, None] <= kc) & (kz[None, None, :] <= kc)
    m &= (kx[:, None, None] < N[0] // 2) & (ky[None, :, None] < N[1] // 2) & (kz[None, None, :] < N[2] // 2)
    return m


def coarse_state(plan_c, ubar, kc=None):
    """Spectral state of the coarse plan from a filtered physical velocity:
    Nyquist masked and, with kc, band-limited to |k_i| <= kc."""
    import torch
    cd = torch.complex64 if plan_c.precision == "single" else torch.complex128
    u_hat = torch.empty((3,) + plan_c.spectral_shape, dtype=cd, device="cuda")
    plan_c.forward(ubar.to(plan_c.real_dtype).contiguous(), u_hat)
    plan_c.mask_nyquist(u_hat)
    if kc is not None:
        u_hat *= keep_mask(plan_c, kc)
    return u_hat


def band_limit(plan_c, f, kc):
    """Physical fields f (c, n0, n1, n2) with every mode |k_i| > kc removed
    (forward transform, mask, backward transform on the coarse plan)."""
    import torch
    out = torch.empty_like(f)
    cd = torch.complex64 if plan_c.precision == "single" else torch.complex128
    m = keep_mask(plan_c, kc)
    for c0 in range(0, f.shape[0], 3):
        blk = f[c0:c0 + 3].contiguous()
        h = torch.empty((blk.shape[0],) + plan_c.spectral_shape, dtype=cd, device="cuda")
        plan_c.forward(blk, h)
        h *= m
        o = torch.empty_like(blk)
        plan_c.backward(h, o)
        out[c0:c0 + 3] = o
    return out


def filtered_pair(run, r, plan_c, kc):
    """Training pair for the composed filter G = (spectral cutoff kc) o (box r):
    bar u = G u and tau_ij = G(u_i u_j) - bar u_i bar u_j (deviatoric), from the
    box filter's block means of u and u_i u_j on the coarse grid."""
    import torch
    ub, tau = run.filtered(r)
    ub = ub.float().contiguous()
    T = tau.float().clone()
    for m, (i, j) in enumerate(_PAIRS):
        T[m] += ub[i] * ub[j]  # box means of the products
    ubl = band_limit(plan_c, ub, kc)
    Tl = band_limit(plan_c, T, kc)
    for m, (i, j) in enumerate(_PAIRS):
        Tl[m] -= ubl[i] * ubl[j]
    return ubl.contiguous(), deviatoric(Tl).contiguous()


def deviatoric(tau):
    """tau_ij - delta_ij tau_kk / 3 (in place on the (6, ...) xx, yy, zz, xy, xz, yz
    fields).  The isotropic part only adds a gradient to the Source, which the
    projection removes, so the network is trained on the deviatoric part."""
    tr = (tau[0] + tau[1] + tau[2]) / 3.0
    tau[:3] -= tr
    return tau


# component pairs of the six stress fields
_PAIRS = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))
_INDEX = {p: m for m, p in enumerate(_PAIRS)}
_INDEX.update({(j, i): m for (i, j), m in _INDEX.items()})


def symmetry(u, tau, perm, signs):
    """Image of a training pair under the cube symmetry x'_i = s_i x_{perm[i]}
    (periodic; NS is invariant under rotations and reflections):
    u'_i(x') = s_i u_{perm[i]}(x), tau'_ij(x') = s_i s_j tau_{perm[i] perm[j]}(x)."""
    import torch
    u2 = u[list(perm)].permute(0, *[1 + p for p in perm])
    t2 = torch.stack([tau[_INDEX[(perm[i], perm[j])]] for i, j in _PAIRS]).permute(0, *[1 + p for p in perm])
    for ax in range(3):
        if signs[ax] < 0:
            # x -> -x on the periodic lattice: index i -> (-i) mod N
            u2 = torch.roll(torch.flip(u2, dims=(1 + ax,)), 1, dims=1 + ax)
            t2 = torch.roll(torch.flip(t2, dims=(1 + ax,)), 1, dims=1 + ax)
    sg = torch.tensor(signs, dtype=u.dtype, device=u.device)
    u2 = u2 * sg.view(3, 1, 1, 1)
    st = torch.tensor([signs[i] * signs[j] for i, j in _PAIRS], dtype=tau.dtype, device=tau.device)
    t2 = t2 * st.view(6, 1, 1, 1)
    return u2.contiguous(), t2.contiguous()


def augment(pairs, count, seed=0):
    """`count` random cube-symmetry images of randomly chosen pairs."""
    import itertools
    rng = np.random.Generator(np.random.PCG64(seed))
    perms = list(itertools.permutations(range(3)))
    out = []
    for _ in range(count):
        u, t = pairs[int(rng.integers(len(pairs)))]
        perm = perms[int(rng.integers(6))]
        signs = tuple(int(x) for x in rng.choice([-1, 1], 3))
        out.append(symmetry(u, t, perm, signs))
    return out


def scales(pairs):
    """(in_scale, out_scale): 1 / rms of the filtered velocity, rms of tau."""
    import torch
    su = torch.sqrt(torch.stack([torch.mean(u.double() ** 2) for u, _ in pairs]).mean()).item()
    st = torch.sqrt(torch.stack([torch.mean(t.double() ** 2) for _, t in pairs]).mean()).item()
    return 1.0 / su, st


def train(closure, pairs, iters, lr0=2e-3, lr1=1e-4, log_every=100, val=None):
    """Adam over the pairs (cycled), learning rate decayed geometrically from
    lr0 to lr1.  Returns [(iteration, mean training loss since the last log,
    validation loss or None)]."""
    import torch
    hist = []
    acc = torch.zeros(1, dtype=torch.float64, device="cuda")
    cnt = 0
    for it in range(iters):
        u, t = pairs[it % len(pairs)]
        acc += closure.train_step(u, t)
        cnt += 1
        lr = lr0 * (lr1 / lr0) ** (it / max(iters - 1, 1))
        closure.adam(lr)
        if (it + 1) % log_every == 0 or it == iters - 1:
            vl = None
            if val is not None:
                vl = float(np.mean([validation_loss(closure, u2, t2) for u2, t2 in val]))
            hist.append((it + 1, float(acc.item()) / cnt, vl))
            acc.zero_()
            cnt = 0
    return hist


def validation_loss(closure, u, t):
    """Normalised loss of one pair without touching the gradients: mean over
    points and components of ((tau_pred - tau) / out_scale)^2."""
    tau = closure.forward(u)
    return float(torch_mean_sq((tau - t) / closure.out_scale))


def zero_model_loss(closure, t):
    return float(torch_mean_sq(t / closure.out_scale))


def torch_mean_sq(x):
    return (x.double() ** 2).mean().item()
