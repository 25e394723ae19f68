"""Fourier spaces over a HIP plan: the parts of shenfun's TensorProductSpace /
VectorSpace that spectralDNS/solvers/NS.py and the demos use (forward,
backward, shape, local_mesh, local_wavenumbers, get_mask_nyquist,
mask_nyquist) plus ``Array``, ``Function`` and ``energy_fourier``.

Arrays are torch tensors on the GPU.  Every transform runs in libsdns_hip.
"""
import numpy as np

from ._abi import Plan, SdnsError

__all__ = ["FourierSpace", "Array", "Function", "energy_fourier", "spectrum", "Wavenumbers", "NyquistMask",
           "WorkArrays"]


class Wavenumbers(list):
    """List of broadcastable wavenumber tensors (K[0] (N0,1,1), K[1] (1,N1,1),
    K[2] (1,1,Nh)) that remembers its plan, so cross2(c, K, a) can run on the GPU."""

    def __init__(self, items, plan):
        super().__init__(items)
        self.plan = plan


class NyquistMask:
    """Marker returned by get_mask_nyquist(): the mask itself is an index test
    inside the kernels (NS.py:34, 253-254)."""

    def __init__(self, plan):
        self.plan = plan


class FourierSpace:
    """rank 0 (scalar) or rank 1 (3-vector) Fourier space of one box.
    ``dealiased`` marks the padded / truncated space of the nonlinear term
    (NS.py:29-32): Tp / VTp, whose transforms are sdns_backward_dealiased and
    sdns_forward_dealiased (the Vortex RHS runs them fused in its passes)."""

    def __init__(self, plan, rank=0, dealiased=False):
        self.plan = plan
        self.rank = rank
        self.dealiased = dealiased

    # -- shapes ---------------------------------------------------------------
    def shape(self, forward_output=False):
        if forward_output:
            s = self.plan.spectral_shape
        elif self.dealiased:
            s = self.plan.dealiased_shape
        else:
            s = self.plan.N
        return ((3,) if self.rank else ()) + tuple(s)

    @property
    def dimensions(self):
        return 3

    # -- transforms -----------------------------------------------------------
    # The dealiased space (Tp, VTp) transforms between the plan's modes and its
    # dealiased grid: the padded 3N/2 grid of the 3/2 rule, or the plan's grid
    # with the 2/3-rule input truncation (backward) and a plain forward.
    def forward(self, u, u_hat=None):
        import torch
        if u_hat is None:
            u_hat = torch.empty(u.shape[:-3] + self.plan.spectral_shape, dtype=self.plan.complex_dtype,
                                device=u.device)
        if self.dealiased:
            return self.plan.forward_dealiased(u, u_hat)
        return self.plan.forward(u, u_hat)

    def backward(self, u_hat, u=None):
        import torch
        if u is None:
            shp = self.plan.dealiased_shape if self.dealiased else self.plan.physical_shape
            u = torch.empty(u_hat.shape[:-3] + shp, dtype=self.plan.real_dtype, device=u_hat.device)
        if self.dealiased:
            return self.plan.backward_dealiased(u_hat, u)
        return self.plan.backward(u_hat, u)

    # -- mesh / wavenumbers ---------------------------------------------------
    def local_mesh(self, broadcast=True):
        """x_i = j L_i / N_i (the Fourier points of shenfun's local_mesh)."""
        import torch
        X = []
        for i in range(3):
            x = (np.arange(self.plan.N[i]) * self.plan.L[i] / self.plan.N[i])
            t = torch.tensor(x, dtype=self.plan.real_dtype, device=self.plan.device)
            shp = [1, 1, 1]
            shp[i] = self.plan.N[i]
            X.append(t.reshape(shp) if broadcast else t)
        return X

    def local_wavenumbers(self, broadcast=True, scaled=True):
        """(2 pi/L_i) fftfreq(N_i) N_i on axes 0, 1 and rfftfreq on axis 2, cast to
        the working precision (NS.py:37-41)."""
        import torch
        N = self.plan.N
        ks = [np.fft.fftfreq(N[0], 1.0 / N[0]), np.fft.fftfreq(N[1], 1.0 / N[1]),
              np.arange(N[2] // 2 + 1, dtype=float)]
        K = []
        for i in range(3):
            k = ks[i] * (2 * np.pi / self.plan.L[i]) if scaled else ks[i]
            t = torch.tensor(k.astype(np.float64), device=self.plan.device).to(self.plan.real_dtype)
            shp = [1, 1, 1]
            shp[i] = len(k)
            K.append(t.reshape(shp) if broadcast else t)
        return Wavenumbers(K, self.plan)

    def get_mask_nyquist(self):
        return NyquistMask(self.plan)

    def mask_nyquist(self, u_hat, mask=None):
        """Zero the Nyquist planes in place (shenfun Function.mask_nyquist)."""
        if u_hat.dim() == 4 and u_hat.shape[0] == 3:
            return self.plan.mask_nyquist(u_hat)
        N = self.plan.N
        u_hat[..., N[0] // 2, :, :] = 0
        u_hat[..., :, N[1] // 2, :] = 0
        u_hat[..., :, :, N[2] // 2] = 0
        return u_hat


def Array(space):
    """Zero physical-space array of a space (shenfun Array)."""
    import torch
    return torch.zeros(space.shape(False), dtype=space.plan.real_dtype, device=space.plan.device)


def Function(space):
    """Zero spectral-space array of a space (shenfun Function)."""
    import torch
    return torch.zeros(space.shape(True), dtype=space.plan.complex_dtype, device=space.plan.device)


class WorkArrays(dict):
    """Cache of scratch arrays keyed (array, index, zero_fill) like shenfun's
    CachedArrayDict used by NS.py (work[(a, 0, False)])."""

    def __getitem__(self, key):
        a, i, fill = key
        k = (tuple(a.shape), a.dtype, i)
        if dict.__contains__(self, k):
            arr = dict.__getitem__(self, k)
            if fill:
                arr.zero_()
            return arr
        import torch
        arr = torch.zeros_like(a) if fill else torch.empty_like(a)
        dict.__setitem__(self, k, arr)
        return arr



def energy_fourier(u_hat, T):
    """sum over components and modes of w |u_hat|^2, w = 1 on kz = 0 and kz = N/2
    and 2 elsewhere, reduced in FP64 on the GPU (shenfun.fourier.energy_fourier
    as used by demo/Isotropic.py:67, 167-168)."""
    import torch
    plan = T.plan
    out = torch.empty(2, dtype=torch.float64, device=plan.device)
    plan.energy(u_hat.contiguous(), 0.0, out)
    return float(out[0].item()) * 2.0


def spectrum(u_hat, T, N=None):
    """Shell-binned energy spectrum of demo/Isotropic.py:88-115 on the GPU.

    Returns (Ek, bins) as NumPy arrays with the reference's definition: shells
    (k - 1/2, k + 1/2] for k = 1..Nb-1 with Nb = int(sqrt(sum((N/2)^2)/3)),
    Ek = (k_hi^3 - k_lo^3) * (4 pi/3) * sum(uiui) / count, Ek[Nb-1] = 0."""
    plan = T.plan
    N = np.array(plan.N if N is None else N)
    Nb = int(np.sqrt(np.sum((N / 2) ** 2) / 3))
    bins = np.array(range(0, Nb)) + 0.5
    Ek = np.zeros(Nb)
    if Nb < 2:
        return Ek, bins
    sums, counts = plan.spectrum_sums(u_hat.contiguous(), Nb - 1)
    sums = sums.cpu().numpy() * (4.0 / 3.0 * np.pi)
    counts = counts.cpu().numpy()
    for i, k in enumerate(bins[1:]):
        k0 = bins[i]
        Ek[i] = (k ** 3 - k0 ** 3) * sums[i]
        if counts[i] != 0:
            Ek[i] = Ek[i] / counts[i]
    return Ek, bins
