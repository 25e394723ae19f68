"""Isotropic turbulence initial field (the construction of demo/Isotropic.py:29-76).

Rogallo's random-phase solenoidal field: for every mode
    u_hat = alpha e1 + beta e2,  alpha = A e^{i theta1} cos(phi),  beta = A e^{i theta2} sin(phi)
with (e1, e2) the unit vectors normal to k used at Isotropic.py:50-52 and
A = sqrt(E(|k|) / (4 pi |k|^2)).  The spectrum is E(k) ∝ k^4 exp(-2 (k/k0)^2)
and the phases theta1, theta2, phi are uniform in [0, 2 pi) (the reference
multiplies its samples by 2j*pi, which makes the phases imaginary; that quirk
is not reproduced).  Then, as the reference does: Nyquist mask, Hermitian
symmetry by a backward/forward round trip, projection onto k.u_hat = 0, zero
mean and a rescale to the requested kinetic energy 0.5*sum w|u_hat|^2.

This is host code: the field is generated once with NumPy (seeded PCG64, so
the GPU run and any checker start from bit-identical data) and then copied to
the device.
"""
import numpy as np

__all__ = ["random_solenoidal_field", "hermitian_weights"]


def hermitian_weights(N2):
    """w = 1 on kz = 0 and kz = N2/2, 2 elsewhere (demo/Isotropic.py:90-93)."""
    w = np.full(N2 // 2 + 1, 2.0)
    w[0] = 1.0
    w[-1] = 1.0
    return w


def random_solenoidal_field(N, L=(2 * np.pi,) * 3, seed=1234, k0=4.0, energy=0.5):
    """complex128 array (3, N0, N1, N2//2+1) with the properties listed above."""
    N = tuple(int(n) for n in N)
    Nh = N[2] // 2 + 1
    rng = np.random.Generator(np.random.PCG64(seed))
    shape = (N[0], N[1], Nh)
    theta1 = rng.uniform(0.0, 2 * np.pi, shape)
    theta2 = rng.uniform(0.0, 2 * np.pi, shape)
    phi = rng.uniform(0.0, 2 * np.pi, shape)
    k1 = (np.fft.fftfreq(N[0], 1.0 / N[0]) * 2 * np.pi / L[0])[:, None, None]
    k2 = (np.fft.fftfreq(N[1], 1.0 / N[1]) * 2 * np.pi / L[1])[None, :, None]
    k3 = (np.arange(Nh) * 2 * np.pi / L[2])[None, None, :]
    K2 = k1 ** 2 + k2 ** 2 + k3 ** 2
    k = np.sqrt(K2)
    kk = np.where(K2 == 0, 1.0, K2)
    kn = np.where(k == 0, 1.0, k)
    ksq = np.sqrt(k1 ** 2 + k2 ** 2)
    ksq = np.where(ksq == 0, 1.0, ksq)
    Ek = k ** 4 * np.exp(-2.0 * (k / k0) ** 2)
    amp = np.sqrt(Ek / (4.0 * np.pi * kk))
    alpha = amp * np.exp(1j * theta1) * np.cos(phi)
    beta = amp * np.exp(1j * theta2) * np.sin(phi)
    u = np.empty((3,) + shape, dtype=np.complex128)
    u[0] = (alpha * kn * k2 + beta * k1 * k3) / (kn * ksq)
    u[1] = (beta * k2 * k3 - alpha * kn * k1) / (kn * ksq)
    u[2] = beta * ksq / kn
    # Nyquist planes
    u[:, N[0] // 2] = 0
    u[:, :, N[1] // 2] = 0
    u[:, :, :, N[2] // 2] = 0
    # Hermitian symmetry on kz = 0 (and kz = N/2) by a round trip (Isotropic.py:56-57)
    n = np.prod(N)
    phys = np.fft.irfftn(u, s=N, axes=(1, 2, 3)) * n
    u = np.fft.rfftn(phys, axes=(1, 2, 3)) / n
    # projection onto divergence-free fields (Isotropic.py:61)
    Kv = (k1, k2, k3)
    div = Kv[0] * u[0] + Kv[1] * u[1] + Kv[2] * u[2]
    for i in range(3):
        u[i] -= div * Kv[i] / kk
    u[:, 0, 0, 0] = 0.0
    u[:, N[0] // 2] = 0
    u[:, :, N[1] // 2] = 0
    u[:, :, :, N[2] // 2] = 0
    w = hermitian_weights(N[2])
    e = 0.5 * float(np.sum(w * (u.real ** 2 + u.imag ** 2)))
    u *= np.sqrt(energy / e)
    return u


def random_solenoidal_field_gpu(plan, seed=1234, k0=4.0, energy=0.5):
    """The same construction on the GPU, for boxes whose host generation would
    take minutes (1024^3): Rogallo amplitudes and basis, phases from torch's
    seeded device generator (so the draws differ from the PCG64 host field),
    Hermitian symmetry by a backward/forward round trip with the plan's
    transforms, projection by the plan's pressure kernel (nu = 0), Nyquist
    planes and mean mode zeroed, energy scaled with the FP64 energy kernel.
    Returns a (3, N0, N1, N2//2+1) tensor of the plan's complex dtype."""
    import torch
    N, L = plan.N, plan.L
    Nh = N[2] // 2 + 1
    dev = plan.device
    rd, cd = plan.real_dtype, plan.complex_dtype
    g = torch.Generator(device=dev)
    g.manual_seed(int(seed))
    k1 = (torch.fft.fftfreq(N[0], 1.0 / N[0], device=dev, dtype=rd) * (2 * np.pi / L[0])).view(-1, 1, 1)
    k2 = (torch.fft.fftfreq(N[1], 1.0 / N[1], device=dev, dtype=rd) * (2 * np.pi / L[1])).view(1, -1, 1)
    k3 = (torch.arange(Nh, device=dev, dtype=rd) * (2 * np.pi / L[2])).view(1, 1, -1)
    shape = (N[0], N[1], Nh)
    u = torch.empty((3,) + shape, dtype=cd, device=dev)
    K2 = k1 * k1 + k2 * k2 + k3 * k3
    k = torch.sqrt(K2)
    kk = torch.where(K2 == 0, torch.ones_like(K2), K2)
    kn = torch.where(k == 0, torch.ones_like(k), k)
    ksq = torch.sqrt(k1 * k1 + k2 * k2).expand(shape)
    ksq = torch.where(ksq == 0, torch.ones_like(ksq), ksq)
    amp = torch.sqrt(k ** 4 * torch.exp(-2.0 * (k / k0) ** 2) / (4.0 * np.pi * kk))
    del K2, k
    two_pi = 2 * np.pi
    th1 = torch.rand(shape, generator=g, device=dev, dtype=rd) * two_pi
    th2 = torch.rand(shape, generator=g, device=dev, dtype=rd) * two_pi
    phi = torch.rand(shape, generator=g, device=dev, dtype=rd) * two_pi
    alpha = amp * torch.cos(phi) * torch.polar(torch.ones_like(th1), th1)
    beta = amp * torch.sin(phi) * torch.polar(torch.ones_like(th2), th2)
    del th1, th2, phi, amp
    u[0] = (alpha * kn * k2 + beta * k1 * k3) / (kn * ksq)
    u[1] = (beta * k2 * k3 - alpha * kn * k1) / (kn * ksq)
    u[2] = beta * ksq / kn
    del alpha, beta, ksq, kn, kk
    plan.mask_nyquist(u)
    phys = torch.empty((3,) + tuple(N), dtype=rd, device=dev)
    plan.backward(u, phys)
    plan.forward(phys, u)
    del phys
    plan.add_pressure_diffusion(u, u, 0.0)  # u -= K (K . u) / K^2 (nu = 0)
    u[:, 0, 0, 0] = 0
    plan.mask_nyquist(u)
    e = torch.zeros(2, dtype=torch.float64, device=dev)
    plan.energy(u, 0.0, e)
    u *= float(np.sqrt(energy / float(e[0].item())))
    return u
