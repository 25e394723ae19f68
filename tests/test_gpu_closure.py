"""Learned-closure kernels on the GPU against the NumPy closure oracle.

Bars: network output within 5e-3 relative L2 of the oracle run with the same
bf16 storage points (bf16 rounding of intermediate values may differ by one
unit where FP32 accumulation orders differ); every gradient group within 2e-2
relative L2 and the loss within 1e-3; Adam within 1e-6; the box filter within
1e-12 (FP64) / 1e-6 (FP32); zero parameters reproduce the plain RK4 step
bit for bit; one closure RK4 step within 1e-2 of the oracle's increment.
"""
import numpy as np
import pytest

from oracle import closure as OC
from oracle.ns import Grid

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
TWO_PI = 2 * np.pi


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def split(vec):
    """canonical vector -> oracle ([W1, W2, W3], [b1, b2, b3])"""
    W, b, o = [], [], 0
    for cin, cout in OC.LAYERS:
        W.append(vec[o:o + 27 * cout * cin].reshape(27, cout, cin).astype(np.float64))
        o += 27 * cout * cin
        b.append(vec[o:o + cout].astype(np.float64))
        o += cout
    return W, b


def smooth_field(N, seed, nc=3):
    from spectraldns_amd.isotropic import random_solenoidal_field
    g = Grid(N, (TWO_PI,) * 3, "double", "None")
    u = g.backward(random_solenoidal_field(N, seed=seed, k0=3.0, energy=0.5))
    if nc == 3:
        return u
    return np.concatenate([u, u[:nc - 3] ** 2])


def make(N, seed=3, bias_scale=0.1):
    from spectraldns_amd.closure import Closure, init_params
    print("\n[case] closure N=%s seed=%d" % (N, seed), flush=True)
    p = init_params(seed, scale=1.5)
    rng = np.random.default_rng(seed)
    o = 0
    for cin, cout in OC.LAYERS:  # nonzero biases so that they are exercised
        o += 27 * cout * cin
        p[o:o + cout] = rng.uniform(-bias_scale, bias_scale, cout)
        o += cout
    return Closure(N, params=p), p


N0 = (32, 32, 32)


def test_closure_forward_vs_oracle():
    c, p = make(N0)
    u = smooth_field(N0, 5)
    c.in_scale, c.out_scale = 2.0, 0.5
    tau = c.forward(torch.from_numpy(u).float().cuda().contiguous())
    torch.cuda.synchronize()
    W, b = split(p)
    x = np.moveaxis(u.astype(np.float32) * np.float32(2.0), 0, -1)
    y, _ = OC.forward(x, W, b, bf16_storage=True)
    ref = 0.5 * np.moveaxis(y, -1, 0)
    err = rel(tau.cpu().numpy(), ref)
    print("forward rel err %.2e" % err)
    assert err < 5e-3


def test_closure_gradients_vs_oracle():
    c, p = make(N0, seed=4)
    u = smooth_field(N0, 6)
    t = smooth_field(N0, 7, nc=6) * 0.3
    c.in_scale, c.out_scale = 1.5, 2.0
    ud = torch.from_numpy(u).float().cuda().contiguous()
    td = torch.from_numpy(t).float().cuda().contiguous()
    loss = c.train_step(ud, td)
    torch.cuda.synchronize()
    g = c.grad.cpu().numpy()
    W, b = split(p)
    x = np.moveaxis(u.astype(np.float32) * np.float32(1.5), 0, -1)
    tt = np.moveaxis(t.astype(np.float32) * np.float32(0.5), 0, -1)
    y, saved = OC.forward(x, W, b, bf16_storage=True)
    dW, db = OC.backward(y, tt, saved, bf16_storage=True)
    assert abs(float(loss.item()) - OC.loss(y, tt)) <= 1e-3 * OC.loss(y, tt)
    gW, gb = split(g)
    for layer in range(3):
        ew, eb = rel(gW[layer], dW[layer]), rel(gb[layer], db[layer])
        print("layer %d: dW rel %.2e, db rel %.2e" % (layer + 1, ew, eb))
        assert ew < 2e-2 and eb < 2e-2


def test_adam_step():
    from spectraldns_amd.closure import Closure, NPARAM
    c = Closure(N0)
    rng = np.random.default_rng(1)
    p0 = c.params.cpu().numpy().astype(np.float64)
    m = np.zeros(NPARAM)
    v = np.zeros(NPARAM)
    lr, b1, b2, eps = 1e-2, 0.9, 0.999, 1e-8
    for t in range(1, 4):
        gr = rng.standard_normal(NPARAM).astype(np.float32)
        c.grad.copy_(torch.from_numpy(gr))
        c.adam(lr, b1, b2, eps)
        m = b1 * m + (1 - b1) * gr
        v = b2 * v + (1 - b2) * gr.astype(np.float64) ** 2
        p0 = p0 - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    torch.cuda.synchronize()
    assert np.max(np.abs(c.params.cpu().numpy() - p0)) < 1e-6


@pytest.mark.parametrize("precision", ["double", "single"])
def test_box_filter_vs_oracle(precision):
    from spectraldns_amd._abi import Plan
    N = (64, 32, 48)
    plan = Plan(N, (TWO_PI,) * 3, precision, "None", True, 0, tuned=False)
    u = smooth_field(N, 9)
    dt = torch.float64 if precision == "double" else torch.float32
    ub, tau = plan.box_filter(torch.from_numpy(u).to(dt).cuda().contiguous(), 4)
    torch.cuda.synchronize()
    tol = 1e-12 if precision == "double" else 2e-6
    assert rel(ub.cpu().numpy(), OC.box_filter(u, 4)) < tol
    assert rel(tau.cpu().numpy(), OC.sgs_stress(u, 4)) < (tol if precision == "double" else 1e-5)


def test_zero_closure_equals_plain_step():
    from spectraldns_amd._abi import Plan
    from spectraldns_amd.closure import Closure, NPARAM
    from spectraldns_amd.isotropic import random_solenoidal_field
    plan = Plan(N0, (TWO_PI,) * 3, "single", "2/3-rule", True, 0, tuned=False)
    c = Closure(N0, params=np.zeros(NPARAM, dtype=np.float32))
    u0 = random_solenoidal_field(N0, seed=2, k0=3.0, energy=0.5)
    ua = torch.from_numpy(u0).to(torch.complex64).cuda()
    ub = ua.clone()
    plan.rk4_steps(ua, [0.005] * 3, 0.005)
    c.rk4_steps(plan, ub, [0.005] * 3, 0.005)
    torch.cuda.synchronize()
    assert torch.equal(ua, ub)


@pytest.mark.parametrize("per_stage", [True, False])
def test_closure_step_vs_oracle(per_stage):
    """One RK4 step with the learned Source: GPU increment vs the oracle's, where
    the oracle evaluates the same network (bf16 storage) on every stage, or
    once from the step's initial state and holds it (per_stage False)."""
    from spectraldns_amd._abi import Plan
    from spectraldns_amd.isotropic import random_solenoidal_field
    c, p = make(N0, seed=8, bias_scale=0.0)
    c.in_scale, c.out_scale = 1.0, 0.05
    plan = Plan(N0, (TWO_PI,) * 3, "single", "2/3-rule", True, 0, tuned=False)
    u0 = random_solenoidal_field(N0, seed=3, k0=3.0, energy=0.5)
    nu, dt = 0.01, 0.01
    ud = torch.from_numpy(u0).to(torch.complex64).cuda()
    c.rk4_steps(plan, ud, [dt], nu, per_stage=per_stage)
    torch.cuda.synchronize()
    g = Grid(N0, (TWO_PI,) * 3, "double", "2/3-rule", True, 0)
    W, b = split(p)
    K = g.K

    def source(uh):
        u = g.backward(uh).astype(np.float32)
        x = np.moveaxis(u * np.float32(c.in_scale), 0, -1)
        y, _ = OC.forward(x, W, b, bf16_storage=True)
        tau = np.moveaxis(y, -1, 0).astype(np.float32) * np.float32(c.out_scale)
        th = g.forward(tau.astype(np.float64))
        idx = [(0, 3, 4), (3, 1, 5), (4, 5, 2)]
        f = np.array([-1j * sum(K[j] * th[idx[i][j]] for j in range(3)) for i in range(3)])
        K2 = sum(k * k for k in K)
        kf = sum(K[j] * f[j] for j in range(3)) / np.where(K2 == 0, 1, K2)
        f = np.array([f[i] - K[i] * kf for i in range(3)])
        mask = np.ones(K2.shape, dtype=bool)
        for ax in range(3):
            kmax = max(g.keep[ax][0])  # input keep range (neg, pos) of the dealiased transforms
            shape = [1, 1, 1]
            shape[ax] = -1
            mask &= (np.abs(g.kint[ax]) <= kmax).reshape(shape)
        N = g.N
        mask[N[0] // 2] = False
        mask[:, N[1] // 2] = False
        mask[:, :, N[2] // 2] = False
        return f * mask

    a = [1 / 6, 1 / 3, 1 / 3, 1 / 6]
    bb = [0.5, 0.5, 1.0]
    u1 = u0.copy()
    u2 = u0.copy()
    us = u0.copy()
    held = source(u0)
    for k in range(4):
        r = g.compute_rhs(us, nu, source=source(us) if per_stage else held)
        if k < 3:
            us = u1 + bb[k] * dt * r
        u2 = u2 + a[k] * dt * r
    du_gpu = ud.cpu().numpy() - u0
    du_ref = u2 - u0
    s_only = [source(u0)]
    print("increment rel err %.2e; |source|/|rhs| %.2e" % (
        rel(du_gpu, du_ref), np.linalg.norm(s_only[0]) / np.linalg.norm(g.compute_rhs(u0, nu))))
    assert np.linalg.norm(s_only[0]) > 1e-3 * np.linalg.norm(g.compute_rhs(u0, nu))
    assert rel(du_gpu, du_ref) < 1e-2


def test_training_beats_zero_model_on_held_out_pairs():
    """Pairs from a forced 64^3 run filtered by 2; after a short training the
    held-out loss is well below the zero-model loss (what the network gains
    over a coarse run without a model, a priori)."""
    from spectraldns_amd import closure as CL
    print("\n[case] closure training 64^3 -> 32^3", flush=True)
    dns = CL.ForcedRun((64, 64, 64), 0.004, 0.004, seed=11, kf=2.0, k0=3.0)
    dns.step(60)
    pairs = []
    for _ in range(8):
        dns.step(5)
        ub, tau = dns.filtered(2)
        pairs.append((ub.contiguous(), tau.contiguous()))
    tr, va = pairs[:6], pairs[6:]
    net = CL.Closure((32, 32, 32), seed=1)
    net.in_scale, net.out_scale = CL.scales(tr)
    zero = float(np.mean([CL.zero_model_loss(net, t) for _, t in va]))
    first = float(np.mean([CL.validation_loss(net, u, t) for u, t in va]))
    hist = CL.train(net, tr, 300, lr0=3e-3, lr1=3e-4, log_every=100, val=va)
    print("zero-model %.4e, untrained %.4e, history %s" % (zero, first, hist))
    assert np.isfinite(hist[-1][1]) and hist[-1][1] < hist[0][1]
    assert hist[-1][2] < 0.5 * zero
