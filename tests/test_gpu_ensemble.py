"""Batched uneven-size ensembles on the GPU (sdns_batch_* through ctypes).

Bars: every box of an FP64 batch within 1e-12 relative L2 of the oracle after
N RK4 steps (the oracle advances each box on its own); an FP32 batch within
1e-5 of the FP64 oracle and within 1e-5 of the single-box FP32 path; graph
replay identical (bitwise) to eager launches; per-box FP64 energy and
dissipation equal to the oracle's.
"""
import numpy as np
import pytest

from oracle.ns import Grid

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
TWO_PI = 2 * np.pi


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def init(N, L, i, seed=7):
    from spectraldns_amd.isotropic import random_solenoidal_field
    from spectraldns_amd.ensemble import taylor_green_hat
    if i % 2 == 0 and L == (TWO_PI,) * 3:
        return taylor_green_hat(N, L)
    return random_solenoidal_field(N, L, seed=seed + i, k0=min(N) / 8 + 1, energy=0.5)


def run_batch(boxes, precision, rule, nus, dts, nsteps, graph=True, mode=0, split=None):
    from spectraldns_amd._abi import Batch
    print("\n[case] batch %s %s %s steps=%d graph=%s" % ([b[0] for b in boxes], precision, rule, nsteps, graph),
          flush=True)
    cd = torch.complex64 if precision == "single" else torch.complex128
    u0 = [init(N, L, i) for i, (N, L) in enumerate(boxes)]
    u = [torch.from_numpy(a).to(cd).cuda().contiguous() for a in u0]
    b = Batch(boxes, precision, rule, True, mode)
    b.use_graph(graph)
    b.set_state(u)
    b.set_params(nus, dts)
    diag = torch.zeros(2 * len(boxes), dtype=torch.float64, device="cuda")
    for n in (split or [nsteps]):
        b.rk4_steps(n, diag)
    torch.cuda.synchronize()
    return u0, [t.cpu().numpy() for t in u], diag.cpu().numpy()


BOXES_23 = [((16, 16, 16), (TWO_PI,) * 3), ((32, 32, 32), (TWO_PI,) * 3), ((48, 48, 48), (TWO_PI,) * 3),
            ((24, 32, 16), (TWO_PI, 4.0, 3.0)), ((64, 64, 64), (TWO_PI,) * 3), ((32, 48, 64), (TWO_PI,) * 3),
            ((128, 32, 16), (2 * TWO_PI, TWO_PI, 3.0))]
BOXES_32 = [((16, 16, 16), (TWO_PI,) * 3), ((32, 32, 32), (TWO_PI,) * 3), ((64, 64, 64), (TWO_PI,) * 3),
            ((32, 16, 64), (TWO_PI, 3.0, TWO_PI))]


@pytest.mark.parametrize("rule,boxes,mode", [("2/3-rule", BOXES_23, 0), ("2/3-rule", BOXES_23[:4], 1),
                                             ("3/2-rule", BOXES_32, 0)])
def test_batch_fp64_vs_oracle(rule, boxes, mode):
    nsteps = 3
    nus = [0.002 * (i + 1) for i in range(len(boxes))]
    dts = [0.004 + 0.001 * (i % 3) for i in range(len(boxes))]
    u0, u, diag = run_batch(boxes, "double", rule, nus, dts, nsteps, mode=mode)
    for i, (N, L) in enumerate(boxes):
        g = Grid(N, L, "double", rule, True, mode)
        ref = u0[i]
        for _ in range(nsteps):
            ref = g.rk4_step(ref, dts[i], nus[i])
        err = rel(u[i], ref)
        assert err < 1e-12, (i, N, err)
        assert abs(diag[2 * i] - g.kinetic_energy(ref)) <= 1e-12 * g.kinetic_energy(ref), i
        assert abs(diag[2 * i + 1] - g.dissipation(ref, nus[i])) <= 1e-11 * g.dissipation(ref, nus[i]), i


def test_batch_fp32_vs_oracle_and_single_box():
    from spectraldns_amd._abi import Plan
    boxes = [((32, 32, 32), (TWO_PI,) * 3), ((48, 48, 48), (TWO_PI,) * 3), ((64, 64, 64), (TWO_PI,) * 3),
             ((32, 32, 32), (TWO_PI,) * 3), ((48, 48, 48), (TWO_PI,) * 3), ((64, 64, 64), (TWO_PI,) * 3)]
    nsteps = 10
    nus = [1e-3, 2e-3, 5e-3, 1e-2, 3e-3, 7e-3]
    dts = [0.005] * len(boxes)
    u0, u, diag = run_batch(boxes, "single", "2/3-rule", nus, dts, nsteps)
    for i, (N, L) in enumerate(boxes):
        g = Grid(N, L, "double", "2/3-rule", True, 0)
        ref = u0[i]
        for _ in range(nsteps):
            ref = g.rk4_step(ref, dts[i], nus[i])
        assert rel(u[i], ref) < 1e-5, (i, rel(u[i], ref))
        assert abs(diag[2 * i] - g.kinetic_energy(ref)) <= 1e-5 * g.kinetic_energy(ref)
        p = Plan(N, L, "single", "2/3-rule", True, 0, tuned=False)
        us = torch.from_numpy(u0[i]).to(torch.complex64).cuda()
        p.rk4_steps(us, dts[:1] * nsteps, nus[i])
        torch.cuda.synchronize()
        assert rel(u[i], us.cpu().numpy()) < 1e-5, i


def test_batch_graph_equals_eager():
    boxes = [((32, 32, 32), (TWO_PI,) * 3), ((48, 48, 48), (TWO_PI,) * 3), ((64, 64, 64), (TWO_PI,) * 3)]
    nus, dts = [1e-3, 4e-3, 9e-3], [0.005] * 3
    _, ug, dg = run_batch(boxes, "single", "2/3-rule", nus, dts, 6, graph=True, split=[2, 1, 3])
    _, ue, de = run_batch(boxes, "single", "2/3-rule", nus, dts, 6, graph=False)
    for a, b in zip(ug, ue):
        assert np.array_equal(a, b)
    assert np.array_equal(dg, de)


def test_batch_params_and_state_update_after_capture():
    """New nu/dt and new state pointers after the graph exists take effect."""
    from spectraldns_amd._abi import Batch
    boxes = [((32, 32, 32), (TWO_PI,) * 3), ((48, 48, 48), (TWO_PI,) * 3)]
    u0 = [init(N, L, i) for i, (N, L) in enumerate(boxes)]
    b = Batch(boxes, "double", "2/3-rule")
    ua = [torch.from_numpy(a).cuda() for a in u0]
    b.set_state(ua)
    b.set_params([1e-3, 1e-3], [0.005, 0.005])
    b.rk4_steps(1)
    ub = [torch.from_numpy(a).cuda() for a in u0]
    b.set_state(ub)
    b.set_params([4e-3, 8e-3], [0.003, 0.002])
    b.rk4_steps(2)
    torch.cuda.synchronize()
    for i, (N, L) in enumerate(boxes):
        g = Grid(N, L, "double", "2/3-rule")
        ref = u0[i]
        for _ in range(2):
            ref = g.rk4_step(ref, [0.003, 0.002][i], [4e-3, 8e-3][i])
        assert rel(ub[i].cpu().numpy(), ref) < 1e-12
        ref1 = g.rk4_step(u0[i], 0.005, 1e-3)
        assert rel(ua[i].cpu().numpy(), ref1) < 1e-12


def test_batch_rejects_unsupported_lengths():
    from spectraldns_amd._abi import Batch, SdnsError
    with pytest.raises(SdnsError, match="kernel for line length 40"):  # no single-box kernel either
        Batch([((40, 40, 40), (TWO_PI,) * 3)], "single", "2/3-rule")
    with pytest.raises(SdnsError, match="batched kernel"):  # 192-point lines: single-box only
        Batch([((192, 32, 32), (TWO_PI,) * 3)], "single", "2/3-rule")
    with pytest.raises(SdnsError, match="batched kernel"):
        Batch([((48, 48, 48), (TWO_PI,) * 3)], "single", "3/2-rule")  # 72-point lines
    with pytest.raises(SdnsError):
        Batch([((32, 32, 32), (TWO_PI,) * 3)], "single", "None")


def test_default_ensemble_96_boxes():
    """The 96-box uneven ensemble (32 x 32^3, 32 x 48^3, 32 x 64^3): a few
    steps, energies finite and decaying, spot boxes against the oracle."""
    from spectraldns_amd.ensemble import Ensemble
    print("\n[case] default ensemble", flush=True)
    ens = Ensemble(precision="single")
    assert ens.n == 96 and ens.points() == 32 * (32 ** 3 + 48 ** 3 + 64 ** 3)
    diag = torch.zeros(2 * ens.n, dtype=torch.float64, device="cuda")
    nsteps = 4
    ens.steps(nsteps, diag)
    torch.cuda.synchronize()
    d = diag.cpu().numpy()
    assert np.all(np.isfinite(d))
    for i in (0, 33, 95):
        N = ens.sizes[i]
        g = Grid((N, N, N), (TWO_PI,) * 3, "double", "2/3-rule")
        ref = ens.u0[i]
        e0 = g.kinetic_energy(ref)
        for _ in range(nsteps):
            ref = g.rk4_step(ref, ens.dt[i], ens.nu[i])
        assert rel(ens.u_hat[i].cpu().numpy(), ref) < 1e-5, i
        assert d[2 * i] < e0
        assert abs(d[2 * i] - g.kinetic_energy(ref)) <= 1e-5 * e0


@pytest.mark.parametrize("rule,boxes", [("2/3-rule", BOXES_23), ("3/2-rule", BOXES_32)])
def test_batch_schedule_variants_fp64(rule, boxes):
    """Every (pass, length) schedule set to variant 1 (half the points per
    thread, see sdns_batch_set_variant): FP64 boxes within 1e-12 of the oracle
    after three steps and within 1e-13 of the default schedules, graph replay
    after the new layout included; an invalid variant is rejected."""
    from spectraldns_amd._abi import Batch, SdnsError
    nsteps = 3
    nus = [0.01 + 0.002 * i for i in range(len(boxes))]
    dts = [0.004] * len(boxes)
    u0, ref_gpu, _ = run_batch(boxes, "double", rule, nus, dts, nsteps)
    print("\n[case] batch variants %s" % rule, flush=True)
    u = [torch.from_numpy(a).to(torch.complex128).cuda().contiguous() for a in u0]
    b = Batch(boxes, "double", rule, True, 0)
    b.set_state(u)
    b.set_params(nus, dts)
    b.rk4_steps(1)                          # a captured step of the default layout first
    for t, a in zip(u, u0):
        t.copy_(torch.from_numpy(a))
    for k, lengths in b.line_lengths().items():
        for length in lengths:
            b.set_variant(k, length, 1)
    with pytest.raises(SdnsError):
        b.set_variant(5, 32, 1)
    b.rk4_steps(nsteps)
    torch.cuda.synchronize()
    for i, (N, L) in enumerate(boxes):
        g = Grid(N, L, "double", rule)
        ref = u0[i]
        for _ in range(nsteps):
            ref = g.rk4_step(ref, dts[i], nus[i])
        got = u[i].cpu().numpy()
        assert rel(got, ref) < 1e-12, (i, rel(got, ref))
        assert rel(got, ref_gpu[i]) < 1e-13, (i, rel(got, ref_gpu[i]))


@pytest.mark.timeout(300)
def test_default_ensemble_500_steps():
    """The ensemble benchmark configuration in full: 96 boxes (32 x 32^3,
    32 x 48^3, 32 x 64^3), RK4, 2/3 rule, 500 steps (T = 2.5), the stored
    schedules and graph replay of the bench.  Reference: the FP64 single-box
    GPU path over the same 500 steps (held to 1e-12 of the oracle in the other
    tests) and, for one 32^3 box, the oracle itself.
      * FP64 batch: every box within 1e-11 relative L2 (the batched passes are
        the single-box arithmetic, so only rounding differs);
      * FP32 batch: the Taylor-Green boxes within 3e-6 relative L2, every box
        within 5e-5, median below 5e-6, FP64 energy and dissipation within
        1e-5.  The random-field boxes (k0 = 4, nu down to 1e-3) amplify FP32
        rounding over the 2.5 time units (measured: Taylor-Green boxes
        <= 1.6e-6, random boxes up to 3.7e-5 at 64^3, median 2.9e-6), so the
        1e-5 bar of a single 256^3 run after a few steps is 5e-5 here after
        500; the FP64 batch shows the batched arithmetic itself is exact."""
    from spectraldns_amd._abi import Plan
    from spectraldns_amd.ensemble import Ensemble
    nsteps = 500
    refs, plans = {}, {}
    errs = {}
    for precision in ("double", "single"):
        print("\n[case] default ensemble, %s, 500 steps" % precision, flush=True)
        ens = Ensemble(precision=precision)
        diag = torch.zeros(2 * ens.n, dtype=torch.float64, device="cuda")
        ens.steps(nsteps, diag)
        torch.cuda.synchronize()
        d = diag.cpu().numpy()
        errs[precision] = []
        for i, N in enumerate(ens.sizes):
            if i not in refs:
                p = plans.setdefault(N, Plan((N, N, N), (TWO_PI,) * 3, "double", "2/3-rule"))
                ref = torch.from_numpy(ens.u0[i]).to(torch.complex128).cuda()
                dd = torch.zeros(2 * nsteps, dtype=torch.float64, device="cuda")
                p.rk4_steps(ref, [ens.dt[i]] * nsteps, ens.nu[i], diag=dd)
                refs[i] = (ref.cpu().numpy(), dd[-2:].cpu().numpy())
            ref, (e, eps) = refs[i]
            errs[precision].append(rel(ens.u_hat[i].cpu().numpy().astype(np.complex128), ref))
            tol = 1e-12 if precision == "double" else 1e-5
            assert abs(d[2 * i] - e) <= tol * e and abs(d[2 * i + 1] - eps) <= tol * eps, (precision, i)
        if precision == "double":
            last = ens.u_hat[1].cpu().numpy()
    e64, e32 = np.array(errs["double"]), np.array(errs["single"])
    print("FP64 batch worst %.3e; FP32 batch worst %.3e (box %d), median %.3e" %
          (e64.max(), e32.max(), int(e32.argmax()), float(np.median(e32))), flush=True)
    assert e64.max() < 1e-11
    tg = np.array([k == "tg" for k in ens.kind])
    assert e32[tg].max() < 3e-6
    assert e32.max() < 5e-5 and np.median(e32) < 5e-6
    g = Grid((32, 32, 32), (TWO_PI,) * 3, "double", "2/3-rule")
    ref = ens.u0[1]
    for _ in range(nsteps):
        ref = g.rk4_step(ref, ens.dt[1], ens.nu[1])
    assert rel(last, ref) < 1e-11
    assert rel(ens.u_hat[1].cpu().numpy().astype(np.complex128), ref) < 5e-5
