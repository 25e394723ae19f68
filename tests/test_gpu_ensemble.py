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
