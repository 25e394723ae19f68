"""The reference's host API on the GPU: config.update, get_solver,
get_context, set_velocity / get_velocity / get_curl / get_pressure, the RK4
integrator hook and solve(), driven exactly like tests/test_NSVV.py with the
Taylor-Green regression hook of tests/TG.py:116-126."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def tg_initialize(solver, context):
    from spectraldns_amd import config
    U, X = context.U, context.X
    U[0] = torch.sin(X[0]) * torch.cos(X[1]) * torch.cos(X[2])
    U[1] = -torch.cos(X[0]) * torch.sin(X[1]) * torch.cos(X[2])
    U[2] = 0
    solver.set_velocity(**context)
    config.params.t = 0.0
    config.params.tstep = 0


def tg_regression(context):
    from spectraldns_amd import config
    params = config.params
    solver = config.solver
    U = solver.get_velocity(**context)
    curl = solver.get_curl(**context)
    w = float(torch.sum(curl.double() * curl.double()).item()) / np.prod(params.N) / 2
    k = float(torch.sum(U.double() * U.double()).item()) / np.prod(params.N) / 2
    context.regression = (k, w)
    assert round(w - 0.375249930801, params.ntol) == 0, w
    assert round(k - 0.124953117517, params.ntol) == 0, k


@pytest.mark.parametrize("args", [["--M", "4", "4", "4", "--L", "2*pi", "2*pi", "2*pi"],
                                  ["--M", "5", "5", "5", "--L", "2*pi", "2*pi", "2*pi"],
                                  ["--M", "6", "5", "4", "--L", "6*pi", "4*pi", "2*pi"]])
@pytest.mark.parametrize("dealias", ["2/3-rule", "3/2-rule"])
def test_taylor_green_solve(args, dealias):
    from spectraldns_amd import config, get_solver, solve
    print("\n[case] TG %s %s" % (" ".join(args), dealias), flush=True)
    config.update({"nu": 0.000625, "dt": 0.01, "T": 0.1, "convection": "Vortex", "dealias": dealias,
                   "precision": "double"})
    solver = get_solver(regression_test=tg_regression, parse_args=args + ["NS"])
    context = solver.get_context()
    tg_initialize(solver, context)
    solve(solver, context)
    assert config.params.tstep == 10


def test_taylor_green_single_precision():
    from spectraldns_amd import config, get_solver, solve
    config.update({"nu": 0.000625, "dt": 0.01, "T": 0.1, "dealias": "2/3-rule", "precision": "single"})
    solver = get_solver(regression_test=tg_regression,
                        parse_args=["--M", "5", "5", "5", "--ntol", "5", "NS"])
    context = solver.get_context()
    assert context.U_hat.dtype == torch.complex64
    tg_initialize(solver, context)
    solve(solver, context)
    config.update({"precision": "double", "ntol": 7})


def test_context_keys_and_accessors():
    from spectraldns_amd import config, get_solver
    from oracle.ns import Grid
    config.update({"nu": 0.01, "dt": 0.01, "T": 0.02, "dealias": "2/3-rule", "precision": "double"})
    solver = get_solver(parse_args=["--M", "4", "4", "4", "NS"])
    c = solver.get_context()
    for key in ("T", "VT", "Tp", "VTp", "mask", "X", "K", "U", "U_hat", "P", "P_hat", "u", "dU", "curl",
                "Source", "work", "hdf5file", "float", "complex"):
        assert key in c
    assert c.K2.shape == (16, 16, 9) and c.K_over_K2.shape == (3, 16, 16, 9)
    tg_initialize(solver, c)
    g = Grid((16, 16, 16))
    assert np.max(abs(c.U_hat.cpu().numpy() - g.taylor_green())) < 1e-15
    solver.conv = solver.getConvection("Vortex")
    rhs = solver.ComputeRHS(c.dU, c.U_hat, solver, **c)
    assert np.max(abs(rhs.cpu().numpy() - g.compute_rhs(g.taylor_green(), 0.01))) < 1e-14
    P = solver.get_pressure(**c)
    _, P_ref = g.add_pressure_diffusion(g.mask_nyquist(g.conv(g.taylor_green())), g.taylor_green(), 0.01)
    assert np.max(abs(P.cpu().numpy() - g.backward(-1j * P_ref))) < 1e-13
    div = solver.get_divergence(**c)
    assert float(div.abs().max()) < 1e-14


def test_compute_rhs_granular_path():
    """ComputeRHS with a user-supplied conv (not the fused Vortex one) runs
    conv -> Nyquist mask -> add_pressure_diffusion -> Source step by step
    (NS.py:219-261) and agrees with the oracle."""
    from spectraldns_amd import config, get_solver
    from spectraldns_amd.isotropic import random_solenoidal_field
    from oracle.ns import Grid
    config.update({"nu": 0.01, "dt": 0.01, "T": 0.02, "dealias": "2/3-rule", "precision": "double"})
    solver = get_solver(parse_args=["--M", "4", "4", "4", "NS"])
    c = solver.get_context()
    u = random_solenoidal_field((16, 16, 16), seed=5, k0=3.0, energy=0.5)
    c.U_hat.copy_(torch.from_numpy(u))
    base = solver.getConvection("Vortex")

    def conv(rhs, u_hat, work, Tp, VTp, K, dummy):
        return base(rhs, u_hat, work, Tp, VTp, K, dummy)

    solver.conv = conv
    rhs = solver.ComputeRHS(c.dU, c.U_hat, solver, **c)
    g = Grid((16, 16, 16))
    ref = g.compute_rhs(u, 0.01)
    assert np.max(abs(rhs.cpu().numpy() - ref)) < 1e-13 * np.max(abs(ref))


def test_checkpoint_restart_continues_bitwise(tmp_path, monkeypatch):
    """A run checkpointing every 5 steps (HDF5File.update, HDF5File.py:57-120,
    with the reference's keys U/3D/0, t, tstep), restarted from its step-5
    checkpoint, reaches t = 0.1 with the same U_hat, bit for bit, as the run
    itself."""
    import shutil
    from spectraldns_amd import config, get_solver, solve
    from spectraldns_amd.h5io import read_checkpoint
    monkeypatch.chdir(tmp_path)

    def keep_step5(context):  # update runs before the checkpoint of its step
        if config.params.tstep == 6:
            shutil.copy("NS_c.npz", "step5_c.npz")

    config.update({"nu": 0.000625, "dt": 0.01, "T": 0.1, "dealias": "2/3-rule", "precision": "double",
                   "checkpoint": 5, "write_result": 100000})
    solver = get_solver(update=keep_step5, regression_test=lambda c: None,
                        parse_args=["--M", "4", "4", "4", "NS"])
    c = solver.get_context()
    tg_initialize(solver, c)
    solve(solver, c)
    straight = c.U_hat.clone()
    f10 = read_checkpoint("NS")
    assert f10["tstep"] == 10 and abs(f10["t"] - 0.1) < 1e-12
    assert np.array_equal(f10["U/3D/0"], straight.cpu().numpy())
    f = read_checkpoint("step5")
    assert f["tstep"] == 5
    solver.update = lambda c: None
    c2 = solver.get_context()
    c2.U_hat.copy_(torch.from_numpy(f["U/3D/0"]))
    config.params.t, config.params.tstep = f["t"], f["tstep"]
    solve(solver, c2)
    assert torch.equal(c2.U_hat, straight)
    config.update({"checkpoint": 100000, "write_result": 100000000})
