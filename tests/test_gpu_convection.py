"""The Standard, Divergence and Skewed convection forms of the NS solver
(spectralDNS/solvers/NS.py:138-199) on the GPU, through the reference's API
(config convection, get_solver, solve; Taylor-Green pins at ntol 7 as in
tests/TG.py:116-126) on the uniform and the non-uniform box, and against the
oracle's restatement (oracle/convection.py) on the same inputs: FP64 within
1e-12 relative L2 after the ten RK4 steps; the same with the 3/2 rule
(padded 3N/2 transforms, sdns_backward_dealiased / sdns_forward_dealiased),
and those padded transforms against the oracle's."""
import numpy as np
import pytest

from oracle import convection as OCV
from oracle.ns import Grid, dt_schedule

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
TWO_PI = 2 * np.pi
CASES = [["--M", "4", "4", "4", "--L", "2*pi", "2*pi", "2*pi"],
         ["--M", "6", "5", "4", "--L", "6*pi", "4*pi", "2*pi"]]


def tg_initialize(solver, context):
    from spectraldns_amd import config
    U, X = context.U, context.X
    U[0] = torch.sin(X[0]) * torch.cos(X[1]) * torch.cos(X[2])
    U[1] = -torch.cos(X[0]) * torch.sin(X[1]) * torch.cos(X[2])
    U[2] = 0
    solver.set_velocity(**context)
    config.params.t = 0.0
    config.params.tstep = 0


def regression(context):
    from spectraldns_amd import config
    solver = config.solver
    U = solver.get_velocity(**context)
    curl = solver.get_curl(**context)
    n = np.prod(config.params.N)
    w = float(torch.sum(curl.double() * curl.double()).item()) / n / 2
    k = float(torch.sum(U.double() * U.double()).item()) / n / 2
    assert round(w - 0.375249930801, 7) == 0, w
    assert round(k - 0.124953117517, 7) == 0, k


@pytest.mark.parametrize("form", ["Standard", "Divergence", "Skewed"])
@pytest.mark.parametrize("args", CASES)
def test_convection_forms_fp64(args, form):
    from spectraldns_amd import config, get_solver, solve
    print("\n[case] NS %s %s" % (form, " ".join(args)), flush=True)
    config.update({"nu": 0.000625, "dt": 0.01, "T": 0.1, "convection": form, "dealias": "2/3-rule",
                   "precision": "double"})
    solver = get_solver(regression_test=regression, parse_args=args + ["NS"])
    c = solver.get_context()
    tg_initialize(solver, c)
    solve(solver, c)
    u = c.U_hat.cpu().numpy()
    config.update({"convection": "Vortex"})
    M = [int(a) for a in args[1:4]]
    L = (TWO_PI * 3, TWO_PI * 2, TWO_PI) if M[0] == 6 else (TWO_PI,) * 3
    g = Grid(tuple(2 ** m for m in M), L, "double", "2/3-rule")
    ref = g.taylor_green()
    for dt in dt_schedule(0.1, 0.01):
        ref = OCV.rk4_step(g, ref, dt, 0.000625, form)
    err = float(np.linalg.norm((u - ref).ravel()) / np.linalg.norm(ref.ravel()))
    assert err < 1e-12, err


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("N", [(16, 16, 16), (32, 16, 64), (64, 32, 16)])
def test_padded_transforms_vs_oracle(N, precision):
    """3/2-rule VTp.backward onto the 3N/2 grid and Tp.forward back to N modes
    (FP64 1e-14, FP32 1e-6 relative L2), on random data."""
    from spectraldns_amd._abi import Plan
    print("\n[case] padded transforms %s %s" % (N, precision), flush=True)
    L = (TWO_PI, 3.0, 5.0)
    g = Grid(N, L, "double", "3/2-rule")
    p = Plan(N, L, precision, "3/2-rule")
    assert p.dealiased_shape == tuple(3 * n // 2 for n in N)
    cd = torch.complex128 if precision == "double" else torch.complex64
    rd = torch.float64 if precision == "double" else torch.float32
    tol = 1e-14 if precision == "double" else 1e-6
    rng = np.random.default_rng(11)
    shp = (3,) + g.K2.shape
    a = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    a = g.forward(g.backward(a))          # a Hermitian-consistent spectrum
    phys = torch.empty((3,) + p.dealiased_shape, dtype=rd, device="cuda")
    p.backward_dealiased(torch.from_numpy(a).to(cd).cuda(), phys)
    ref = g.backward(a, dealiased=True)
    err = float(np.linalg.norm((phys.cpu().numpy() - ref).ravel()) / np.linalg.norm(ref.ravel()))
    assert err < tol, err
    x = rng.standard_normal((3,) + p.dealiased_shape)
    out = torch.empty(shp, dtype=cd, device="cuda")
    p.forward_dealiased(torch.from_numpy(x).to(rd).cuda(), out)
    ref = g.forward(x, dealiased=True)
    err = float(np.linalg.norm((out.cpu().numpy() - ref).ravel()) / np.linalg.norm(ref.ravel()))
    assert err < tol, err


@pytest.mark.parametrize("form", ["Standard", "Divergence", "Skewed"])
def test_convection_forms_32_rule_fp64(form):
    """The forms with the 3/2 rule through the API: Taylor-Green pins at ntol 7
    and the oracle (3/2 rule) within 1e-12 after ten steps."""
    from spectraldns_amd import config, get_solver, solve
    print("\n[case] NS %s 16^3 3/2-rule" % form, flush=True)
    config.update({"nu": 0.000625, "dt": 0.01, "T": 0.1, "convection": form, "dealias": "3/2-rule",
                   "precision": "double"})
    solver = get_solver(regression_test=regression, parse_args=["--M", "4", "4", "4", "NS"])
    c = solver.get_context()
    assert tuple(c.u_dealias.shape) == (3, 24, 24, 24)
    tg_initialize(solver, c)
    solve(solver, c)
    u = c.U_hat.cpu().numpy()
    config.update({"convection": "Vortex", "dealias": "2/3-rule"})
    g = Grid((16, 16, 16), (TWO_PI,) * 3, "double", "3/2-rule")
    ref = g.taylor_green()
    for dt in dt_schedule(0.1, 0.01):
        ref = OCV.rk4_step(g, ref, dt, 0.000625, form)
    err = float(np.linalg.norm((u - ref).ravel()) / np.linalg.norm(ref.ravel()))
    assert err < 1e-12, err


@pytest.mark.parametrize("form", ["Standard", "Skewed"])
def test_convection_forms_fp32(form):
    """Single precision against the FP64 oracle on the 32^3 box (1e-5)."""
    from spectraldns_amd import config, get_solver, solve
    print("\n[case] NS %s 32^3 single" % form, flush=True)
    config.update({"nu": 0.000625, "dt": 0.01, "T": 0.1, "convection": form, "dealias": "2/3-rule",
                   "precision": "single"})
    solver = get_solver(update=lambda c: None, regression_test=lambda c: None,
                        parse_args=["--M", "5", "5", "5", "NS"])
    c = solver.get_context()
    tg_initialize(solver, c)
    solve(solver, c)
    u = c.U_hat.cpu().numpy().astype(np.complex128)
    config.update({"convection": "Vortex", "precision": "double"})
    g = Grid((32, 32, 32), (TWO_PI,) * 3, "double", "2/3-rule")
    ref = g.taylor_green()
    for dt in dt_schedule(0.1, 0.01, precision="single"):
        ref = OCV.rk4_step(g, ref, dt, 0.000625, form)
    err = float(np.linalg.norm((u - ref).ravel()) / np.linalg.norm(ref.ravel()))
    assert err < 1e-5, err
