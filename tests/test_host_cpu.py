"""CPU tests of the host side: configuration semantics (spectralDNS/config.py),
the solve() loop hook order and dt sequence (spectralDNS/__init__.py:69-123,
NS.py:112-122), the C-ABI library exports, and the isotropic generator."""
import ctypes
import os
import re
import types

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_params_semantics():
    from spectraldns_amd.config import Params
    p = Params()
    p.precision = "single"
    p.M = [4, 5, 6]
    p.nu = 0.000625
    p.dt = 0.01
    p.L = ["2*pi", "4*pi", 3]
    assert list(p.N) == [16, 32, 64] and not p.M.flags.writeable
    assert isinstance(p.nu, np.float32) and isinstance(p["nu"], float)
    assert np.allclose(p.L, [2 * np.pi, 4 * np.pi, 3.0]) and not p.L.flags.writeable
    p.precision = "double"
    assert isinstance(p.dt, np.float64)
    p.N = [8, 8, 8]
    assert list(p.N) == [8, 8, 8]
    assert np.allclose(p.dx, p.L / p.N)
    with pytest.raises(ValueError):
        p.L = ["__import__('os')", 1, 1]


def test_update_and_parse():
    from spectraldns_amd import config
    config.update({"nu": 0.002, "dt": 0.005, "T": 0.2, "planner_effort": {"fft": "FFTW_ESTIMATE"}})
    args = config.triplyperiodic.parse_args(["--M", "5", "5", "5", "--dealias", "3/2-rule", "NS"])
    assert args.nu == 0.002 and args.dt == 0.005 and args.solver == "NS" and args.dealias == "3/2-rule"
    assert args.planner_effort["fft"] == "FFTW_ESTIMATE" and args.optimization == "hip"
    config.update({"nu": 0.000625, "dt": 0.01, "T": 0.1})


def test_solve_loop_hook_order_and_dt_sequence():
    """solve() with a stub solver module: the hooks run in the reference's order
    and the dt sequence equals the oracle's restatement (last step adjusted)."""
    import spectraldns_amd as sd
    from spectraldns_amd import config
    from spectraldns_amd.solvers import spectralinit
    from oracle.ns import dt_schedule
    for prec in ("double", "single"):
        config.params.update(dict(precision=prec, dt=0.01, T=0.1, t=0.0, tstep=0, make_profile=0,
                                  convection="Vortex", verbose=False))
        calls, dts = [], []
        mod = types.SimpleNamespace()
        mod.params = config.params
        mod.Timer = lambda: types.SimpleNamespace(__call__=None, final=lambda v: calls.append("final"))

        class T:
            def __call__(self):
                calls.append("timer")

            def final(self, v):
                calls.append("final")
        mod.Timer = T
        mod.profiler = types.SimpleNamespace(getstats=lambda: [], enable=lambda: None)
        mod.getConvection = lambda c: "conv"

        def getintegrator(rhs, u, solver, context):
            def f():
                dts.append(config.params.dt)
                calls.append("integrate")
                return u, config.params.dt, config.params.dt
            return f
        mod.getintegrator = getintegrator
        mod.update = lambda c: calls.append("update")
        mod.end_of_tstep = lambda c: (calls.append("end"), NSend(c))[1]
        mod.regression_test = lambda c: calls.append("regression")
        from spectraldns_amd.solvers import NS as _unused  # noqa: F401  (import needs no GPU)
        NSend = __import__("spectraldns_amd.solvers.NS", fromlist=["end_of_tstep"]).end_of_tstep
        hfile = types.SimpleNamespace(update=lambda p, **k: calls.append("hdf5"),
                                      close=lambda: calls.append("close"))
        ctx = config.AttributeDict(dU=None, u=None, hdf5file=hfile)
        sd.solve(mod, ctx)
        assert calls[:5] == ["integrate", "update", "hdf5", "timer", "end"]
        assert calls[-3:] == ["final", "regression", "close"]
        ref = dt_schedule(0.1, 0.01, precision=prec)
        assert len(dts) == len(ref) and all(a == b for a, b in zip(dts, ref))
        assert config.params.dt == 0.01 or abs(config.params.dt - 0.01) < 1e-9


def test_library_exports_every_declared_symbol():
    from spectraldns_amd import _abi
    if not os.path.exists(_abi.LIB_PATH):
        pytest.skip("libsdns_hip.so not built")
    hdr = open(os.path.join(ROOT, "include", "sdns.h")).read()
    names = set(re.findall(r"\b(sdns_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) >= 18
    lib = ctypes.CDLL(_abi.LIB_PATH)
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert names == set(_abi._SIGS), names ^ set(_abi._SIGS)


def test_isotropic_generator_properties():
    from spectraldns_amd.isotropic import random_solenoidal_field, hermitian_weights
    from oracle.ns import Grid
    N = (32, 32, 32)
    u = random_solenoidal_field(N, seed=7, k0=4.0, energy=0.5)
    g = Grid(N)
    assert abs(g.kinetic_energy(u) - 0.5) < 1e-14
    div = g.K[0] * u[0] + g.K[1] * u[1] + g.K[2] * u[2]
    assert np.max(abs(div)) < 1e-14
    assert np.all(u[:, 16] == 0) and np.all(u[:, :, 16] == 0) and np.all(u[..., 16] == 0)
    assert np.all(u[:, 0, 0, 0] == 0)
    # Hermitian on kz = 0: the inverse transform is real
    assert np.max(abs(np.fft.ifft2(u[:, :, :, 0]).imag)) < 1e-15
    # same seed, same field; spectrum peaks near k0 = 4 (E ~ k^4 exp(-2 (k/4)^2) peaks at k = 4)
    assert np.array_equal(u, random_solenoidal_field(N, seed=7, k0=4.0, energy=0.5))
    Ek, bins = g.spectrum(u)
    assert 3 <= int(np.argmax(Ek)) + 1 <= 5
    assert list(hermitian_weights(8)) == [1, 2, 2, 2, 1]


def test_closure_symmetry_images_match_analytic_fields():
    """symmetry(): u'_i(x') = s_i u_{p(i)}(x), tau'_ij(x') = s_i s_j tau_{p(i)p(j)}(x)
    with x'_i = s_i x_{p(i)} on the periodic grid, checked against analytic fields."""
    import itertools
    import torch
    from spectraldns_amd.closure import symmetry, _PAIRS
    N = 8
    xs = 2 * np.pi * np.arange(N) / N

    def u_of(x):  # x: (3, ...) coordinates
        return np.stack([np.sin(x[1]) * np.cos(2 * x[2]), 2 * np.cos(x[0]) + np.sin(x[2]),
                         np.sin(x[0] + 2 * x[1])])

    X = np.stack(np.meshgrid(xs, xs, xs, indexing="ij"))
    u = u_of(X)
    tau = np.stack([u[i] * u[j] + (i == j) for i, j in _PAIRS])
    ut, tt = torch.from_numpy(u), torch.from_numpy(tau)
    for perm in itertools.permutations(range(3)):
        for signs in ((1, 1, 1), (-1, 1, 1), (1, -1, -1), (-1, -1, 1)):
            u2, t2 = symmetry(ut, tt, perm, signs)
            x = np.empty_like(X)
            for i in range(3):
                x[perm[i]] = signs[i] * X[i]  # old coordinates of the new grid point X
            uo = u_of(x)
            ue = np.stack([signs[i] * uo[perm[i]] for i in range(3)])
            te = np.stack([signs[i] * signs[j] * (uo[perm[i]] * uo[perm[j]] + (perm[i] == perm[j]))
                           for i, j in _PAIRS])
            assert np.allclose(u2.numpy(), ue, atol=1e-12), (perm, signs)
            assert np.allclose(t2.numpy(), te, atol=1e-12), (perm, signs)


@pytest.mark.parametrize("name,conv", [("NS", "Vortex"), ("VV", "Vortex"), ("MHD", "Divergence")])
def test_solver_modules_keep_the_reference_protocol(name, conv):
    """get_solver imports spectraldns_amd.solvers.<name> with the reference's
    module protocol (spectralDNS/__init__.py:19-66, solvers/NS.py, VV.py,
    MHD.py); the convection forms a module does not have raise."""
    from spectraldns_amd import config, get_solver
    solver = get_solver(parse_args=["--M", "4", "4", "4", name])
    assert config.params.solver == name and config.solver is solver
    for fn in ("get_context", "getConvection", "ComputeRHS", "end_of_tstep", "getintegrator", "update",
               "regression_test", "additional_callback", "get_velocity", "get_curl", "get_divergence"):
        if name == "MHD" and fn in ("get_velocity", "get_curl"):
            continue
        assert callable(getattr(solver, fn)), fn
    assert callable(solver.getConvection(conv))
    bad = {"NS": "Magnetic", "VV": "Standard", "MHD": "Vortex"}[name]
    with pytest.raises(NotImplementedError):
        solver.getConvection(bad)
    if name == "MHD":
        assert config.params.eta == np.float64(0.01)


def test_unbuilt_solver_and_integrator_raise():
    from spectraldns_amd import get_solver
    from spectraldns_amd.maths.integrators import getintegrator
    with pytest.raises(NotImplementedError):
        get_solver(parse_args=["--M", "4", "4", "4", "Bq"])
    from spectraldns_amd.config import AttributeDict
    solver = types.SimpleNamespace(params=AttributeDict(integrator="Heun"))
    with pytest.raises(NotImplementedError):
        getintegrator(None, np.zeros(1), solver, AttributeDict(float=np.float64))


def test_bs5_tableau_in_the_package_matches_the_oracle():
    from oracle import integrators as OI
    from spectraldns_amd.maths import integrators as PI
    assert np.array_equal(np.array(PI._BS5_A), OI.BS5_A)
    assert np.array_equal(np.array(PI._BS5_B), OI.BS5_B)
    assert np.array_equal(np.array(PI._BS5_BHAT), OI.BS5_BHAT)


def _file_params(tstep, t=0.5, checkpoint=10, write_result=4):
    from spectraldns_amd.config import AttributeDict
    return AttributeDict(tstep=tstep, t=t, checkpoint=checkpoint, write_result=write_result)


def test_kill_file_checkpoints_and_exits(tmp_path, monkeypatch):
    """HDF5File.update with 'killspectraldns' in the working directory writes
    the checkpoint at once (whatever params.checkpoint says), removes the kill
    file and exits with status 1 (h5io/HDF5File.py:82-92, 107-120)."""
    from spectraldns_amd.h5io import HDF5File, read_checkpoint
    monkeypatch.chdir(tmp_path)
    u = np.arange(24, dtype=np.complex128).reshape(3, 2, 2, 2) * (1 - 0.5j)
    f = HDF5File(str(tmp_path / "NS"), checkpoint={"data": {"0": {"U": [u]}}})
    f.update(_file_params(tstep=3))               # neither a checkpoint step nor a kill file
    assert not os.path.exists(tmp_path / "NS_c.npz")
    open("killspectraldns", "w").close()
    with pytest.raises(SystemExit) as e:
        f.update(_file_params(tstep=7, t=0.07))
    assert e.value.code == 1
    assert not os.path.exists("killspectraldns")
    c = read_checkpoint(str(tmp_path / "NS"))
    assert c["tstep"] == 7 and c["t"] == 0.07
    assert np.array_equal(c["U/3D/0"], u)
    assert sorted(os.listdir(tmp_path)) == ["NS_c.npz"]   # no temporary file left behind


def test_checkpoint_is_replaced_atomically(tmp_path, monkeypatch):
    """A checkpoint write that fails half way leaves the previous checkpoint
    readable and no temporary file."""
    from spectraldns_amd import h5io
    monkeypatch.chdir(tmp_path)
    u = np.ones((3, 2, 2, 2), dtype=np.complex128)
    f = h5io.HDF5File(str(tmp_path / "NS"), checkpoint={"data": {"0": {"U": [u]}}})
    f.update(_file_params(tstep=10, t=0.1))
    real_savez = np.savez

    def broken_savez(fobj, **data):
        fobj.write(b"PK\x03\x04 partial")
        raise OSError("disk full")
    monkeypatch.setattr(h5io.np, "savez", broken_savez)
    with pytest.raises(OSError):
        f.update(_file_params(tstep=20, t=0.2))
    monkeypatch.setattr(h5io.np, "savez", real_savez)
    c = h5io.read_checkpoint(str(tmp_path / "NS"))
    assert c["tstep"] == 10 and np.array_equal(c["U/3D/0"], u)
    assert sorted(os.listdir(tmp_path)) == ["NS_c.npz"]


def test_results_written_once_per_step(tmp_path, monkeypatch):
    """Results go to one archive per written step; nothing accumulates in the
    writer between writes, and read_results merges them under the reference's
    dataset names."""
    from spectraldns_amd.h5io import HDF5File, read_results
    monkeypatch.chdir(tmp_path)
    U = np.zeros((3, 4, 4, 4))
    f = HDF5File(str(tmp_path / "NS"), results={"data": {"U": [U]}})
    for step in range(1, 9):
        U[:] = step
        f.update(_file_params(tstep=step))
    assert not hasattr(f, "_w")
    r = read_results(str(tmp_path / "NS"))
    assert sorted(r) == ["U/3D/4", "U/3D/8"]
    assert np.all(r["U/3D/4"] == 4) and np.all(r["U/3D/8"] == 8)
