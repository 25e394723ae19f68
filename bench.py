"""Throughput of the RK4 step on one MI355X.

Workload (one "step" = one RK4 step of the whole box): decaying isotropic
turbulence at 256^3, 3/2-rule dealiasing (384^3 padded), nu = 0.001,
dt = 0.001, FP32 with FP64 energy/dissipation reductions; the initial field
is the seeded random-phase solenoidal field with E(k) ~ k^4 exp(-2 (k/4)^2),
energy 0.5 (synthetic data, generated on the host, resident in HBM before the
timed region).  The same workload is also timed in FP64.

Prints ONE JSON line with steps/s, the per-kernel HBM roofline of the
dominant pass (algorithmic bytes per launch / its average HIP-event duration
on the plan's stream, recorded around that pass's launches inside the timed
region), and a CPU baseline: the NumPy/SciPy oracle timed on this host on one
RK4 step of the same workload.  The per-pass table comes from a short run
with every pass timed, before the timed region.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "RK4 steps/s at 256^3 FP32/FP64, % HBM roofline; closure vs DNS speedup @err"


def pass_bytes(N, dealias, esz):
    """Algorithmic bytes per launch of each pass (canonical fused graph).

    F = N0*N1*(N2/2+1) complex elements * esz bytes.  With M = 3N/2 padding
    (3/2 rule) fields that went through the x (y) inverse are M0/N0 (M1/N1)
    times larger.  P1 reads u_hat (3F) and writes 5 fields; P2 reads 5 and
    writes 6; P3 reads 6 and writes 3; P4 reads 3 and writes 3; P5 reads 3 and
    streams the RK4 state: stage 0 reads u (3F) and writes u0, u2 (6F);
    stages 1-2 read u1, u0, u2 (9F) and write u0, u2 (6F); stage 3 reads u0,
    u2 (6F) and writes u (3F)."""
    F = N[0] * N[1] * (N[2] // 2 + 1) * esz
    rx = 1.5 if dealias == "3/2-rule" else 1.0
    rxy = rx * rx
    p = {
        "x_inverse": 3 * F + 5 * F * rx,
        "y_inverse_curl": 5 * F * rx + 6 * F * rxy,
        "z_cross": 6 * F * rxy + 3 * F * rxy,
        "y_forward": 3 * F * rxy + 3 * F * rx,
    }
    rk = [9 * F, 15 * F, 15 * F, 9 * F]
    p["x_forward_rk"] = 3 * F * rx + sum(rk) / 4.0
    step = 4 * (p["x_inverse"] + p["y_inverse_curl"] + p["z_cross"] + p["y_forward"]) + 4 * 3 * F * rx + sum(rk)
    return p, step, F


def gpu_run(N, precision, dealias, steps, warmup, nu=0.001, dt=0.001, seed=1234):
    import torch
    from spectraldns_amd._abi import Plan
    from spectraldns_amd.isotropic import random_solenoidal_field
    L = (2 * np.pi,) * 3
    plan = Plan(N, L, precision, dealias)
    cd = torch.complex64 if precision == "single" else torch.complex128
    u = torch.from_numpy(random_solenoidal_field(N, L, seed=seed, k0=4.0, energy=0.5)).to(cd).cuda()
    diag = torch.zeros(2 * max(steps, warmup, 1), dtype=torch.float64, device="cuda")
    plan.rk4_steps(u, [dt] * warmup, nu, diag=diag)
    torch.cuda.synchronize()
    # per-pass breakdown from a short run with every pass timed; it also
    # names the dominant pass
    plan.timing_read()
    plan.timing_enable(True)
    plan.rk4_steps(u, [dt] * min(steps, 10), nu, diag=diag)
    torch.cuda.synchronize()
    plan.timing_enable(False)
    passes = plan.timing_read()
    dom = max((k for k in passes if k != "other"), key=lambda k: passes[k][0])
    # timed region: HIP events only around the dominant pass's launches, so the
    # other launches stay back to back
    plan.timing_passes([dom])
    plan.timing_enable(True)
    t0 = time.perf_counter()
    plan.rk4_steps(u, [dt] * steps, nu, diag=diag)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    plan.timing_enable(False)
    passes[dom] = plan.timing_read()[dom]
    plan.timing_passes(None)
    e = diag[: 2 * steps].cpu().numpy().reshape(-1, 2)
    variants = (plan.schedule or {}).get("variants", {})
    return elapsed, passes, e, variants, dom


def ensemble_run(steps, warmup, graph=True):
    """The 96-box uneven ensemble (32 x 32^3, 32 x 48^3, 32 x 64^3, FP32,
    2/3 rule): seconds for `steps` RK4 steps of every box, and the FP64
    per-box (energy, dissipation) after them."""
    import torch
    from spectraldns_amd.ensemble import Ensemble
    ens = Ensemble(precision="single")
    ens.batch.use_graph(graph)
    diag = torch.zeros(2 * ens.n, dtype=torch.float64, device="cuda")
    ens.steps(warmup, diag)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.steps(steps, diag)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, ens, diag.cpu().numpy().reshape(-1, 2)


def ensemble_main(args):
    el, ens, diag = ensemble_run(args.steps, args.warmup, graph=not args.no_graph)
    sps = args.steps / el
    b_step = sum(pass_bytes((n, n, n), "2/3-rule", 8)[1] for n in ens.sizes)
    out = {
        "metric": "ensemble RK4 steps/s (every box advanced one step)",
        "value": round(sps, 3), "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1000.0 * el / args.steps, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (Taylor-Green and seeded random-phase fields, nu log-uniform in [1e-3, 1e-2])",
        "config": {"workload": "ensemble of 96 boxes: 32 x 32^3, 32 x 48^3, 32 x 64^3, 2/3 rule, RK4, FP32, "
                               "dt=0.005, %s" % ("hipGraph replay" if not args.no_graph else "eager launches"),
                   "boxes": ens.n, "points": ens.points()},
        "box_steps_per_s": round(sps * ens.n, 1),
        "step_roofline": {"algorithmic_bytes_per_step": b_step, "achieved_GBps": round(b_step * sps / 1e9, 1),
                          "frac": round(b_step * sps / 1e9 / HBM_PEAK_GBS, 4)},
        "energy_range": [float(diag[:, 0].min()), float(diag[:, 0].max())],
    }
    print(json.dumps(out), flush=True)


def forced_main(args):
    """Forced isotropic turbulence, compensated FP32: one step = one RK4 step
    (u_hat carried as the pair u_hat + u_lo) followed by the large-scale
    forcing of demo/Isotropic.py:154-184 at |k| <= 2."""
    import torch
    from spectraldns_amd._abi import Plan
    from spectraldns_amd.isotropic import random_solenoidal_field_gpu
    N = (args.N,) * 3
    nu, dt, kf = 0.0002, 0.0005, 2.0
    plan = Plan(N, (2 * np.pi,) * 3, "single", "2/3-rule")
    t0 = time.perf_counter()
    u = random_solenoidal_field_gpu(plan, seed=1234, k0=4.0, energy=0.5)
    ulo = torch.zeros_like(u)
    e = torch.zeros(2, dtype=torch.float64, device="cuda")
    plan.energy(u, nu, e)
    target = 2.0 * float(e[0].item())
    torch.cuda.synchronize()
    t_init = time.perf_counter() - t0
    diag = torch.zeros(2 * max(args.steps, args.warmup, 1), dtype=torch.float64, device="cuda")

    def run(n):
        for s in range(n):
            plan.rk4_steps(u, [dt], nu, diag=diag[2 * s:], u_lo=ulo)
            plan.force(u, target, kf, u_lo=ulo)

    run(args.warmup)
    torch.cuda.synchronize()
    plan.timing_read()
    plan.timing_enable(True)
    t0 = time.perf_counter()
    run(args.steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    plan.timing_enable(False)
    passes = plan.timing_read()
    F = N[0] * N[1] * (N[2] // 2 + 1) * 8
    b_step = 214 * F  # 2/3 rule with compensated u_hat and u2 (hi, lo)
    sps = args.steps / el
    d = diag[: 2 * args.steps].cpu().numpy().reshape(-1, 2)
    out = {
        "metric": "forced isotropic %d^3 RK4 steps/s (FP32 compensated, FP64 reductions)" % args.N,
        "value": round(sps, 4), "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1000 * el / args.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded random-phase solenoidal field generated on the GPU, E(k)~k^4 exp(-2(k/4)^2))",
        "config": {"workload": "forced isotropic turbulence %d^3, 2/3 rule, RK4, nu=0.0002, dt=0.0005, forcing "
                               "|k|<=2, compensated u_hat" % args.N, "N": args.N},
        "step_roofline": {"algorithmic_bytes_per_step": b_step, "achieved_GBps": round(b_step * sps / 1e9, 1),
                          "frac": round(b_step * sps / 1e9 / HBM_PEAK_GBS, 4)},
        "passes": {k: {"ms_per_launch": round(v[0] / max(v[1], 1), 3), "launches": v[1]} for k, v in passes.items()},
        "energy_first_last": [float(d[0][0]), float(d[-1][0])], "dissipation_last": float(d[-1][1]),
        "init_s": round(t_init, 2),
        "memory_GB": round(torch.cuda.max_memory_allocated() / 1e9, 1),
    }
    print(json.dumps(out), flush=True)


def measured_traffic(N, dealias, dom, variants):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3
    PMC summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated), when that summary
    covers this workload, pass and schedule variant; otherwise None."""
    rule = {"3/2-rule": "32rule", "2/3-rule": "23rule"}[dealias]
    path = os.path.join(ROOT, "profiles", "traffic_%dcube_%s_fp32.json" % (N[0], rule))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    if t.get("pass") != dom or int(variants.get(dom, 0)) != int(t.get("variant", -1)):
        return None
    return {"bytes_per_launch": t["hbm_bytes_per_launch"], "kernel": t["kernel"],
            "source": "profiles/" + t["source"]}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(N, precision, dealias, nu=0.001, dt=0.001, warmup=1, timed=5):
    """The NumPy/SciPy oracle on this host's cores: `warmup` untimed RK4 steps,
    then the median wall time of `timed` RK4 steps of the same workload."""
    from oracle.ns import Grid, _WORKERS
    from spectraldns_amd.isotropic import random_solenoidal_field
    g = Grid(N, (2 * np.pi,) * 3, precision, dealias)
    u = random_solenoidal_field(N, seed=1234, k0=4.0, energy=0.5).astype(g.complex)
    for _ in range(warmup):
        u = g.rk4_step(u, dt, nu)
    times = []
    for _ in range(timed):
        t0 = time.perf_counter()
        u = g.rk4_step(u, dt, nu)
        times.append(time.perf_counter() - t0)
    return float(np.median(times)), _WORKERS, times


# Closure experiment figures of the committed summary
# profiles/closure_64cube_from_256cube_fp32.md (three demo/Closure.py runs, seeds
# 1234, 1234, 4321: 256^3 forced DNS filtered by 4, 2000 coarse steps): means
# over the runs with the per-run values; quoted, not re-measured by bench.py.
CLOSURE_SUMMARY = {
    "source": "profiles/closure_64cube_from_256cube_fp32.md",
    "runs": 5,
    "spectrum_error_time_averaged": 0.050, "spectrum_error_time_averaged_runs": [0.045, 0.078, 0.019, 0.089, 0.020],
    "spectrum_error_instantaneous_mean": 0.089,
    "dissipation_error_time_averaged": 0.068, "dissipation_error_instantaneous_mean": 0.072,
    "no_model_spectrum_error_time_averaged": 0.243,
}


def closure_timing(steps=20, warmup=3):
    """GPU time per step of the two runs the closure metric compares, measured
    here with HIP events on the solver stream: the 256^3 forced DNS (FP32, 2/3
    rule, nu = 0.0008, dt = 0.002, RK4 + forcing) and the 64^3 coarse run with
    the learned closure (FP32, 2/3 rule with the output mask, dt = 0.008, RK4
    whose stages add the network's Source, + forcing).  The network's cost
    does not depend on its weights, so canonical initial weights are used.
    ratio = GPU time of the DNS / GPU time of the coarse run over the same
    simulated time (4 DNS steps per coarse step)."""
    import torch
    from spectraldns_amd import closure as CL

    def per_step(run, n_warm, n):
        for _ in range(n_warm):
            run.step()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(n):
            run.step()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / n

    from spectraldns_amd._abi import Plan
    from spectraldns_amd.isotropic import random_solenoidal_field_gpu
    u0 = random_solenoidal_field_gpu(Plan((256,) * 3, (2 * np.pi,) * 3, "single", "2/3-rule"), seed=1234)
    dns = CL.ForcedRun((256,) * 3, 0.0008, 0.002, u_hat=u0)
    ms_dns = per_step(dns, warmup, steps)
    del dns
    torch.cuda.empty_cache()
    net = CL.Closure((64,) * 3)
    coarse = CL.ForcedRun((64,) * 3, 0.0008, 0.008, seed=1234, closure=net, dealias_mode=1)
    ms_les = per_step(coarse, warmup, 5 * steps)
    plain = CL.ForcedRun((64,) * 3, 0.0008, 0.008, seed=1234, dealias_mode=1)
    ms_plain = per_step(plain, warmup, 5 * steps)
    ratio = (ms_dns * 4.0) / ms_les
    out = {"dns_256cube_ms_per_step": round(ms_dns, 4), "dns_dt": 0.002,
           "les_64cube_ms_per_step": round(ms_les, 4), "les_dt": 0.008,
           "les_64cube_no_model_ms_per_step": round(ms_plain, 4),
           "gpu_time_ratio_dns_over_les": round(ratio, 2),
           "errors_quoted": dict(CLOSURE_SUMMARY)}
    return out


# Algorithmic HBM bytes per time step of the solver workloads, in units of
# F = N0*N1*(N2/2+1)*esz (one complex spectral field; a real physical field is
# counted as F too), 2/3 rule.  VV RHS: K/K^2 curl-inverse 6F, fused
# convection passes 40F (P1-P4 34F + convection-only x-forward 6F), the
# one-launch tail (cross2, mask, add_linear: reads F(u x w) and w_hat, writes
# rhs) 9F = 55F; RK4 stage by stage: three stage values u0 + c rhs 27F, the
# accumulator u2 = u0 + c rhs then three u2 (+)= c rhs 36F = 63F.  MHD RHS:
# six inverse transforms of the kept modes 36F (x, y, z pass: 6F per field,
# the 2/3-rule truncation inside them), Elsasser products 15F, nine forward
# transforms 54F, set_Elsasser 15F, pressure/diffusion of U 10F, diffusion
# of B 9F = 139F; RK4 over six
# components 126F.  NS integrators: fused RHS 44F (P1-P4 34F; x-forward with
# the projection reads W4 and u_hat, writes rhs and P_hat: 10F);
# ForwardEuler update 9F; AB2 updates 12F + 6F; BS5 (FSAL: 7 RHS per step
# after the first) stage sums sum_{i=1..7} 3(i+2)F = 126F, solution sum 30F,
# error sum 27F, error norm 9F, copy of the new state 6F.
SOLVER_BYTES_F = {
    ("VV", "RK4"): 4 * 55 + 63,
    ("MHD", "RK4"): 4 * 139 + 126,
    ("NS", "ForwardEuler"): 44 + 9,
    ("NS", "AB2"): 44 + 18,
    ("NS", "BS5_fixed"): 7 * 44 + 126 + 30 + 27 + 9 + 6,
}


def solver_main(args):
    """One time step of a solver/integrator pair through the reference's API
    (get_solver, get_context, getintegrator), Taylor-Green start, 2/3 rule."""
    import torch
    from spectraldns_amd import config, get_solver
    m = int(round(np.log2(args.N)))
    conv = "Divergence" if args.solver == "MHD" else "Vortex"
    config.update({"nu": 0.000625, "dt": 0.001, "T": 1e9, "eta": 0.01, "convection": conv,
                   "dealias": "2/3-rule", "precision": args.precision, "integrator": args.integrator})
    solver = get_solver(parse_args=["--M", str(m), str(m), str(m), args.solver])
    c = solver.get_context()
    X = c.X
    c.U[0] = torch.sin(X[0]) * torch.cos(X[1]) * torch.cos(X[2])
    c.U[1] = -torch.cos(X[0]) * torch.sin(X[1]) * torch.cos(X[2])
    c.U[2] = 0
    if args.solver == "MHD":
        c.B[0] = torch.sin(X[0]) * torch.sin(X[1]) * torch.cos(X[2])
        c.B[1] = torch.cos(X[0]) * torch.cos(X[1]) * torch.cos(X[2])
        c.VM.forward(c.UB, c.UB_hat)
    else:
        solver.set_velocity(**c)
        if args.solver == "VV":
            solver.cross2(c.W_hat, c.K, c.U_hat)
    config.params.t = 0.0
    config.params.tstep = 0
    solver.conv = solver.getConvection(conv)
    integrate = solver.getintegrator(c.dU, c.u, solver, c)

    def run(n):
        for _ in range(n):
            integrate()
            config.params.tstep += 1

    run(args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run(args.steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    sps = args.steps / el
    esz = 8 if args.precision == "single" else 16
    F = args.N * args.N * (args.N // 2 + 1) * esz
    nf = SOLVER_BYTES_F.get((args.solver, args.integrator))
    out = {
        "metric": "%s %s time steps/s at %d^3" % (args.solver, args.integrator, args.N),
        "value": round(sps, 3), "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1000 * el / args.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if args.precision == "single" else "f64",
        "data": "synthetic (Taylor-Green field)",
        "config": {"workload": "%s solver, %s, %d^3, 2/3 rule, nu=0.000625, dt=0.001%s"
                               % (args.solver, args.integrator, args.N, ", eta=0.01" if args.solver == "MHD" else ""),
                   "N": args.N},
    }
    if nf is not None:
        out["step_roofline"] = {"algorithmic_bytes_per_step": nf * F, "F_units": nf,
                                "achieved_GBps": round(nf * F * sps / 1e9, 1),
                                "frac": round(nf * F * sps / 1e9 / HBM_PEAK_GBS, 4)}
    config.update({"convection": "Vortex", "integrator": "RK4", "precision": "double"})
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--N", type=int, default=256)
    ap.add_argument("--dealias", default="3/2-rule", choices=("3/2-rule", "2/3-rule"))
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 measurement")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-closure", action="store_true", help="skip the closure/DNS step timing")
    ap.add_argument("--workload", default="isotropic", choices=("isotropic", "ensemble", "forced", "solver"),
                    help="isotropic: the 256^3 box (default); ensemble: the 96-box uneven ensemble; "
                         "forced: forced isotropic turbulence with compensated FP32 (use --N 1024); "
                         "solver: time steps of --solver with --integrator through the solver API")
    ap.add_argument("--solver", default="VV", choices=("NS", "VV", "MHD"))
    ap.add_argument("--integrator", default="RK4", choices=("RK4", "ForwardEuler", "AB2", "BS5_fixed"))
    ap.add_argument("--precision", default="single", choices=("single", "double"))
    ap.add_argument("--no-graph", action="store_true", help="ensemble: eager launches instead of graph replay")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MiB",
                    help="first copy a buffer of this size 3 times (reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * 2 ** 20 // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.workload == "ensemble":
        ensemble_main(args)
        return
    if args.workload == "forced":
        forced_main(args)
        return
    if args.workload == "solver":
        solver_main(args)
        return
    N = (args.N,) * 3
    elapsed, passes, diag, variants, dom = gpu_run(N, "single", args.dealias, args.steps, args.warmup)
    ms_step = 1000.0 * elapsed / args.steps
    sps = args.steps / elapsed
    pb, b_step, F = pass_bytes(N, args.dealias, 8)
    traffic = measured_traffic(N, args.dealias, dom, variants)
    dom_ms, dom_n = passes[dom]
    avg_s = dom_ms / 1000.0 / max(dom_n, 1)
    achieved = pb[dom] / avg_s / 1e9
    per_pass = {k: {"ms_per_launch": round(v[0] / max(v[1], 1), 4), "launches": v[1],
                    "GBps": round(pb[k] / (v[0] / 1000.0 / max(v[1], 1)) / 1e9, 1) if k in pb and v[1] else None}
                for k, v in passes.items()}
    out = {
        "metric": METRIC,
        "value": round(sps, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded random-phase solenoidal field, E(k)~k^4 exp(-2(k/4)^2), energy 0.5)",
        "config": {"workload": "decaying isotropic turbulence %d^3, %s%s, RK4, nu=0.001, dt=0.001, FP32 "
                               "state with FP64 energy/dissipation reductions"
                               % (args.N, args.dealias,
                                  " (%d^3 padded)" % (3 * args.N // 2) if args.dealias == "3/2-rule" else ""),
                   "N": args.N, "dealias": args.dealias, "precision": "fp32"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": (round(traffic["bytes_per_launch"]) if traffic else None),
                     "algorithmic_bytes": pb[dom], "variant": int(variants.get(dom, 0)),
                     "traffic_source": traffic["source"] if traffic else None},
        "step_roofline": {"algorithmic_bytes_per_step": b_step, "achieved_GBps": round(b_step * sps / 1e9, 1),
                          "frac": round(b_step * sps / 1e9 / HBM_PEAK_GBS, 4)},
        "passes": per_pass,
        "energy_first_last": [float(diag[0][0]), float(diag[-1][0])],
    }
    if not args.no_fp64:
        s64 = min(args.steps, 20)
        el64, pass64, _, _, _ = gpu_run(N, "double", args.dealias, s64, min(args.warmup, 2))
        pb64, b64, _ = pass_bytes(N, args.dealias, 16)
        sps64 = s64 / el64
        out["fp64"] = {"steps_per_s": round(sps64, 3), "ms_per_step": round(1000 * el64 / s64, 4),
                       "step_roofline_frac": round(b64 * sps64 / 1e9 / HBM_PEAK_GBS, 4),
                       "passes": {k: {"ms_per_launch": round(v[0] / max(v[1], 1), 4), "launches": v[1]}
                                  for k, v in pass64.items()}}
    if not args.no_closure:
        out["closure"] = closure_timing()
    if not args.no_cpu:
        t_cpu, cores, times = cpu_baseline(N, "single", args.dealias)
        out["cpu_baseline"] = {"value": round(1.0 / t_cpu, 5), "unit": "steps/s", "cores": cores, "kind": "port",
                               "cpu_model": cpu_model(),
                               "step_s": [round(t, 3) for t in times],
                               "sample": "1 warm-up + median of 5 RK4 steps of the same %d^3 %s FP32 workload "
                                         "with the NumPy/SciPy oracle (scipy.fft workers=%d)"
                                         % (args.N, args.dealias, cores)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
