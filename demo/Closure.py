"""Learned subgrid closure: trained on box-filtered forced DNS, evaluated against it.

All steps run on the GPU through libsdns_hip:
  1. forced isotropic DNS at N^3 (FP32, 2/3 rule, seeded random-phase start,
     forcing at |k| <= kf holding the energy, demo/Isotropic.py:154-184);
  2. after a spin-up, every ``snap_every`` steps the DNS velocity is filtered
     into training pairs (bar u, tau) on the (N/r)^3 grid by the box filter r
     followed by the coarse grid's spectral cutoff kc (the filter the coarse
     run represents), tau deviatoric;
  3. the closure network is trained with Adam on the pairs (the last ``val``
     pairs are held out; their loss is reported next to the zero-model loss);
  4. from the filtered DNS state, the coarse grid (2/3 rule with the output
     mask, so the state stays band-limited to |k_i| <= kc) runs with and
     without the closure at dt_c = dt_ratio * dt while the DNS continues; every
     ``sample_every`` coarse steps the shell spectra E(k), k = 1..kc, and the
     resolved dissipation are compared with those of the filtered DNS;
  5. GPU time per unit of simulated time of the DNS and of the coarse run with
     the closure (HIP events on the solver stream).
Prints one JSON report.

    python demo/Closure.py --N 256 --r 4 --les_steps 2000
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from spectraldns_amd import closure as CL  # noqa: E402
from spectraldns_amd._abi import Plan  # noqa: E402


class GpuTimer:
    def __init__(self):
        self.ms = 0.0
        self.pending = []

    def __enter__(self):
        self.a = torch.cuda.Event(enable_timing=True)
        self.b = torch.cuda.Event(enable_timing=True)
        self.a.record()
        return self

    def __exit__(self, *exc):
        self.b.record()
        self.pending.append((self.a, self.b))

    def total_ms(self):
        torch.cuda.synchronize()
        for a, b in self.pending:
            self.ms += a.elapsed_time(b)
        self.pending = []
        return self.ms


def run(args):
    t_start = time.time()
    N = (args.N,) * 3
    n = args.N // args.r
    Nc = (n, n, n)
    kc = (n + 2) // 3 - 1  # largest |k| of the coarse 2/3-rule keep set
    log = lambda *a: print("[closure]", *a, flush=True)  # noqa: E731
    dns = CL.ForcedRun(N, args.nu, args.dt, seed=args.seed, kf=args.kf)
    log("DNS %s target energy_fourier %.6f; spin-up %d steps" % (N, dns.target, args.spinup))
    dns.step(args.spinup)
    # coarse plan: 2/3 rule with the output mask |k| < 2/3 (n/2 + 1) (spectralDNS3D_short.py:44-46),
    # so the coarse state stays band-limited to |k_i| <= kc
    plan_c = Plan(Nc, (2 * np.pi,) * 3, "single", "2/3-rule", True, 1)
    pairs = []
    for s in range(args.snapshots):
        dns.step(args.snap_every)
        pairs.append(CL.filtered_pair(dns, args.r, plan_c, kc))
    log("collected %d pairs after %d DNS steps" % (len(pairs), dns.steps_done))
    train_pairs, val_pairs = pairs[:-args.val], pairs[-args.val:]
    if args.augment:
        train_pairs = train_pairs + CL.augment(train_pairs, args.augment, seed=args.seed)
    net = CL.Closure(Nc, seed=args.seed)
    net.in_scale, net.out_scale = CL.scales(train_pairs)
    zero = float(np.mean([CL.zero_model_loss(net, t) for _, t in val_pairs]))
    hist = CL.train(net, train_pairs, args.iters, lr0=args.lr, lr1=args.lr_end, log_every=max(args.iters // 10, 1),
                    val=val_pairs)
    for h in hist:
        log("iter %d train %.4e val %.4e (zero model %.4e)" % (h[0], h[1], h[2], zero))
    # coarse runs from the filtered DNS state
    ub, _ = dns.filtered(args.r)
    u0 = CL.coarse_state(plan_c, ub.float(), kc)
    dtc = args.dt * args.dt_ratio
    with_model = CL.ForcedRun(Nc, args.nu, dtc, u_hat=u0.clone(), kf=args.kf, closure=net,
                              per_stage=args.per_stage, dealias_mode=1)
    no_model = CL.ForcedRun(Nc, args.nu, dtc, u_hat=u0.clone(), kf=args.kf, target=with_model.target,
                            dealias_mode=1)
    t_dns, t_les, t_nom = GpuTimer(), GpuTimer(), GpuTimer()
    ediag = torch.zeros(2, dtype=torch.float64, device="cuda")
    samples = []
    for s in range(1, args.les_steps + 1):
        with t_dns:
            dns.step(args.dt_ratio)
        with t_les:
            with_model.step(1)
        with t_nom:
            no_model.step(1)
        if s % args.sample_every == 0:
            ubf, _ = dns.filtered(args.r)
            uf = CL.coarse_state(plan_c, ubf.float(), kc)
            ef = CL.shell_energy(plan_c, uf, kc)
            em = CL.shell_energy(with_model.plan, with_model.u_hat, kc)
            en = CL.shell_energy(no_model.plan, no_model.u_hat, kc)
            plan_c.energy(uf, args.nu, ediag)
            eps_f = float(ediag[1].item())
            eps_m = with_model.diagnostics()[1]
            eps_n = no_model.diagnostics()[1]
            samples.append({"step": s, "eps_f": eps_f, "eps_m": eps_m, "eps_n": eps_n,
                            "spec_err_model": float(np.abs(em - ef).sum() / ef.sum()),
                            "spec_err_nomodel": float(np.abs(en - ef).sum() / ef.sum()),
                            "eps_err_model": abs(eps_m - eps_f) / eps_f, "eps_err_nomodel": abs(eps_n - eps_f) / eps_f,
                            "E_filtered": ef.tolist(), "E_model": em.tolist(), "E_nomodel": en.tolist()})
            log("step %d: spectrum error model %.3f no model %.3f; dissipation error model %.3f no model %.3f" % (
                s, samples[-1]["spec_err_model"], samples[-1]["spec_err_nomodel"], samples[-1]["eps_err_model"],
                samples[-1]["eps_err_nomodel"]))
    ms_dns, ms_les, ms_nom = t_dns.total_ms(), t_les.total_ms(), t_nom.total_ms()
    T = args.les_steps * dtc
    # network evaluation alone (one Source evaluation = transforms + three conv layers)
    t_net = GpuTimer()
    ub_c = torch.empty((3,) + Nc, dtype=torch.float32, device="cuda")
    plan_c.backward(with_model.u_hat, ub_c)
    for _ in range(20):
        with t_net:
            net.forward(ub_c)
    ms_net = t_net.total_ms() / 20

    def mean(key):
        return float(np.mean([x[key] for x in samples])) if samples else None

    # time-averaged spectra over the second half of the comparison window
    late = samples[len(samples) // 2:]
    avg = {k: np.mean([x[k] for x in late], axis=0) for k in ("E_filtered", "E_model", "E_nomodel")} if late else {}
    eps_avg = {k: float(np.mean([x[k] for x in late])) for k in ("eps_f", "eps_m", "eps_n")} if late else {}

    def avg_err(key):
        return float(np.abs(avg[key] - avg["E_filtered"]).sum() / avg["E_filtered"].sum()) if late else None

    report = {
        "dns": {"N": args.N, "nu": args.nu, "dt": args.dt, "kf": args.kf, "seed": args.seed,
                "precision": "fp32", "dealias": "2/3-rule"},
        "filter": args.r, "coarse_N": n, "coarse_dt": dtc, "shells": kc,
        "training": {"pairs": len(train_pairs), "held_out": len(val_pairs), "iters": args.iters,
                     "final_val_loss": hist[-1][2], "zero_model_val_loss": zero,
                     "history": [[h[0], h[1], h[2]] for h in hist]},
        "comparison": {"coarse_steps": args.les_steps, "simulated_time": T, "samples": len(samples),
                       "spectrum_error_model": mean("spec_err_model"),
                       "spectrum_error_nomodel": mean("spec_err_nomodel"),
                       "dissipation_error_model": mean("eps_err_model"),
                       "dissipation_error_nomodel": mean("eps_err_nomodel"),
                       "time_averaged_window": [late[0]["step"], late[-1]["step"]] if late else None,
                       "avg_spectrum_error_model": avg_err("E_model"),
                       "avg_spectrum_error_nomodel": avg_err("E_nomodel"),
                       "avg_dissipation_error_model": abs(eps_avg["eps_m"] - eps_avg["eps_f"]) / eps_avg["eps_f"] if late else None,
                       "avg_dissipation_error_nomodel": abs(eps_avg["eps_n"] - eps_avg["eps_f"]) / eps_avg["eps_f"] if late else None,
                       "avg_spectra": {k: v.tolist() for k, v in avg.items()}},
        "gpu_time": {"dns_ms": ms_dns, "coarse_with_closure_ms": ms_les, "coarse_no_model_ms": ms_nom,
                     "dns_ms_per_step": ms_dns / (args.les_steps * args.dt_ratio),
                     "coarse_closure_ms_per_step": ms_les / args.les_steps,
                     "coarse_no_model_ms_per_step": ms_nom / args.les_steps,
                     "closure_eval_ms": ms_net, "source_per_stage": bool(args.per_stage),
                     "speedup_vs_dns": ms_dns / ms_les},
        "samples": [{k: v for k, v in x.items()} for x in samples],

        "wall_s": time.time() - t_start,
    }
    return report


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--N", type=int, default=256)
    ap.add_argument("--r", type=int, default=4)
    ap.add_argument("--nu", type=float, default=0.0008)
    ap.add_argument("--dt", type=float, default=0.002)
    ap.add_argument("--kf", type=float, default=2.0)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--spinup", type=int, default=1500)
    ap.add_argument("--snapshots", type=int, default=36)
    ap.add_argument("--snap_every", type=int, default=20)
    ap.add_argument("--val", type=int, default=4)
    ap.add_argument("--iters", type=int, default=3000)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr_end", type=float, default=1e-4)
    ap.add_argument("--les_steps", type=int, default=2000)
    ap.add_argument("--dt_ratio", type=int, default=4)
    ap.add_argument("--sample_every", type=int, default=100)
    ap.add_argument("--augment", type=int, default=96, help="cube-symmetry images added to the training set")
    ap.add_argument("--per_stage", action="store_true", help="evaluate the closure at every RK stage")
    ap.add_argument("--out", default="")
    args = ap.parse_args()
    rep = run(args)
    txt = json.dumps(rep)
    if args.out:
        with open(args.out, "w") as f:
            f.write(txt)
    summary = {k: rep[k] for k in ("training", "comparison", "gpu_time")}
    summary["training"] = {k: v for k, v in rep["training"].items() if k != "history"}
    summary["comparison"] = {k: v for k, v in rep["comparison"].items() if k != "avg_spectra"}
    print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    main()
