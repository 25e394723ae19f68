"""Schedule autotuner for the pass kernels.

Every pass (0 x-inverse, 1 y-inverse + curl, 2 z + cross product, 3
y-forward, 4 x-forward + RK stage) is compiled with up to two line-FFT
schedules (points per thread and the radix sequence that follows from it,
see sdns_device.hpp); the z pass also has three storage schemes (all in
registers, LDS-assisted, LDS-DMA staging), six variants in all, and the
x-forward + RK pass has its run-time row maps or its 3/2-rule stage kernels
(kept rows fixed at compile time), four variants.  The tuner times each variant of each pass with HIP
events inside real RK4 steps of the target box and keeps the fastest; the
result is stored per box in ``schedules.json`` next to this file and applied
by ``apply_schedule`` whenever a plan for that box is created.

    python -m spectraldns_amd.autotune --N 256 256 256 --precision single --dealias 3/2-rule

Batched ensembles (``_abi.Batch``) have a schedule per (pass, line length):
variant 0 or 1 (half the points per thread, twice the threads per line).
``tune_batch`` picks them by timing whole ensemble steps (graph replay, the
production path), one (pass, length) choice at a time, and stores them under
the ensemble's size list; ``Ensemble`` applies them on creation.

    python -m spectraldns_amd.autotune --ensemble --precision single --dealias 2/3-rule
"""
import argparse
import json
import os

import numpy as np

SCHEDULE_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "schedules.json")
PASSES = ("x_inverse", "y_inverse_curl", "z_cross", "y_forward", "x_forward_rk", "transforms")


def key_of(N, precision, dealias, mode=0):
    return "%dx%dx%d/%s/%s/%d" % (tuple(int(n) for n in N) + (precision, dealias, int(mode)))


def load_table(path=SCHEDULE_FILE):
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def apply_schedule(plan, path=SCHEDULE_FILE):
    """Set the stored variants for this plan's box (no-op if none stored)."""
    entry = load_table(path).get(key_of(plan.N, plan.precision, plan.dealias, plan.dealias_mode))
    if not entry:
        return None
    for name, var in entry.get("variants", {}).items():
        plan.set_variant(PASSES.index(name), int(var))
    return entry


def tune(N, L=(2 * np.pi,) * 3, precision="single", dealias="3/2-rule", mode=0, steps=3, seed=1234):
    """Time every variant of every pass on RK4 steps of a random solenoidal
    field; return {"variants": {pass: v}, "ms": {pass: [ms_v0, ms_v1, ...]}}."""
    import torch
    from ._abi import Plan
    from .isotropic import random_solenoidal_field
    plan = Plan(N, L, precision, dealias, True, mode)
    cd = torch.complex64 if precision == "single" else torch.complex128
    u0 = torch.from_numpy(random_solenoidal_field(N, L, seed=seed, k0=4.0, energy=0.5)).to(cd).cuda()
    res = {"variants": {}, "ms": {}}
    for pid, name in enumerate(PASSES[:5]):
        times = []
        for var in ((0, 1, 2, 3, 4, 5) if pid == 2 else (0, 1, 2, 3) if pid == 4 else (0, 1)):
            plan.set_variant(pid, var)
            u = u0.clone()
            plan.rk4_steps(u, [1e-3], 1e-3)       # warm-up (first launch of this variant)
            torch.cuda.synchronize()
            plan.timing_read()
            plan.timing_enable(True)
            plan.rk4_steps(u, [1e-3] * steps, 1e-3)
            torch.cuda.synchronize()
            plan.timing_enable(False)
            ms, n = plan.timing_read()[name]
            times.append(ms / max(n, 1))
        best = int(np.argmin(times))
        plan.set_variant(pid, best)
        res["variants"][name] = best
        res["ms"][name] = [round(t, 5) for t in times]
    # plain transforms (pass 5): a forward and a backward transform of the
    # three velocity components, timed with events on the plan's stream
    rd = torch.float32 if precision == "single" else torch.float64
    phys = torch.empty((3,) + tuple(N), dtype=rd, device="cuda")
    spec = u0.clone()
    stream = torch.cuda.current_stream()
    times = []
    for var in (0, 1):
        plan.set_variant(5, var)
        plan.backward(spec, phys)
        plan.forward(phys, spec)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(steps):
            plan.backward(spec, phys)
            plan.forward(phys, spec)
        b.record(stream)
        torch.cuda.synchronize()
        times.append(a.elapsed_time(b) / steps)
    best = int(np.argmin(times))
    plan.set_variant(5, best)
    res["variants"]["transforms"] = best
    res["ms"]["transforms"] = [round(t, 5) for t in times]
    return res


def batch_key(sizes, precision, dealias, mode=0):
    """Key of an ensemble: its box sizes with multiplicities, e.g. 32x32,48x32."""
    counts = {}
    for n in sizes:
        counts[int(n)] = counts.get(int(n), 0) + 1
    desc = ",".join("%dx%d" % (n, c) for n, c in sorted(counts.items()))
    return "batch/%s/%s/%s/%d" % (desc, precision, dealias, int(mode))


def apply_batch_schedule(batch, sizes, precision, dealias, mode=0, path=SCHEDULE_FILE):
    """Set the stored (pass, length) variants of this ensemble (no-op if none)."""
    entry = load_table(path).get(batch_key(sizes, precision, dealias, mode))
    if not entry:
        return None
    for item in entry.get("variants", []):
        batch.set_variant(int(item[0]), int(item[1]), int(item[2]))
    return entry


def _time_ensemble(ens, steps, reps=3):
    import torch
    ens.steps(1)                         # captures the step graph for the current layout
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        ens.steps(steps)
        b.record()
        b.synchronize()
        times.append(a.elapsed_time(b) / steps)
    return float(np.median(times))


def tune_batch(sizes=None, precision="single", dealias="2/3-rule", mode=0, steps=10):
    """One sweep over the (pass, length) schedule choices of an ensemble:
    keep variant 1 where it makes the ensemble step faster.  Returns
    {"variants": [[pass, length, v], ...], "ms": {"pass/length": [ms_v0, ms_v1]},
    "step_ms": [before, after]}."""
    from .ensemble import Ensemble
    ens = Ensemble(sizes, precision=precision, dealias=dealias, dealias_mode=mode, tuned=False)
    base = _time_ensemble(ens, steps)
    start = base
    chosen, ms = [], {}
    for k, lengths in sorted(ens.batch.line_lengths().items()):
        for length in lengths:
            ens.reset()
            ens.batch.set_variant(k, length, 1)
            t1 = _time_ensemble(ens, steps)
            ms["%d/%d" % (k, length)] = [round(base, 5), round(t1, 5)]
            if t1 < base:
                base = t1
                chosen.append([k, length, 1])
            else:
                ens.batch.set_variant(k, length, 0)
    return {"variants": chosen, "ms": ms, "step_ms": [round(start, 5), round(base, 5)]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--N", type=int, nargs=3, default=[256, 256, 256])
    ap.add_argument("--precision", default="single", choices=("single", "double"))
    ap.add_argument("--dealias", default="3/2-rule", choices=("3/2-rule", "2/3-rule", "None"))
    ap.add_argument("--mode", type=int, default=0)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--out", default=SCHEDULE_FILE)
    ap.add_argument("--ensemble", action="store_true", help="tune the default 96-box ensemble instead")
    args = ap.parse_args()
    if args.ensemble:
        from .ensemble import default_sizes
        sizes = default_sizes()
        res = tune_batch(sizes, precision=args.precision, dealias=args.dealias, mode=args.mode)
        key = batch_key(sizes, args.precision, args.dealias, args.mode)
    else:
        res = tune(tuple(args.N), precision=args.precision, dealias=args.dealias, mode=args.mode, steps=args.steps)
        key = key_of(args.N, args.precision, args.dealias, args.mode)
    table = load_table(args.out)
    table[key] = res
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    print(json.dumps({key: res}), flush=True)


if __name__ == "__main__":
    main()
