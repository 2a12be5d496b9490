#!/usr/bin/env python3
"""bench.py -- retrograde throughput of the MI355X king-slice engine.

Workload (default, one GPU): the KQRvKR table -- 462 king slices x 226,920
positions per side (62*61*60), exhaustive, no input data -- generated
end to end on the device: capture exits resolved against the resident
KQvKR / KRvKR / KQRvK tables (generated beforehand, untimed), illegal and
mate init, then the ply loop of the reference's main() to its fixpoint.
One step = one such generation.

value = position-side-plies per second over the whole step: every executed
half-step (L or W, one side) covers 462 * 226,920 positions for half a ply,
so a full ply of both sides is 2 * 104,837,040 position-side-plies.
roofline: the kernel with the largest device time; algorithmic bytes per
launch = its bits per position (pred 2, verify 3, W sweep 4; 11 per
position-side-ply in all) * positions it covers / 8, over its average
launch time measured with HIP events on the engine's stream.
cpu_baseline: the CPU oracle (a plain-C restatement of the reference's
algorithm) on the host cores, timed over plies 1..2 of the same material.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--form batched|plain]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BITS = {"pred": 2, "verify": 3, "sweep": 4}
METRIC = "retro positions/s per GPU + HBM GB/s fraction; p50/p99 slice latency vs plain"


def cpu_threads():
    n = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    return max(1, min(n, 16))


def run_steps(table, subs, form, steps, tbg):
    """Time `steps` full generations; returns (seconds list, timing, half_steps, lines)."""
    secs = []
    table.set_profiling(True)
    for _ in range(steps):
        t0 = time.perf_counter()
        table.resolve(subs)
        lines = table.generate(form)
        secs.append(time.perf_counter() - t0)
    timing = table.timing()
    table.set_profiling(False)
    return secs, timing, lines


def run_synth(args, tbg):
    """One king slice at real shape with seeded Bernoulli(0.03) win / loss
    exits for captures and king moves; one step = capture exits + init + the
    ply loop to the fixpoint."""
    material = args.material if args.material != "KQRvKR" else "KQRBvKQRN"
    t0 = time.perf_counter()
    table = tbg.Table(material, synth_slice=args.slice, seed=args.seed)
    table.set_scatter(tbg.SCATTER_BLOCK if args.scatter == "block" else tbg.SCATTER_GLOBAL)
    alloc_s = time.perf_counter() - t0
    n = table.size
    print(f"synth {material} slice {args.slice}: {n} positions per side, "
          f"{table.info.device_bytes / 2**30:.1f} GiB resident, alloc {alloc_s:.1f} s",
          file=sys.stderr, flush=True)
    secs = []
    table.set_profiling(True)
    for step in range(args.warmup + args.steps):
        t0 = time.perf_counter()

        def progress(line, _u, t0=t0):
            print(f"  [{time.perf_counter() - t0:8.2f} s] {line.decode()}", file=sys.stderr,
                  flush=True)

        fn = tbg.LINE_FN(progress)
        table.resolve([])
        form = {"batched": tbg.FORM_BATCHED, "graph": tbg.FORM_GRAPH,
                "plain": tbg.FORM_PLAIN}[args.form]
        tbg.check(tbg.lib().tbg_generate(table.ptr, form, fn, None,
                                         tbg.ctypes.byref(table.counts)), "generate")
        dt = time.perf_counter() - t0
        if step == args.warmup - 1:
            table.set_profiling(False)
            table.set_profiling(True)
        if step >= args.warmup:
            secs.append(dt)
    tm = table.timing()
    steps = max(1, args.steps)
    half_steps = tm.half_steps / steps
    psp = half_steps * n / 2.0
    cats = {"pred": (tm.ms_pred, tm.n_pred), "verify": (tm.ms_verify, tm.n_verify),
            "sweep": (tm.ms_sweep, tm.n_sweep)}
    dom = max(cats, key=lambda k: cats[k][0])
    avg_ms = cats[dom][0] / max(1, cats[dom][1])
    alg = BITS[dom] * n / 8.0
    achieved = alg / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    sweep_ms = tm.ms_sweep / max(1, tm.n_sweep)
    sweep_gbs = BITS["sweep"] * n / 8.0 / (sweep_ms * 1e-3) / 1e9 if sweep_ms > 0 else 0.0
    out = {
        "metric": METRIC, "value": psp / (sum(secs) / len(secs)),
        "unit": "position-side-plies/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1000 * max(secs), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": f"{material} single king slice {args.slice} at real shape, "
                               f"seeded hash exits (seed {args.seed}), full fixpoint",
                   "material": material, "slice": args.slice, "positions_per_side": n,
                   "half_steps_per_step": half_steps, "plies": table.counts.nplies},
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "alg_bytes_per_launch": alg, "avg_launch_ms": avg_ms},
        "sweep_roofline": {"achieved": sweep_gbs, "frac": sweep_gbs / HBM_PEAK_GBS,
                           "avg_launch_ms": sweep_ms},
        "kernels_ms_per_step": {k: getattr(tm, "ms_" + k) / steps for k in
                                ("pred", "verify", "sweep", "uncap", "init", "resolve")},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--material", default="KQRvKR")
    ap.add_argument("--form", default="batched", choices=["batched", "graph", "plain"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-form comparison")
    ap.add_argument("--cpu-plies", type=int, default=2)
    ap.add_argument("--scatter", default="block", choices=["block", "global"])
    ap.add_argument("--workload", default="table", choices=["table", "synth"],
                    help="table: a whole material (default KQRvKR); synth: one king slice of "
                         "--material with seeded hash exits (default KQRBvKQRN slice 170)")
    ap.add_argument("--slice", type=int, default=170)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against known bytes)")
    args = ap.parse_args()

    if args.copy_probe:
        import torch
        n = args.copy_probe << 20
        a = torch.ones(n, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n, "copies": 3}))
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this engine is measured on one GPU; multi-GPU slice sharding is "
              "designed (DESIGN.md) but not built, so --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import tb8gen_amd as tbg

    tbg.init(0)
    if args.workload == "synth":
        return run_synth(args, tbg)
    form = {"batched": tbg.FORM_BATCHED, "graph": tbg.FORM_GRAPH, "plain": tbg.FORM_PLAIN}[args.form]
    reg = tbg.Registry(tbg.FORM_BATCHED)
    for sub in tbg.sub_materials(args.material):
        reg.get(sub)
    subs = list(reg.tables.values())
    table = tbg.Table(args.material)
    table.set_scatter(tbg.SCATTER_BLOCK if args.scatter == "block" else tbg.SCATTER_GLOBAL)
    n_total = tbg.NSLICES * table.size

    for _ in range(args.warmup):
        table.resolve(subs)
        table.generate(form)
    secs, tm, lines = run_steps(table, subs, form, args.steps, tbg)
    ms_step = 1000.0 * max(secs)
    half_steps = tm.half_steps / max(1, args.steps)
    psp = half_steps * n_total / 2.0  # position-side-plies per step
    value = psp / (sum(secs) / len(secs))

    cats = {"pred": (tm.ms_pred, tm.n_pred), "verify": (tm.ms_verify, tm.n_verify),
            "sweep": (tm.ms_sweep, tm.n_sweep)}
    dom = max(cats, key=lambda k: cats[k][0])
    dom_ms, dom_n = cats[dom]
    avg_ms = dom_ms / max(1, dom_n)
    covered = table.size if form == tbg.FORM_PLAIN else n_total
    alg_bytes = BITS[dom] * covered / 8.0
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{args.material}_{args.form}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get(dom)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "position-side-plies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.material} table generation, {args.form} form: 462 king slices x "
                        f"{table.size} positions per side, capture exits from resident smaller "
                        "tables, init + ply loop to the fixpoint",
            "material": args.material, "form": args.form, "scatter": args.scatter, "slices": 462,
            "positions_per_side": n_total, "half_steps_per_step": half_steps,
            "plies": table.counts.nplies,
        },
        "roofline": {
            "kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "alg_bytes_per_launch": alg_bytes, "avg_launch_ms": avg_ms, "launches": dom_n / max(1, args.steps),
        },
        "kernels_ms_per_step": {
            "pred": tm.ms_pred / args.steps, "verify": tm.ms_verify / args.steps,
            "sweep": tm.ms_sweep / args.steps, "uncap": tm.ms_uncap / args.steps,
            "init": tm.ms_init / args.steps, "resolve": tm.ms_resolve / args.steps,
            "launches": tm.launches / args.steps,
        },
    }

    if not args.no_plain and form != tbg.FORM_PLAIN:
        t0 = time.perf_counter()
        table.resolve(subs)
        plines = table.generate(tbg.FORM_PLAIN)
        pt = time.perf_counter() - t0
        out["plain"] = {"ms_per_step": 1000 * pt, "value": psp / pt,
                        "batched_over_plain": (psp / (sum(secs) / len(secs))) / (psp / pt),
                        "same_count_lines": plines == lines}

    if not args.no_cpu:
        from oracle import oracle as O
        threads = cpu_threads()
        O.set_threads(threads)
        sec, hs = O.time_plies(args.material, args.cpu_plies)
        out["cpu_baseline"] = {
            "value": hs * n_total / 2.0 / sec, "unit": "position-side-plies/s", "cores": threads,
            "kind": "port",
            "sample": f"CPU oracle (C restatement of the reference algorithm, OpenMP) on plies "
                      f"1..{args.cpu_plies} of {args.material}: {hs} half-steps over all 462 "
                      f"slices in {sec:.2f} s; smaller tables generated untimed",
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
