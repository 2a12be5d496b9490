#!/usr/bin/env python3
"""bench.py -- retrograde throughput of the MI355X king-slice engine.

Workload (default, one GPU): one king slice of the 8-piece material
KQRBvKQRN at real shape -- slice 170 (white king d1, black king e8), six
distinct non-kings over 62 squares, 44,261,653,680 positions per side,
1-bit bitmaps -- with capture and king-move exits drawn from a seeded
splitmix64 hash (win below 0.03, loss in [0.03, 0.06)), run to its full
fixpoint.  One step = capture exits + init + the ply loop to the fixpoint.

value = position-side-plies per second, the mean over the timed steps:
every executed half-step (L or W, one side) covers the slice's positions
for half a ply.
roofline: the kernel with the largest device time (pred, verify, sweep,
init or the static refutations' kernel xr): its algorithmic bytes per launch
(bits per position: pred 2, verify 3, W sweep 4 per executed half-step; init
10 once per step: ten output planes; xr 6: two planes read and written back,
two read) over its average launch time from HIP events on the engine's
stream; traffic = HBM bytes per launch from the committed rocprof counter
summary (profiles/).
secondary: the KQRvKR table (462 slices x 226,920 positions per side)
generated end to end in the batched form.
cpu_baseline: the CPU oracle (a plain-C restatement of the reference's
algorithm, OpenMP) on the host cores, on a bounded sample of the same kind
of job: the full fixpoint, init included, of the 6-piece synthetic slice
KQRvKRN 170 with the same exit hash (also timed on the GPU for a
like-for-like ratio), plus the init of 2^24 positions of the 8-piece slice
itself; the extrapolation to the whole 8-piece slice is stated.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
       [--workload synth|table|stream|served] [--form batched|graph|plain] ...
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BITS = {"pred": 2, "verify": 3, "sweep": 4}
XR_BITS = 6
METRIC = "retro positions/s per GPU + HBM GB/s fraction; p50/p99 slice latency vs plain"


def cpu_threads():
    """Threads for the CPU oracle: every core the process may use, within
    the OMP_NUM_THREADS share when one is set."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    lim = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return max(1, min(n, lim) if lim > 0 else n)


def host_cpu():
    """nproc and the lscpu model line of this host."""
    import subprocess
    info = {"nproc": os.cpu_count()}
    try:
        info["nproc"] = int(subprocess.run(["nproc"], capture_output=True, text=True).stdout)
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for l in out.splitlines():
            k, _, v = l.partition(":")
            if k.strip() in ("Model name", "CPU(s)", "Socket(s)", "Thread(s) per core"):
                info[k.strip()] = v.strip()
    except OSError:
        pass
    return info


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


def kernel_rooflines(tm, lines, positions, steps, traffic=None, init_bits=10):
    """Per kernel: algorithmic bytes per launch = its bits per position x
    positions x the half-steps it executed (from the count lines: pred runs
    for every W step and every L step after ply 1, verify for L steps, sweep
    for W steps) / its launches; the init kernel writes init_bits bits per
    position once per step (illegal and mates per side, and for a synthetic
    job the four exit-candidate planes and the exit part of the static
    refutations), the refutation kernel (xr) XR_BITS (it reads and writes
    both refutation planes and reads both wins planes).  Average launch time
    from the HIP events; achieved GB/s = the two divided.  traffic: measured
    HBM bytes per launch (copy-calibrated rocprof counters) where a profile
    is committed."""
    lh = [l for l in lines if re.match(r"l\d+_[wb] = ", l) and not l.startswith("l0_")]
    wh = [l for l in lines if re.match(r"w\d+_[wb] = ", l)]
    l1 = [l for l in lh if l.startswith("l1_")]
    hs = {"pred": len(lh) - len(l1) + len(wh), "verify": len(lh), "sweep": len(wh)}
    out = {}
    for k in ("pred", "verify", "sweep", "init", "xr"):
        ms = getattr(tm, "ms_" + k) / steps
        n = getattr(tm, "n_" + k) / steps
        if k == "xr" and n == 0:
            continue
        if k in ("init", "xr"):
            alg = (init_bits if k == "init" else XR_BITS) * positions / 8.0
        else:
            alg = BITS[k] * positions * hs[k] / 8.0 / max(n, 1.0)
        avg = ms / max(n, 1.0)
        ach = alg / (avg * 1e-3) / 1e9 if avg > 0 else 0.0
        out[k] = {"kernel": k, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
                  "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                  "traffic": (traffic or {}).get(k), "alg_bytes_per_launch": alg,
                  "avg_launch_ms": avg, "launches": n, "half_steps": hs.get(k, 1), "ms": ms}
        if k in ("pred", "verify", "sweep"):
            # A sparse ply reads and writes only the chunks its summaries
            # mark, so the dense-plane rate is an equivalent, not bytes
            # moved (it can exceed the peak); achieved/frac are then the
            # profiled bytes per launch over the live launch time, or None
            # without a profile.
            t = out[k]["traffic"]
            mea = t / (avg * 1e-3) / 1e9 if t and avg > 0 else None
            out[k].update({"dense_equiv_gbs": ach, "dense_equiv_frac": ach / HBM_PEAK_GBS,
                           "achieved": mea,
                           "frac": mea / HBM_PEAK_GBS if mea is not None else None})
    return out


def load_traffic(name):
    path = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


def scatter_mode(name, tbg):
    mode = tbg.SCATTER_BLOCK if name.startswith("block") else tbg.SCATTER_GLOBAL
    return mode | (tbg.SCATTER_COMBINE if name.endswith("+combine") else 0)


def pawnless_materials(extra):
    """Every pawnless material with `extra` non-king pieces, in the colour
    orientation the engine generates (stronger side white)."""
    import itertools

    import tb8gen_amd as tbg
    out = []
    for w in range(extra, -1, -1):
        b = extra - w
        if b > w:
            continue
        for ws in itertools.combinations_with_replacement("QRBN", w):
            for bs in itertools.combinations_with_replacement("QRBN", b):
                m = tbg.canonical("K" + "".join(ws) + "vK" + "".join(bs))
                if m not in out:
                    out.append(m)
    return out


def run_stream(args, tbg):
    """Config 3: every pawnless 3-, 4- and 5-piece material (4 + 20 + 60)
    as a stream of king-slice jobs, generated level by level in dependency
    order, each level as one batch (one launch per phase covers all of its
    materials x 462 slices x both sides).  Optionally the 4-piece level again
    in the plain form (one slice per launch) for the ratio."""
    form = {"batched": tbg.FORM_BATCHED, "graph": tbg.FORM_GRAPH, "plain": tbg.FORM_PLAIN}[args.form]
    levels = [pawnless_materials(1), pawnless_materials(2), pawnless_materials(3)]
    tables = {}
    for lev in levels:  # allocation is not part of the timed region
        for m in lev:
            tables[m] = tbg.Table(m)
    geoms = {}
    if args.autotune:
        # untimed pass: each level generated once, its launch geometry tuned
        # on its largest material and given to all of its materials
        done = []
        for lev in levels:
            subs = [tables[m] for m in done]
            for m in lev:
                tables[m].resolve(subs)
            big = max(lev, key=lambda m: tables[m].size)
            g, _ = tables[big].autotune()
            for m in lev:
                tables[m].set_geometry(g)
            tbg.generate_batch([tables[m] for m in lev], form)
            geoms[len(lev[0]) - 1] = g.as_dict()
            done += lev
    done = []
    per_level = []
    total_s, total_psp = 0.0, 0.0
    for lev in levels:
        t0 = time.perf_counter()
        subs = [tables[m] for m in done]
        for m in lev:
            tables[m].resolve(subs)
        batch = [tables[m] for m in lev]
        lines = tbg.generate_batch(batch, form)
        dt = time.perf_counter() - t0
        psp = sum((len(l) - 4) * tbg.NSLICES * t.size / 2.0 for l, t in zip(lines, batch))
        per_level.append({"pieces": len(lev[0]) - 1, "materials": len(lev), "s": dt,
                          "position_side_plies": psp, "value": psp / dt,
                          "geometry": geoms.get(len(lev[0]) - 1)})
        print(f"level {len(lev[0]) - 1}: {len(lev)} materials, {dt:.2f} s, "
              f"{psp / dt:.3e} position-side-plies/s", file=sys.stderr, flush=True)
        total_s += dt
        total_psp += psp
        done += lev
    out = {
        "metric": METRIC, "value": total_psp / total_s, "unit": "position-side-plies/s",
        "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": 1000 * total_s,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": f"stream of all {sum(len(l) for l in levels)} pawnless 3/4/5-piece "
                               f"materials, {args.form} form, level by level",
                   "form": args.form, "levels": per_level},
    }
    if args.stream_plain:
        # the plain form (one launch per slice and phase, a host round trip
        # per ply) on the whole 4-piece level and on a sample of six 5-piece
        # materials, against the batched rate of the same level
        for key, li, lev in (("plain_4piece_level", 1, levels[1]),
                             ("plain_5piece_sample", 2, levels[2][::10])):
            base = [tables[m] for m in sum(levels[:li], [])]
            fresh = [tbg.Table(m) for m in lev]
            t0 = time.perf_counter()
            for t in fresh:
                t.resolve(base)
            plines = tbg.generate_batch(fresh, tbg.FORM_PLAIN)
            pt = time.perf_counter() - t0
            psp = sum((len(l) - 4) * tbg.NSLICES * t.size / 2.0 for l, t in zip(plines, fresh))
            same = all(pl == tables[m].lines for pl, m in zip(plines, lev))
            out[key] = {"materials": lev, "s": pt, "value": psp / pt,
                        "batched_over_plain": per_level[li]["value"] / (psp / pt),
                        "same_count_lines": same}
            print(f"{key}: {out[key]}", file=sys.stderr, flush=True)
            for t in fresh:
                t.close()
    print(json.dumps(out), flush=True)
    for t in tables.values():
        t.close()


def draw_jobs(args, tbg, pieces, rng):
    """King-slice jobs of the pawnless materials with the given piece
    counts (1 to --max-slices slices each, seeded hash exits), drawn until
    their device bytes reach --fill of the budget."""
    mats = sum((pawnless_materials(p - 2) for p in pieces), [])
    budget = int(args.budget_gb * (1 << 30))
    sizes = {}
    for m in mats:
        t = tbg.Table(m, synth_slice=0, seed=1, keep_codes=True)
        sizes[m] = (t.info.device_bytes, t.size)
        t.close()
    jobs, total = [], 0
    while total < args.fill * budget and len(jobs) < args.jobs:
        m = mats[int(rng.integers(len(mats)))]
        k = int(rng.integers(1, args.max_slices + 1))
        jobs.append((m, int(rng.integers(0, 441)), int(rng.integers(1, 1 << 62)), k))
        total += sizes[m][0] * k
    positions = [sizes[m][1] * k for m, _, _, k in jobs]
    return jobs, total, positions


def run_served(args, tbg):
    """Config 4: seeded arrivals of king-slice jobs of 6-piece pawnless
    materials (1 to --max-slices king slices each, seeded hash exits),
    served with continuous batching under an HBM budget, against the plain
    server (one slice at a time, FIFO).  The job list is drawn until its
    admission estimates reach --fill of the budget, so that the capacity run
    (every job queued at once) loads HBM to about that fraction; latency
    (arrival -> codes on the host, p50 / p99) is measured at seeded Poisson
    arrivals of 70% of the batched capacity, and once more under a budget
    that forces preemption to pinned host memory.  The capacity comparison
    is repeated on a mixed stream of 5- and 6-piece jobs (--mixed-pieces),
    the stream the batched-over-plain target is stated for.  The server runs
    with --max-active slice slots (32: faster than more, see
    capacity_resident, the capacity run at --residency-slots, where the
    budget alone bounds admission and most of the job state is resident)."""
    import numpy as np
    pieces = [int(x) for x in args.served_pieces.split(",")]
    rng = np.random.default_rng(args.seed)
    budget = int(args.budget_gb * (1 << 30))
    jobs, total, positions = draw_jobs(args, tbg, pieces, rng)
    print(f"served: {len(jobs)} jobs, {sum(k for *_, k in jobs)} slices, "
          f"{min(positions):.3g}-{max(positions):.3g} positions per job, "
          f"{total / 2**30:.1f} GiB of job state", file=sys.stderr, flush=True)

    def serve(arrivals, form, slots, bud, jl=None):
        jl = jobs if jl is None else jl
        srv = tbg.Server(bud, slots)
        ids = [srv.submit(m, s, seed, a, k) for (m, s, seed, k), a in zip(jl, arrivals)]
        st = srv.run(form)
        res = [srv.result(i) for i in ids]
        srv.close()
        lat = np.array([r.done_ms - r.arrival_ms for r in res if r.status == 0])
        return st, res, lat

    def summary(st, lat):
        return {"wall_ms": st.wall_ms, "p50_ms": float(np.percentile(lat, 50)),
                "p99_ms": float(np.percentile(lat, 99)),
                "mean_resident_gb": st.mean_resident_bytes / 2**30,
                "peak_resident_gb": st.peak_resident_bytes / 2**30,
                "peak_active_slices": st.peak_active, "plies": st.plies,
                "jobs_done": st.jobs_done, "spills": st.spills, "restores": st.restores,
                "spill_ms": st.spill_ms,
                "value": st.position_side_plies / (st.wall_ms / 1000.0)}

    cap = {}
    for name, form in (("batched", tbg.FORM_BATCHED), ("plain", tbg.FORM_PLAIN)):
        st, _, lat = serve([0.0] * len(jobs), form, args.max_active, budget)
        cap[name] = dict(summary(st, lat), jobs_per_s=len(jobs) / (st.wall_ms / 1000.0))
        print(f"capacity {name}: {cap[name]}", file=sys.stderr, flush=True)
    # the same jobs with enough slots that admission is bound by the budget:
    # most of the job state resident at once (slower than a few dozen slots,
    # also when the running batch alone is capped at that: residency itself
    # costs here)
    resident = None
    if args.residency_slots:
        st, _, lat = serve([0.0] * len(jobs), tbg.FORM_BATCHED, args.residency_slots, budget)
        resident = dict(summary(st, lat), jobs_per_s=len(jobs) / (st.wall_ms / 1000.0),
                        slots=args.residency_slots)
        print(f"capacity batched, {args.residency_slots} slots: {resident}", file=sys.stderr,
              flush=True)
    mixed = {}
    if args.mixed_pieces:
        mp = [int(x) for x in args.mixed_pieces.split(",")]
        mjobs, mtotal, mpos = draw_jobs(args, tbg, mp, np.random.default_rng(args.seed + 1))
        for name, form in (("batched", tbg.FORM_BATCHED), ("plain", tbg.FORM_PLAIN)):
            st, _, lat = serve([0.0] * len(mjobs), form, args.max_active, budget, mjobs)
            mixed[name] = dict(summary(st, lat), jobs_per_s=len(mjobs) / (st.wall_ms / 1000.0))
            print(f"mixed capacity {name}: {mixed[name]}", file=sys.stderr, flush=True)
        mixed = {"pieces": mp, "jobs": len(mjobs), "slices": sum(k for *_, k in mjobs),
                 "job_state_gb": mtotal / 2**30, "positions_per_job": [min(mpos), max(mpos)],
                 "capacity": mixed,
                 "batched_over_plain": mixed["batched"]["value"] / mixed["plain"]["value"]}
    lam = 0.7 * cap["batched"]["jobs_per_s"]  # offered load: 70 % of the batched capacity
    arrivals = list(np.cumsum(rng.exponential(1.0 / lam, len(jobs))) * 1000.0)
    out_modes = {}
    for name, form, bud in (("batched", tbg.FORM_BATCHED, budget), ("plain", tbg.FORM_PLAIN, budget),
                            ("batched_small_budget", tbg.FORM_BATCHED, budget // 64)):
        st, _, lat = serve(arrivals, form, args.max_active, bud)
        out_modes[name] = dict(summary(st, lat), budget_gb=bud / 2**30)
        print(f"{name} at {lam:.1f} jobs/s: {out_modes[name]}", file=sys.stderr, flush=True)
    b, p = out_modes["batched"], out_modes["plain"]
    out = {
        "metric": METRIC, "value": cap["batched"]["value"], "unit": "position-side-plies/s",
        "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": cap["batched"]["wall_ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": f"served: {len(jobs)} jobs of 1-{args.max_slices} king slices of "
                               f"{'/'.join(map(str, pieces))}-piece pawnless materials with seeded "
                               f"hash exits ({min(positions):.3g}-{max(positions):.3g} positions "
                               f"per job, {total / 2**30:.1f} GiB of state = "
                               f"{100 * total / budget:.0f}% of the budget); value = capacity "
                               f"(all jobs queued at once); latency at seeded Poisson arrivals of "
                               f"70% of the batched capacity ({lam:.1f} jobs/s); "
                               f"{args.budget_gb:g} GB budget, {args.max_active} slots",
                   "seed": args.seed, "lambda_jobs_per_s": lam, "budget_gb": args.budget_gb,
                   "pieces": pieces, "jobs": len(jobs), "job_state_gb": total / 2**30},
        "capacity": cap,
        "capacity_resident": resident,
        "batched_over_plain": cap["batched"]["value"] / cap["plain"]["value"],
        "latency_ms": {"batched_p50": b["p50_ms"], "batched_p99": b["p99_ms"],
                       "plain_p50": p["p50_ms"], "plain_p99": p["p99_ms"]},
        "served": out_modes,
        "mixed_stream": mixed or None,
    }
    print(json.dumps(out), flush=True)


def synth_steps(table, form, steps, warmup, tbg, progress=True):
    """warmup + steps full generations of a synthetic slice; returns the
    timed steps' seconds and the last step's count lines."""
    secs, lines = [], []
    table.set_profiling(True)
    for step in range(warmup + steps):
        t0 = time.perf_counter()
        lines = []

        def cb(line, _u, t0=t0, lines=lines):
            lines.append(line.decode())
            if progress:
                print(f"  [{time.perf_counter() - t0:8.2f} s] {line.decode()}", file=sys.stderr,
                      flush=True)

        fn = tbg.LINE_FN(cb)
        table.resolve([])
        tbg.check(tbg.lib().tbg_generate(table.ptr, form, fn, None,
                                         tbg.ctypes.byref(table.counts)), "generate")
        dt = time.perf_counter() - t0
        if step == warmup - 1:  # the timings cover the timed steps only
            table.set_profiling(False)
            table.set_profiling(True)
        if step >= warmup:
            secs.append(dt)
    return secs, lines


def cpu_baseline_synth(args, tbg, material8, n8, psp8):
    """The CPU oracle on a bounded sample of the same kind of job, on the
    host's cores: the full fixpoint (init included, no codes) of the 6-piece
    synthetic slice KQRvKRN 170 with the same exit hash -- also generated on
    the GPU for a like-for-like ratio -- and the init of 2^24 positions of the
    8-piece slice itself.  The whole 8-piece slice on the CPU is extrapolated
    from the two rates."""
    from oracle import oracle as O
    threads = cpu_threads()
    info = host_cpu()
    print(f"cpu baseline: {threads} threads; host {info}", file=sys.stderr, flush=True)
    O.set_threads(threads)
    O.set_keep_synth_codes(False)
    m6, s6 = "KQRvKRN", args.slice
    log = f"/tmp/bench_cpu_{m6}.log"
    t0 = time.perf_counter()
    o = O.Table(m6, synth_slice=s6, seed=args.seed, log_path=log)
    sec6 = time.perf_counter() - t0
    with open(log) as f:
        olines = [l.strip() for l in f if l.strip()]
    hs6 = sum(1 for l in olines if re.match(r"[lw]\d+_[wb] = ", l) and not l.startswith("l0_"))
    cpu6 = hs6 * o.size / 2.0 / sec6
    # the same job on the GPU
    t = tbg.Table(m6, synth_slice=s6, seed=args.seed)
    gsecs, glines = synth_steps(t, tbg.FORM_BATCHED, 5, 1, tbg, progress=False)
    t.close()
    gpu6 = hs6 * o.size / 2.0 / (sum(gsecs) / len(gsecs))
    n_init = 1 << 24
    sec_i, _ = O.synth_init_time(material8, args.slice, args.seed, 0, n_init)
    init_rate = n_init / sec_i
    extrap = n8 / init_rate + psp8 / cpu6
    print(f"cpu baseline: {m6} {sec6:.2f} s ({cpu6:.3e}/s; GPU {gpu6:.3e}/s), "
          f"8-piece init sample {init_rate:.3e} positions/s", file=sys.stderr, flush=True)
    return {
        "value": cpu6, "unit": "position-side-plies/s", "cores": threads, "kind": "port",
        "sample": f"CPU oracle (C restatement of the reference algorithm, OpenMP, {threads} "
                  f"threads) on the full fixpoint, init included, of the synthetic 6-piece "
                  f"slice {m6} {s6} (seed {args.seed}, same exit hash): {hs6} half-steps x "
                  f"{o.size} positions in {sec6:.2f} s; the same job on the GPU runs at "
                  f"{gpu6:.3e}/s ({gpu6 / cpu6:.0f}x)",
        "host": info,
        "like_for_like": {"job": f"{m6} slice {s6} synthetic, full fixpoint",
                          "cpu": cpu6, "gpu": gpu6, "gpu_over_cpu": gpu6 / cpu6,
                          "same_count_lines": glines == olines},
        "init_sample": {"job": f"{material8} slice {args.slice}, indices [0, 2^24)",
                        "seconds": sec_i, "positions_per_s": init_rate},
        "extrapolated_8piece_s": extrap,
        "extrapolation": "8-piece slice positions / the init sample's rate + the slice's "
                         "position-side-plies / the 6-piece job's rate (a lower bound: the "
                         "per-position cost grows with the piece count)",
    }


def run_table(args, tbg, steps, warmup, extras=True, material=None):
    """A whole table (default KQRvKR) generated end to end (smaller tables
    resident and generated untimed); returns the bench dict."""
    form = {"batched": tbg.FORM_BATCHED, "graph": tbg.FORM_GRAPH, "plain": tbg.FORM_PLAIN}[args.form]
    material = material or args.material or "KQRvKR"
    reg = tbg.NativeRegistry(tbg.FORM_GRAPH)
    subs = [reg.get(m) for m in tbg.sub_materials(material)]
    table = tbg.Table(material)
    table.set_scatter(scatter_mode(args.scatter, tbg))
    n_total = tbg.NSLICES * table.size
    tuned = None
    if args.autotune:
        table.resolve(subs)
        g, tms = table.autotune()
        tuned = dict(g.as_dict(), trial_ms=tms)
    for _ in range(warmup):
        table.resolve(subs)
        table.generate(form)
    secs, tm, lines = run_steps(table, subs, form, steps, tbg)
    mean = sum(secs) / len(secs)
    half_steps = tm.half_steps / max(1, steps)
    psp = half_steps * n_total / 2.0  # position-side-plies per step
    kr = kernel_rooflines(tm, lines, n_total, steps, load_traffic(f"{material}_{args.form}"))
    dom = max(kr, key=lambda k: kr[k]["ms"])
    out = {
        "metric": METRIC, "value": psp / mean, "unit": "position-side-plies/s", "n_gpus": 1,
        "steps": steps, "warmup": warmup, "ms_per_step": 1000.0 * mean,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": f"{material} table generation, {args.form} form: 462 king slices x "
                        f"{table.size} positions per side, capture exits from resident smaller "
                        "tables, init + ply loop to the fixpoint",
            "material": material, "form": args.form, "slices": 462,
            "geometry": table.geometry().as_dict(), "autotuned": tuned is not None,
            "positions_per_side": n_total, "half_steps_per_step": half_steps,
            "plies": table.counts.nplies,
        },
        "roofline": {k: v for k, v in kr[dom].items() if k not in ("ms", "half_steps")},
        "kernel_rooflines": kr,
        "kernels_ms_per_step": {k: getattr(tm, "ms_" + k) / steps for k in
                                ("pred", "verify", "sweep", "uncap", "init", "xr", "resolve")},
    }
    if extras and not args.no_plain and form != tbg.FORM_PLAIN:
        t0 = time.perf_counter()
        table.resolve(subs)
        plines = table.generate(tbg.FORM_PLAIN)
        pt = time.perf_counter() - t0
        out["plain"] = {"ms_per_step": 1000 * pt, "value": psp / pt,
                        "batched_over_plain": (psp / mean) / (psp / pt),
                        "same_count_lines": plines == lines}
    if extras and not args.no_cpu:
        from oracle import oracle as O
        threads = cpu_threads()
        O.set_threads(threads)
        sec, hs = O.time_plies(material, args.cpu_plies)
        out["cpu_baseline"] = {
            "value": hs * n_total / 2.0 / sec, "unit": "position-side-plies/s", "cores": threads,
            "kind": "port", "host": host_cpu(),
            "sample": f"CPU oracle (C restatement of the reference algorithm, OpenMP) on plies "
                      f"1..{args.cpu_plies} of {material}: {hs} half-steps over all 462 "
                      f"slices in {sec:.2f} s; smaller tables generated untimed",
        }
    table.close()
    reg.close()
    return out


def run_synth(args, tbg):
    """One king slice at real shape with seeded Bernoulli(0.03) win / loss
    exits for captures and king moves; one step = capture exits + init + the
    ply loop to the fixpoint."""
    material = args.material or "KQRBvKQRN"
    t0 = time.perf_counter()
    table = tbg.Table(material, synth_slice=args.slice, seed=args.seed)
    table.set_scatter(scatter_mode(args.scatter, tbg))
    alloc_s = time.perf_counter() - t0
    n = table.size
    print(f"synth {material} slice {args.slice}: {n} positions per side, "
          f"{table.info.device_bytes / 2**30:.1f} GiB resident, alloc {alloc_s:.1f} s",
          file=sys.stderr, flush=True)
    form = {"batched": tbg.FORM_BATCHED, "graph": tbg.FORM_GRAPH, "plain": tbg.FORM_PLAIN}[args.form]
    secs, lines = synth_steps(table, form, args.steps, args.warmup, tbg)
    tm = table.timing()
    steps = max(1, args.steps)
    half_steps = tm.half_steps / steps
    psp = half_steps * n / 2.0
    mean = sum(secs) / len(secs)
    kr = kernel_rooflines(tm, lines, n, steps, load_traffic(f"{material}_synth"), init_bits=10)
    dom = max(kr, key=lambda k: kr[k]["ms"])
    roof = {k: v for k, v in kr[dom].items() if k not in ("ms", "half_steps")}
    table.close()
    out = {
        "metric": METRIC, "value": psp / mean,
        "unit": "position-side-plies/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1000 * mean, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": f"{material} single king slice {args.slice} at real shape, "
                               f"seeded hash exits (seed {args.seed}), full fixpoint",
                   "material": material, "slice": args.slice, "positions_per_side": n,
                   "half_steps_per_step": half_steps, "plies": table.counts.nplies,
                   "form": args.form},
        "roofline": roof,
        "kernel_rooflines": kr,
        "kernels_ms_per_step": {k: getattr(tm, "ms_" + k) / steps for k in
                                ("pred", "verify", "sweep", "uncap", "init", "xr", "resolve")},
        "ms_per_ply": 1000 * mean / max(1, table.counts.nplies),
        "step_s": secs,
    }
    if not args.no_secondary:
        sec = run_table(args, tbg, steps=5, warmup=1, extras=False, material="KQRvKR")
        out["secondary"] = {k: sec[k] for k in ("value", "ms_per_step", "steps", "config",
                                                 "roofline", "kernels_ms_per_step")}
        print(f"secondary KQRvKR table: {sec['ms_per_step']:.1f} ms per table, "
              f"{sec['value']:.3e}/s", file=sys.stderr, flush=True)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_synth(args, tbg, material, n, psp)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--material", default="",
                    help="synth: KQRBvKQRN (default); table: KQRvKR (default)")
    ap.add_argument("--form", default="batched", choices=["batched", "graph", "plain"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-form comparison")
    ap.add_argument("--no-secondary", action="store_true",
                    help="synth: skip the secondary KQRvKR table line")
    ap.add_argument("--cpu-plies", type=int, default=2)
    ap.add_argument("--scatter", default="block+combine",
                    choices=["block", "global", "block+combine", "global+combine"],
                    help="scatter mode when not autotuned")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the built-in launch geometry (table and stream workloads)")
    ap.add_argument("--workload", default="synth", choices=["table", "synth", "stream", "served"],
                    help="synth (default): one king slice of --material with seeded hash "
                         "exits (KQRBvKQRN slice 170); table: a whole material (KQRvKR)")
    ap.add_argument("--slice", type=int, default=170)
    ap.add_argument("--jobs", type=int, default=1000, help="served: number of jobs, at most")
    ap.add_argument("--budget-gb", type=float, default=64.0, help="served: HBM admission budget")
    ap.add_argument("--max-active", type=int, default=32,
                    help="served: slice slots (also a capacity run at --residency-slots)")
    ap.add_argument("--residency-slots", type=int, default=1024,
                    help="served: slots of the extra capacity run that holds most job state "
                         "resident (0: none)")
    ap.add_argument("--max-slices", type=int, default=8, help="served: king slices per job, at most")
    ap.add_argument("--fill", type=float, default=0.7,
                    help="served: job state drawn up to this fraction of the budget")
    ap.add_argument("--served-pieces", default="6",
                    help="served: piece counts of the job materials, e.g. 5,6")
    ap.add_argument("--mixed-pieces", default="5,6",
                    help="served: piece counts of the mixed capacity comparison ('' to skip)")
    ap.add_argument("--stream-plain", action="store_true",
                    help="stream workload: also time the 4-piece level in the plain form")
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--geometry", default="",
                    help="launch geometry of the contexts created, e.g. "
                         "verify_unit_words=1024,pred_wg_per_cu=2 (fields of tbg_geometry; "
                         "the table workload's autotuner replaces it)")
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
    if args.geometry:
        g = tbg.default_geometry()
        for kv in args.geometry.split(","):
            k, v = kv.split("=")
            setattr(g, k.strip(), int(v))
        tbg.set_default_geometry(g)
    if args.workload == "synth":
        return run_synth(args, tbg)
    if args.workload == "stream":
        return run_stream(args, tbg)
    if args.workload == "served":
        return run_served(args, tbg)
    print(json.dumps(run_table(args, tbg, args.steps, args.warmup)), flush=True)


if __name__ == "__main__":
    main()
