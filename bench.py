"""Benchmark of the MI355X mode-split time step (one GPU).

Workload: the 1001x501-node jittered mesh (1,000,000 elements x 40 sigma
layers, KB=41), depth 5+195*(x/Lx)^1.5 with 20 seeded seamounts, M2/S2/K1
tide on the east boundary, a seeded Gaussian random wind field held
constant, MY2.5 closure, DTE = 2 s, ISPLIT = 10 (DTI = 20 s), synthetic
seeded data.

A step is one INTERNAL_STEP (ISPLIT external steps included).  An untimed
survey pass (eager launches, every chain on one stream) with HIP events
around every launch gives each kernel family's own average device time
("roofline_by_kernel") and the dominant family; the timed region then runs
K steps launched eagerly with the independent chains side by side on the
handle's two streams, the state resident in HBM, and HIP events around
one in four of the dominant family's launches (events around every launch
cost about 0.46 ms of a 23.3 ms step), whose average device time is the
"roofline" figure.  The same step replayed as one hipGraph per step is
timed among the variants.  The CPU baseline is the serial FP64 oracle on a
bounded sample of the same workload (the same generator at 251x126 nodes,
62,500 elements x 40 layers), on one core and as one rank per host core (at
most 16), scaled to the metric's unit.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "sim-days/wall-hour @1M elem x 40 layers; elem-layer-steps/s vs plain FP64"
PEAK_HBM = 8.0e12
VARIANTS = ("fp64_graph", "mixed", "mixed_learned_closure")
PROFILE_EVERY = 4   # the timed region's events time one in this many dominant-family launches


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nx", type=int, default=1001)
    ap.add_argument("--ny", type=int, default=501)
    ap.add_argument("--kb", type=int, default=41)
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip timing the mixed-precision and learned-closure forms")
    ap.add_argument("--variant", choices=VARIANTS,
                    help="only time this form (graph replays) and print its line; for profiling one form")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only time a plain device copy of 2^25 doubles (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: runs on one GPU only; the time step would shard by domain "
              "decomposition with halo exchange, which is not provided here", file=sys.stderr)
        sys.exit(2)

    import myfvcomcode_amd as F
    from myfvcomcode_amd import roofline

    if args.copy_calibration:
        dev = F.DeviceModel(F.Model.build(0))
        n = 1 << 25
        us = dev.stream_copy(n, 5)
        print(json.dumps({"copy_doubles": n, "bytes_read": 8 * n, "bytes_written": 8 * n, "avg_us": us,
                          "GBps": 16.0 * n / us / 1e3}), flush=True)
        return

    t0 = time.time()
    model = F.Model.build(1, nx=args.nx, ny=args.ny, kb=args.kb)
    n, kbm1 = model.n, model.kb - 1
    dti = model.hdr.p.dte * model.hdr.p.isplit
    print("bench: built %d nodes / %d elements x %d layers in %.1f s" % (model.m, n, kbm1, time.time() - t0),
          file=sys.stderr, flush=True)
    if args.variant:
        v = time_variants(F, model, args.steps, args.warmup, None, dti, n * kbm1, only=args.variant)
        print(json.dumps({"variant": args.variant, "steps": args.steps, "warmup": args.warmup,
                          "workload": "jittered mesh %dx%d nodes, %d elements x %d layers" % (args.nx, args.ny, n, kbm1),
                          **v[args.variant]}), flush=True)
        return
    dev = F.DeviceModel(model)
    dev.run(args.warmup, graph=False)
    dev.synchronize()
    print("bench: warmup done", file=sys.stderr, flush=True)

    # The survey pass (not timed) runs every chain on the handle's main
    # stream (concurrency 0), eagerly, with HIP events around every launch,
    # so that no family's event time includes another kernel running beside
    # it: each family's average device time and the dominant family.  The
    # timed region then launches eagerly with the chains side by side, with
    # events around the dominant family's launches only -- events around all
    # ~300 launches of a step would add about 5 % to it.  The dominant family
    # (the 2-D momentum stage) runs on the main stream with nothing beside
    # it, so its event time is its own.
    dev.set_concurrency(0)
    survey_steps = max(1, min(10, args.steps))
    dev.synchronize()
    dev.set_profiling(True)
    dev.run(survey_steps, graph=False)
    dev.synchronize()
    dev.set_profiling(False)
    stats = dev.kernel_stats()

    per_kernel = {}
    tot_b = tot_t = 0.0
    for name, (avg_us, count) in stats.items():
        if name not in roofline.FAMILIES or avg_us <= 0:
            continue
        b = roofline.bytes_per_launch(name, model)
        gbps = b / (avg_us * 1e-6) / 1e9
        per_kernel[name] = {"avg_us": round(avg_us, 2), "launches_per_step": count / survey_steps,
                            "bytes": b, "achieved_GBps": round(gbps, 1), "frac": round(gbps * 1e9 / PEAK_HBM, 4)}
        tot_b += b * count / survey_steps
        tot_t += avg_us * 1e-6 * count / survey_steps
    # the families without a byte model (small, pointwise or boundary work)
    unmodelled = {name: {"avg_us": round(avg_us, 2), "launches_per_step": count / survey_steps}
                  for name, (avg_us, count) in sorted(stats.items())
                  if name not in roofline.FAMILIES and count > 0}
    # dominant kernel = the modelled family with the largest device time per step
    dom = max(per_kernel, key=lambda k: per_kernel[k]["avg_us"] * per_kernel[k]["launches_per_step"])

    # timed region: K internal steps, launched eagerly with the independent
    # chains side by side on the handle's two streams, events around one in
    # PROFILE_EVERY launches of the dominant family (HIP events cannot time
    # launches inside a replayed graph here; the graph replay of the same
    # step is timed as a variant)
    dev.set_concurrency(3)
    dev.set_profiling(True, kernel=dom, every=PROFILE_EVERY)
    dev.synchronize()
    t0 = time.perf_counter()
    dev.run(args.steps, graph=False)
    dev.synchronize()
    wall = time.perf_counter() - t0
    dev.set_profiling(False)
    dom_us, dom_count = dev.kernel_stats()[dom]
    diag = dev.diagnostics()
    if diag.nonfinite:
        print("bench: non-finite state", file=sys.stderr)
        sys.exit(1)

    rate = n * kbm1 * args.steps / wall                      # element-layer-steps/s
    sdph = rate * dti / (24.0 * n * kbm1)                     # simulated days per wall-hour

    b = roofline.bytes_per_launch(dom, model)
    gbps = b / (dom_us * 1e-6) / 1e9
    pk = {"avg_us": round(dom_us, 2), "launches_per_step": per_kernel[dom]["launches_per_step"], "bytes": b,
          "achieved_GBps": round(gbps, 1), "frac": round(gbps * 1e9 / PEAK_HBM, 4),
          "timed_launches": dom_count}
    # HBM bytes per launch from the calibrated FETCH_SIZE/WRITE_SIZE runs of
    # this workload (rocprofv3 --pmc), when the default workload is timed
    traffic_tab = {}
    tfile = os.path.join(ROOT, "profiles", "hbm_traffic_1001x501x41_fp64.json")
    if os.path.exists(tfile) and (args.nx, args.ny, args.kb) == (1001, 501, 41):
        try:
            traffic_tab = json.load(open(tfile)).get("families", {})
        except (OSError, ValueError):
            traffic_tab = {}
    for name, rec in per_kernel.items():
        if name in traffic_tab:
            rec["traffic_bytes"] = traffic_tab[name]["bytes_per_launch"]
    traffic = traffic_tab.get(dom, {}).get("bytes_per_launch")
    step_kernel_ms = sum(v[0] * v[1] for v in stats.values()) / 1000.0 / survey_steps

    variants = None
    if not args.no_variants:
        del dev
        variants = time_variants(F, model, args.steps, args.warmup, wall / args.steps, dti, n * kbm1)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_steps, args.kb)

    out = {
        "metric": METRIC,
        "value": sdph,
        "unit": "sim-days/wall-hour",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded jittered mesh, seamounts, M2/S2/K1 tide, Gaussian random wind)",
        "config": {"workload": "FVCOM INTERNAL_STEP, %dx%d-node jittered mesh, %d elements x %d layers, "
                               "DTE=2 s ISPLIT=10, MY2.5, plain FP64" % (args.nx, args.ny, n, kbm1),
                   "elements": n, "layers": kbm1, "dti_s": dti,
                   "mode": "eager launches, independent chains on two streams"},
        "elem_layer_steps_per_s": rate,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": pk["achieved_GBps"], "peak": PEAK_HBM / 1e9,
                     "unit": "GB/s", "frac": pk["frac"], "traffic": traffic,
                     "algorithmic_bytes_per_launch": pk["bytes"], "avg_launch_us": pk["avg_us"],
                     "launches_per_step": pk["launches_per_step"], "timed_launches": pk["timed_launches"],
                     "timed": "HIP events around 1 in %d launches over the timed region" % PROFILE_EVERY},
        "roofline_step": {"modelled_bytes_per_step": tot_b,
                          "achieved_GBps": round(tot_b / tot_t / 1e9, 1) if tot_t else None,
                          "frac": round(tot_b / tot_t / PEAK_HBM, 4) if tot_t else None,
                          "kernel_ms_per_step": round(step_kernel_ms, 3)},
        "roofline_by_kernel": per_kernel,
        "roofline_by_kernel_from": "untimed survey pass of %d steps, HIP events around every launch" % survey_steps,
        "other_kernels": unmodelled,
        "variants": variants,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def time_variants(F, model, steps, warmup, fp64_step_s, dti, units, only=None):
    """FP64 as one hipGraph replay per step, the mixed-precision form, and
    mixed precision with the learned closure (graph replays, chains side by
    side), on the same workload and step count (``only``: that form alone)."""
    from myfvcomcode_amd.closure import load_params
    out = {}
    for name, prec, closure in (("fp64_graph", "fp64", False), ("mixed", "mixed", False),
                                ("mixed_learned_closure", "mixed", True)):
        if only is not None and name != only:
            continue
        dev = F.DeviceModel(model, precision=prec)
        graph = True
        dev.set_concurrency(3)
        if closure:
            dev.set_closure(1, load_params())
        dev.run(warmup, graph=graph)
        dev.synchronize()
        t0 = time.perf_counter()
        dev.run(steps, graph=graph)
        dev.synchronize()
        w = (time.perf_counter() - t0) / steps
        if dev.diagnostics().nonfinite:
            raise RuntimeError("bench: the %s form became non-finite" % name)
        out[name] = {"ms_per_step": 1000.0 * w, "value": units / w * dti / (24.0 * units)}
        if fp64_step_s is not None:
            out[name]["speedup_vs_fp64"] = fp64_step_s / w
        if closure:
            out[name]["fallback_columns_last_step"] = dev.closure_fallbacks()[0]
        print("bench: %s %.2f ms/step" % (name, 1000.0 * w), file=sys.stderr, flush=True)
        del dev
    return out


def _cpu_ranks(nranks, steps, nx, ny, kb):
    """Start `nranks` oracle ranks (oracle/cpu_rank.py) as child processes,
    release them together once all have built their sample, and return
    their timings."""
    import subprocess
    script = os.path.join(ROOT, "oracle", "cpu_rank.py")
    env = dict(os.environ, OMP_NUM_THREADS="1")
    procs = [subprocess.Popen([sys.executable, script, str(nx), str(ny), str(kb), str(steps)],
                              stdin=subprocess.PIPE, stdout=subprocess.PIPE, text=True, env=env)
             for _ in range(nranks)]
    for p in procs:
        if p.stdout.readline().strip() != "ready":
            raise RuntimeError("bench: a CPU baseline rank failed to start")
    for p in procs:
        p.stdin.write("go\n")
        p.stdin.flush()
    out = []
    for p in procs:
        line = p.stdout.readline()
        if p.wait(timeout=600) != 0 or not line:
            raise RuntimeError("bench: a CPU baseline rank failed")
        out.append(json.loads(line))
    return out


def cpu_baseline(steps, kb, nx=251, ny=126):
    """The FP64 oracle (the serial C restatement, built here with -O3
    -march=native) on the 251x126-node sample of the same generator
    (62,500 elements, 1/16 of the 1M-element mesh): one core, median of 3
    runs; and one rank per host core (at most 16, the host share of one
    GPU) each stepping its own sample at once, the stand-in for the
    reference's MPI decomposition of the 1M-element mesh into 62,500-element
    subdomains (no halo exchange, which favours the CPU).  Rates are scaled
    to the metric's 1M elements x 40 layers."""
    import oracle
    import statistics
    oracle.build(native=True)
    single = [_cpu_ranks(1, steps, nx, ny, kb)[0] for _ in range(3)]
    walls = sorted(r["wall_s"] for r in single)
    r0 = single[0]
    units = r0["elements"] * r0["layers"] * r0["steps"]
    dti = r0["dti"]
    rate1 = units / statistics.median(walls)
    ncores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    nranks = max(1, min(16, ncores))
    many = _cpu_ranks(nranks, steps, nx, ny, kb)
    rate_all = sum(units / r["wall_s"] for r in many)

    def sdph(rate):
        return rate * dti / (24.0 * 1000000 * 40)

    return {"value": sdph(rate_all), "unit": "sim-days/wall-hour (scaled to 1M x 40)", "cores": nranks,
            "kind": "port", "elem_layer_steps_per_s": rate_all,
            "single_core": {"value": sdph(rate1), "elem_layer_steps_per_s": rate1, "cores": 1,
                            "wall_s_runs": walls},
            "sample": "oracle (serial FP64 C restatement, -O3 -march=native, built on this host) on the same "
                      "jittered-mesh generator at %dx%d nodes (%d elements x %d layers), %d internal steps after "
                      "one untimed step; one core: median of 3 runs; all cores: %d ranks each stepping its own "
                      "sample at once (stand-in for the reference's MPI domain decomposition, halo exchange "
                      "not included)" % (nx, ny, r0["elements"], r0["layers"], steps, nranks)}


if __name__ == "__main__":
    main()
