"""Batched instances in one step: an ensemble and heterogeneous meshes
advanced together as the disjoint union of their meshes (fv_model_union),
one launch per kernel family for all instances, one captured hipGraph per
step, mixed precision, workgroup sizes from the autotuner.

    python -m myfvcomcode_amd.batched [--members 64] [--steps 5] [--precision mixed]

Workload (defaults): 64 ensemble members of the 501x251-node jittered
generator (250,000 elements x 20 layers, seeded wind and initial-elevation
perturbations) plus 8 instances of the same generator at 10k, 50k and 120k
elements.  Prints one JSON line: aggregate element-layer steps per second,
ms per step, and the check that one member of the batch advanced exactly as
it does alone.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import DeviceModel, Model

FULL_NODES = 501 * 251
HETERO = [(101, 51), (101, 51), (101, 51), (225, 113), (225, 113), (225, 113), (347, 174), (347, 174)]


def instance(nx, ny, seed):
    # seamounts scale with the domain area, so that smaller domains keep the
    # depth range of the full one
    ns = max(1, int(round(20 * nx * ny / FULL_NODES)))
    return Model.build(3, nx=nx, ny=ny, seed=seed, n_seamounts=ns)


def build(members=64, hetero=True, log=None):
    parts = []
    for j in range(members):
        parts.append(instance(501, 251, 1000 + j))
        if log and (j + 1) % 8 == 0:
            log("built %d members" % (j + 1))
    if hetero:
        for j, (nx, ny) in enumerate(HETERO):
            parts.append(instance(nx, ny, 5000 + j))
    union, offs = Model.union(parts)
    return parts, union, offs


def time_analysis(dev, union, offs, members, reps=3):
    """One EnKF ensemble update (A <- A*X5, fvgpu_ensemble_update) of every
    prognostic array of the members: average seconds and the bytes each
    value moves (read once, written once)."""
    from . import enkf
    ens = enkf.Ensemble(union, offs, members)
    X5 = np.eye(members) + 0.01 * np.random.default_rng(0).standard_normal((members, members))
    names = ("el", "ua", "va", "u", "v", "t1", "s1")
    nbytes = 0
    kb = union.kb
    for nm in names:
        per_layer = ens.m if nm in ("el", "t1", "s1") else ens.n
        nbytes += 2 * 8 * members * per_layer * (kb if nm in ("u", "v", "t1", "s1") else 1)
    dev.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        for nm in names:
            ens.update(dev, nm, X5)
    dev.synchronize()
    dt = (time.perf_counter() - t0) / reps
    return {"arrays": list(names), "seconds": dt, "bytes": nbytes, "GBps": nbytes / dt / 1e9}


def run(members=64, steps=5, precision="mixed", hetero=True, tune=True, log=None, analysis=False):
    t0 = time.time()
    parts, union, offs = build(members, hetero, log)
    if log:
        log("union: %d nodes, %d elements, KB=%d, built in %.0f s" % (union.m, union.n, union.kb, time.time() - t0))
    dev = DeviceModel(union, precision=precision)
    dev.run(1, graph=False)
    tuned = dev.autotune(1) if tune else {}
    dev.run(1, graph=True)                      # capture
    dev.synchronize()
    t0 = time.perf_counter()
    dev.run(steps, graph=True)
    dev.synchronize()
    wall = time.perf_counter() - t0
    d = dev.diagnostics()
    kbm1 = union.kb - 1
    out = {"instances": len(parts), "members": members, "elements": union.n, "layers": kbm1,
           "precision": precision, "steps": steps, "ms_per_step": 1000.0 * wall / steps,
           "elem_layer_steps_per_s": union.n * kbm1 * steps / wall, "nonfinite": int(d.nonfinite),
           "autotuned": {k: v[0] for k, v in tuned.items()}}
    # one member alone, same number of steps: bit-identical slice of the batch
    total = 2 + steps
    alone = DeviceModel(parts[0], precision=precision)
    alone.run(total, graph=True)
    om, on, m, n = offs[0]
    same = True
    for f in ("el", "ua", "va", "t1", "s1", "u", "km"):
        a, b = dev.get(f), alone.get(f)
        kb = union.kb
        if a.size == union.m + 1:
            a, b = a[om + 1: om + m + 1], b[1:]
        elif a.size == union.n + 1:
            a, b = a[on + 1: on + n + 1], b[1:]
        elif a.size == (union.m + 1) * kb:
            a, b = a.reshape(kb, -1)[:, om + 1: om + m + 1], b.reshape(kb, -1)[:, 1:]
        else:
            a, b = a.reshape(kb, -1)[:, on + 1: on + n + 1], b.reshape(kb, -1)[:, 1:]
        same = same and bool(np.array_equal(a, b))
    out["member0_matches_alone"] = same
    if analysis:
        out["enkf_update"] = time_analysis(dev, union, offs, members)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--no-hetero", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--analysis", action="store_true", help="also time one EnKF ensemble update of the members")
    args = ap.parse_args(argv)
    log = lambda s: print(s, file=sys.stderr, flush=True)
    print(json.dumps(run(args.members, args.steps, args.precision, not args.no_hetero, not args.no_tune, log,
                         args.analysis)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
