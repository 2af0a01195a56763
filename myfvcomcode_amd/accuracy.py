"""Error of the mixed-precision form against the FP64 path.

    python -m myfvcomcode_amd.accuracy [--case C] [--hours H] [--nx NX --ny NY]

Runs the same seeded case with precision "fp64" (which reproduces the FP64
oracle bit for bit) and "mixed", and reports the RMS and maximum
differences of EL, UA/VA, U/V, T and S, the relative change of the FP64
volume and salt budgets of each run, and the time per internal step.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import DeviceModel, Model

FIELDS = ("el", "ua", "va", "u", "v", "t1", "s1")


def interior(name, a, m):
    if a.size in (m.m + 1, m.n + 1):
        return a[1:]
    return a.reshape(m.kb, -1)[:-1, 1:]


def run(m, precision, nsteps, log=None, chunk=500):
    dev = DeviceModel(m, precision=precision)
    d0 = dev.diagnostics()
    t0 = time.time()
    done = 0
    while done < nsteps:
        n = min(chunk, nsteps - done)
        dev.run(n, graph=True)
        dev.synchronize()
        done += n
        if log:
            log("%s: %d / %d steps, %.1f s" % (precision, done, nsteps, time.time() - t0))
    wall = time.time() - t0
    d1 = dev.diagnostics()
    if d1.nonfinite:
        raise RuntimeError("the %s run became non-finite" % precision)
    st = {f: interior(f, dev.get(f), m) for f in FIELDS}
    budget = {"volume_rel_change": (d1.volume - d0.volume) / d0.volume,
              "salt_rel_change": (d1.salt - d0.salt) / d0.salt}
    return st, wall, budget


def compare(case=0, hours=24.0, log=None, **overrides):
    m = Model.build(case, **overrides)
    dti = m.hdr.p.dte * m.hdr.p.isplit
    nsteps = int(round(hours * 3600.0 / dti))
    ref, w64, b64 = run(m, "fp64", nsteps, log)
    mix, wmx, bmx = run(m, "mixed", nsteps, log)
    out = {"case": case, "hours": hours, "steps": nsteps, "elements": m.n, "layers": m.kb - 1,
           "fp64_ms_per_step": 1000.0 * w64 / nsteps, "mixed_ms_per_step": 1000.0 * wmx / nsteps,
           "fp64_budgets": b64, "mixed_budgets": bmx, "rms": {}, "max": {}}
    for f in FIELDS:
        d = mix[f] - ref[f]
        out["rms"][f] = float(np.sqrt(np.mean(d * d)))
        out["max"][f] = float(np.abs(d).max())
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--case", type=int, default=0)
    ap.add_argument("--hours", type=float, default=24.0)
    ap.add_argument("--nx", type=int, default=None)
    ap.add_argument("--ny", type=int, default=None)
    args = ap.parse_args(argv)
    kw = {k: v for k, v in (("nx", args.nx), ("ny", args.ny)) if v}
    log = lambda s: print(s, file=sys.stderr, flush=True)
    print(json.dumps(compare(args.case, args.hours, log=log, **kw)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
