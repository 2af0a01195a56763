"""Stability trace of a run: Courant numbers and the state of the fastest
surface column, every so many internal steps, until the end or the first
non-finite value.

    python -m myfvcomcode_amd.stability [--config 1] [--hours 6] [--every 30]
        [--fine-after H --fine-every MIN] [--dte S] [--umol K] [--out FILE.json]

Per report (one JSON line each): the largest surface speed and where it is,
the internal step's horizontal advective Courant number there
(|u| DTI / the element's smallest altitude), the largest vertical Courant
number |omega| DTI / (D DZ) over interior interfaces, the external mode's
gravity-wave Courant number (sqrt(g D1) + |UA|) DTE / altitude, KM and Q2
of the two uppermost interior interfaces at the nodes of the fastest
element, and the minimum depth.
"""
import argparse
import json
import sys

import numpy as np


def _altitudes(m):
    n1 = m.n + 1
    x, y = m.field("vx"), m.field("vy")
    nv = m.field("nv").reshape(3, n1)
    xs, ys = x[nv], y[nv]
    side = np.stack([np.hypot(xs[(j + 1) % 3] - xs[j], ys[(j + 1) % 3] - ys[j]) for j in range(3)])
    alt = 2.0 * m.field("art") / np.maximum(side.max(axis=0), 1e-30)
    alt[0] = np.inf
    return alt, nv


def report(dev, m, alt, nv):
    """One line of statistics of the device state."""
    kb, m1, n1 = m.kb, m.m + 1, m.n + 1
    dti, dte = m.hdr.p.dte * m.hdr.p.isplit, m.hdr.p.dte
    d = dev.diagnostics()
    u = dev.get("u").reshape(kb, n1)
    v = dev.get("v").reshape(kb, n1)
    spd = np.hypot(u[0, 1:], v[0, 1:])
    ok = np.isfinite(spd)
    r = {"iint": int(d.iint), "hours": float(d.iint * dti / 3600.0), "nonfinite": int(d.nonfinite),
         "min_depth": float(d.min_depth)}
    if not ok.all():
        r["nonfinite_surface_elements"] = int((~ok).sum())
        return r
    e = int(np.argmax(spd)) + 1
    r.update({"max_surface_speed": float(spd[e - 1]), "element": e,
              "x": float(m.field("xc")[e]), "y": float(m.field("yc")[e]),
              "advective_courant": float(spd[e - 1] * dti / alt[e]),
              "speed_below": float(np.hypot(u[1, e], v[1, e]))})
    dd = dev.get("d")
    wts = dev.get("wts").reshape(kb, m1)
    dz = m.field("dz").reshape(kb, m1)
    vc = np.abs(wts[1:kb - 1, 1:]) * dti / (dd[None, 1:] * dz[:kb - 2, 1:])
    r["vertical_courant"] = float(vc.max())
    ua, va, d1 = dev.get("ua"), dev.get("va"), dev.get("d1")
    cfl = (np.sqrt(9.81 * np.maximum(d1[1:], 0.0)) + np.hypot(ua[1:], va[1:])) * dte / alt[1:]
    r["external_courant"] = float(cfl.max())
    km = dev.get("km").reshape(kb, m1)
    q2 = dev.get("q2").reshape(kb, m1)
    nodes = nv[:, e]
    r["km_interfaces_2_3"] = [float(km[1, nodes].mean()), float(km[2, nodes].mean())]
    r["q2_interfaces_2_3"] = [float(q2[1, nodes].mean()), float(q2[2, nodes].mean())]
    return r


def main(argv=None):
    import myfvcomcode_amd as F
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--hours", type=float, default=6.0)
    ap.add_argument("--every", type=float, default=30.0, help="minutes between reports")
    ap.add_argument("--fine-after", type=float, default=None, help="hours after which reports come every --fine-every")
    ap.add_argument("--fine-every", type=float, default=None, help="minutes between reports after --fine-after")
    ap.add_argument("--dte", type=float, default=None, help="external step (s); DTI = ISPLIT * DTE")
    ap.add_argument("--umol", type=float, default=None, help="background vertical viscosity (m^2/s)")
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    kw = {}
    if args.dte is not None:
        kw["p_dte"] = args.dte
    if args.umol is not None:
        kw["p_umol"] = args.umol
    m = F.Model.build(args.config, **kw)
    alt, nv = _altitudes(m)
    dti = m.hdr.p.dte * m.hdr.p.isplit
    dev = F.DeviceModel(m)
    nsteps = int(round(args.hours * 3600.0 / dti))
    chunk = max(1, int(round(args.every * 60.0 / dti)))
    rows, done = [], 0
    head = {"config": args.config, "elements": m.n, "layers": m.kb - 1, "dte": m.hdr.p.dte, "dti": dti,
            "umol": m.hdr.p.umol, "min_altitude_m": float(alt[1:].min())}
    print(json.dumps(head), flush=True)
    fine_from = nsteps if args.fine_after is None else int(round(args.fine_after * 3600.0 / dti))
    fine = chunk if args.fine_every is None else max(1, int(round(args.fine_every * 60.0 / dti)))
    while done < nsteps:
        step = fine if done >= fine_from else min(chunk, fine_from - done)
        s = min(step, nsteps - done)
        dev.run(s, graph=True)
        done += s
        r = report(dev, m, alt, nv)
        print(json.dumps(r), flush=True)
        rows.append(r)
        if r["nonfinite"]:
            break
    if args.out:
        json.dump({"run": head, "reports": rows}, open(args.out, "w"), indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
