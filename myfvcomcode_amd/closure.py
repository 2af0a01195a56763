"""The learned vertical closure: sample collection, training and the A/B
comparison with MY2.5.

    python -m myfvcomcode_amd.closure train  [--samples N] [--steps S] [--out FILE]
    python -m myfvcomcode_amd.closure judge  [--params FILE] [--case C] [--hours H] [--nx NX --ny NY]

`train` runs the plain MY2.5 model (the 81x21 estuary channel and the
jittered seamount mesh), dumps (features, KM, KH) samples from its own
states, splits them 90/10 with a fixed seed and trains the MLP with the
library's HIP forward/backward/AdamW kernels; `judge` runs the same case
with MY2.5, with the learned closure and with constant vertical mixing and
reports the RMS differences of elevation, surface T/S and currents against
the MY2.5 run.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import DeviceModel, Model, Trainer, gpu_lib

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_PARAMS = os.path.join(HERE, "data", "closure_mlp.npy")
H = 64
NPARAM = 5 * H + H + H * H + H + 2 * H + 2


def init_params(seed=0):
    """He-scaled random weights; the output bias starts at softplus^-1(1e-3)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    o = 0
    p[o:o + 5 * H] = rng.normal(0.0, np.sqrt(2.0 / 5), 5 * H); o += 5 * H
    o += H
    p[o:o + H * H] = rng.normal(0.0, np.sqrt(2.0 / H), H * H); o += H * H
    o += H
    p[o:o + 2 * H] = rng.normal(0.0, np.sqrt(1.0 / H), 2 * H); o += 2 * H
    p[o:o + 2] = np.log(np.expm1(1e-3))
    return p


def load_params(path=DEFAULT_PARAMS):
    p = np.load(path)
    if p.dtype != np.float32 or p.size != NPARAM:
        raise ValueError("%s is not a closure parameter vector" % path)
    return p


def collect(case, nsteps, every, precision="fp64", log=None, **overrides):
    """Run the plain model and dump samples every `every` internal steps."""
    m = Model.build(case, **overrides)
    dev = DeviceModel(m, precision=precision)
    xs, ys = [], []
    done = 0
    while done < nsteps:
        dev.run(every, graph=True)
        done += every
        x, y = dev.closure_samples()
        xs.append(x)
        ys.append(y)
        if log:
            log("case %d: %d steps, %d samples" % (case, done, sum(len(a) for a in xs)))
    return np.concatenate(xs), np.concatenate(ys)


def split(n, seed=0, frac=0.9):
    perm = np.random.default_rng(seed).permutation(n)
    return perm[:int(frac * n)], perm[int(frac * n):]


def val_loss(trainer, lo, n, mfma=False):
    """Mean loss of samples [lo, lo+n) with the current weights (FP32 or MFMA)."""
    y = trainer.y[lo:lo + n].astype(np.float64)
    k = trainer.infer(lo, n, mfma=mfma).astype(np.float64)
    r = np.log(k) - np.log(np.maximum(y, 1e-7))
    return float(np.mean(np.sum(r * r, axis=1)))


def train(x, y, steps=3000, batch=8192, lr=2e-3, wd=1e-4, seed=0, log=None):
    """Seeded 90/10 split; returns (params, report)."""
    tr, va = split(len(x), seed)
    order = np.concatenate([tr, va])
    t = Trainer(x[order], y[order], bmax=batch)
    t.set_params(init_params(seed))
    ntr = len(tr)
    rep = {"samples": int(len(x)), "train": int(ntr), "validation": int(len(va))}
    rep["val_loss_init"] = val_loss(t, ntr, len(va))
    t0 = time.time()
    chunk = max(1, steps // 10)
    done = 0
    while done < steps:
        n = min(chunk, steps - done)
        last = t.train(n, batch=batch, ntrain=ntr, lr=lr * (0.5 ** (3 * done / steps)), wd=wd, seed=seed + done)
        done += n
        if log:
            log("step %d  train loss %.4f" % (done, last))
    rep["train_seconds"] = time.time() - t0
    rep["train_loss_last_batch"] = last
    rep["val_loss"] = val_loss(t, ntr, len(va))
    rep["val_loss_mfma"] = val_loss(t, ntr, len(va), mfma=True)
    return t.params(), rep


def speed(st, kb):
    """Current speed at the element-layers 1..KBM1 (dummy entry 0 dropped)."""
    u = st["u"].reshape(kb, -1)[:-1, 1:]
    v = st["v"].reshape(kb, -1)[:-1, 1:]
    return np.hypot(u, v)


def judge(params, case=1, hours=6.0, log=None, **overrides):
    """A/B run: MY2.5 vs learned vs constant mixing, RMS differences vs MY2.5."""
    m = Model.build(case, **overrides)
    dti = m.hdr.p.dte * m.hdr.p.isplit
    nsteps = int(round(hours * 3600.0 / dti))
    runs = {}
    for name in ("my25", "learned", "constant"):
        mm = m
        if name == "constant":
            mm = Model.build(case, p_vertical_closure=0, **overrides)
        dev = DeviceModel(mm)
        if name == "learned":
            dev.set_closure(1, params)
        t0 = time.time()
        done = 0
        while done < nsteps:
            n = min(500, nsteps - done)
            dev.run(n, graph=True)
            dev.synchronize()
            done += n
            if log:
                log("%s: %d / %d steps, %.1f s" % (name, done, nsteps, time.time() - t0))
        wall = time.time() - t0
        kb = mm.kb
        if dev.diagnostics().nonfinite:
            raise RuntimeError("the %s run became non-finite within %.1f simulated hours" % (name, hours))
        st = {f: dev.get(f) for f in ("el", "ua", "va", "u", "v", "t1", "s1")}
        st["ts"] = st["t1"].reshape(kb, -1)[0, 1:]
        st["ss"] = st["s1"].reshape(kb, -1)[0, 1:]
        fb = dev.closure_fallbacks() if name == "learned" else (0, mm.m)
        runs[name] = (st, wall, fb)
    ref = runs["my25"][0]
    out = {"steps": nsteps, "hours": hours, "elements": m.n, "layers": m.kb - 1}
    for name in ("learned", "constant"):
        st, wall, fb = runs[name]
        kb = m.kb
        rms = lambda a, b: float(np.sqrt(np.mean((a - b) ** 2)))
        out[name] = {"el_rms_m": rms(st["el"][1:], ref["el"][1:]),
                     "surface_t_rms": rms(st["ts"], ref["ts"]), "surface_s_rms": rms(st["ss"], ref["ss"]),
                     "current_rms": rms(speed(st, kb), speed(ref, kb)),
                     "ms_per_step": 1000.0 * wall / nsteps, "fallback_columns_last_step": fb[0]}
    out["my25_ms_per_step"] = 1000.0 * runs["my25"][1] / nsteps
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train", "judge"])
    ap.add_argument("--samples", type=int, default=50_000_000)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=DEFAULT_PARAMS)
    ap.add_argument("--params", default=DEFAULT_PARAMS)
    ap.add_argument("--hours", type=float, default=24.0)
    ap.add_argument("--case", type=int, default=1)
    ap.add_argument("--nx", type=int, default=None)
    ap.add_argument("--ny", type=int, default=None)
    args = ap.parse_args(argv)
    gpu_lib()
    log = lambda s: print(s, file=sys.stderr, flush=True)
    if args.cmd == "train":
        # the estuary channel (every 18 steps of 1 day) and the seamount mesh
        xs, ys = [], []
        x, y = collect(0, 1728, 18, log=log)
        xs.append(x); ys.append(y)
        log("channel samples %d" % len(x))
        need = max(0, args.samples - len(x))
        per = 501501 * 39
        snaps = max(1, int(np.ceil(need / per)))
        x, y = collect(1, 60 * snaps, 60, log=log)
        xs.append(x); ys.append(y)
        log("seamount-mesh samples %d" % len(x))
        x, y = np.concatenate(xs), np.concatenate(ys)
        params, rep = train(x, y, steps=args.steps, batch=args.batch, seed=args.seed, log=log)
        os.makedirs(os.path.dirname(args.out), exist_ok=True)
        np.save(args.out, params)
        print(json.dumps(rep), flush=True)
    else:
        kw = {}
        if args.nx:
            kw["nx"] = args.nx
        if args.ny:
            kw["ny"] = args.ny
        print(json.dumps(judge(load_params(args.params), case=args.case, hours=args.hours, log=log, **kw)),
              flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
