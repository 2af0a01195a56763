"""State snapshot and restart of the mode-split step (FVCOM's DUMP_NC_RST /
hot start, mod_ncdio.F:459-462, mod_ncdio.F:1416-1540, mod_startup.F:75-140)
as a NetCDF classic file (64-bit offset form).

    python -m myfvcomcode_amd.restart save  FILE --config C --steps S   run S steps, write FILE
    python -m myfvcomcode_amd.restart check FILE --config C --steps S   resume FILE, compare

The file holds two views of the state after internal step IINT:

* FVCOM's restart variables in FVCOM's layout, for its tools: `zeta`, `et`,
  `ua`, `va`, `u`, `v`, `ww`, `omega`, `temp`, `salinity`, `rho1`, `rmean1`,
  `tmean1`, `smean1`, `km`, `kh`, `kq`, `q2`, `q2l`, `l`, `viscofm`,
  `viscofh` on the dimensions `time`, `siglay`/`siglev`, `node`/`nele`
  (no dummy entry 0), with the grid `x`, `y`, `h`, `nv`, `siglay`,
  `siglev`, and `time` (days) and `iint`;
* every named state array of the device handle in full (dummy entry 0,
  all KB levels), as `state_<name>`: what a resumed run reads, so that it
  continues bit for bit as the run that wrote the file would have.

`save` takes any state source with `get(name)` (a DeviceModel or the
oracle); `load` returns IINT and the arrays and positions the model at
IINT (fv_params.iint0), after which a new DeviceModel or oracle receives the
arrays with `put`.
"""
import argparse
import json
import sys

import numpy as np
from scipy.io import netcdf_file

# the named state of the device handle (fvgpu_get_state / fvgpu_put_state)
STATE = ["el", "d", "et", "dt", "egf", "dtfa",
         "ua", "va", "el1", "d1", "et1", "dt1", "adx2d", "ady2d", "drx2d", "dry2d", "cbc", "wubot",
         "wvbot", "uard", "vard",
         "u", "v", "w", "ww", "uf", "vf", "drhox", "drhoy", "rho", "rmean", "t", "s", "km1", "viscofm",
         "t1", "s1", "tf1", "sf1", "rho1", "rmean1", "tmean1", "smean1", "q2", "q2l", "q2f", "q2lf", "l",
         "km", "kh", "kq", "wts", "viscofh", "elm1", "elm2"]

# FVCOM restart name, horizontal location, vertical coordinate (None: 2-D)
FVCOM = {"el": ("zeta", "node", None), "et": ("et", "node", None),
         "ua": ("ua", "nele", None), "va": ("va", "nele", None),
         "u": ("u", "nele", "siglay"), "v": ("v", "nele", "siglay"), "ww": ("ww", "nele", "siglay"),
         "wts": ("omega", "node", "siglev"),
         "t1": ("temp", "node", "siglay"), "s1": ("salinity", "node", "siglay"),
         "rho1": ("rho1", "node", "siglay"), "rmean1": ("rmean1", "node", "siglay"),
         "tmean1": ("tmean1", "node", "siglay"), "smean1": ("smean1", "node", "siglay"),
         "km": ("km", "node", "siglev"), "kh": ("kh", "node", "siglev"), "kq": ("kq", "node", "siglev"),
         "q2": ("q2", "node", "siglev"), "q2l": ("q2l", "node", "siglev"), "l": ("l", "node", "siglev"),
         "viscofm": ("viscofm", "nele", "siglay"), "viscofh": ("viscofh", "node", "siglay")}


def _extent_dim(model, size):
    m1, n1, kb = model.m + 1, model.n + 1, model.kb
    dims = {m1: "node0", n1: "nele0", m1 * kb: "node0_kb", n1 * kb: "nele0_kb"}
    if size not in dims:
        raise ValueError("array of %d values matches no extent of the model" % size)
    return dims[size]


def save(path, model, src, iint):
    """Write the state of `src` (get(name) -> float64 array) after internal
    step `iint` of `model`."""
    m, n, kb = model.m, model.n, model.kb
    kbm1 = kb - 1
    dti = model.hdr.p.dte * model.hdr.p.isplit
    f = netcdf_file(path, "w", version=2)
    try:
        f.createDimension("time", None)
        f.createDimension("node", m)
        f.createDimension("nele", n)
        f.createDimension("siglay", kbm1)
        f.createDimension("siglev", kb)
        f.createDimension("three", 3)
        for name, size in (("node0", m + 1), ("nele0", n + 1), ("node0_kb", (m + 1) * kb),
                           ("nele0_kb", (n + 1) * kb)):
            f.createDimension(name, size)
        f.title = "mode-split state after internal step %d" % iint
        f.source = "myfvcomcode_amd"
        # grid
        for name, field in (("x", "vx"), ("y", "vy"), ("h", "h")):
            v = f.createVariable(name, "d", ("node",))
            v[:] = model.field(field)[1:]
        nv = f.createVariable("nv", "i", ("three", "nele"))
        nv[:] = model.field("nv").reshape(3, n + 1)[:, 1:]
        z = model.field("z").reshape(kb, m + 1)[:, 1:]
        zz = model.field("zz").reshape(kb, m + 1)[:kbm1, 1:]
        f.createVariable("siglev", "d", ("siglev", "node"))[:] = z
        f.createVariable("siglay", "d", ("siglay", "node"))[:] = zz
        t = f.createVariable("time", "d", ("time",))
        t.units = "days since the model start"
        t[0] = (model.hdr.p.time0 + iint * dti) / 86400.0
        f.createVariable("iint", "i", ("time",))[0] = iint
        # the state
        for name in STATE:
            a = np.asarray(src.get(name), np.float64)
            v = f.createVariable("state_" + name, "d", ("time", _extent_dim(model, a.size)))
            v[0, :] = a
            if name in FVCOM:
                fname, loc, vert = FVCOM[name]
                ne = m if loc == "node" else n
                if vert is None:
                    fv = f.createVariable(fname, "d", ("time", loc))
                    fv[0, :] = a[1:ne + 1]
                else:
                    nl = kbm1 if vert == "siglay" else kb
                    fv = f.createVariable(fname, "d", ("time", vert, loc))
                    fv[0, :, :] = a.reshape(kb, ne + 1)[:nl, 1:]
    finally:
        f.close()


def load(path, model):
    """(iint, {name: array}) of a file written by `save` for a model of the
    same mesh; positions `model` at IINT for the handle created next."""
    f = netcdf_file(path, "r", mmap=False)
    try:
        dims = f.dimensions
        if (dims["node"], dims["nele"], dims["siglev"]) != (model.m, model.n, model.kb):
            raise ValueError("%s holds a %dx%dx%d state, the model is %dx%dx%d" % (
                path, dims["node"], dims["nele"], dims["siglev"], model.m, model.n, model.kb))
        iint = int(f.variables["iint"][0])
        arrays = {name: np.array(f.variables["state_" + name][0, :], np.float64) for name in STATE
                  if "state_" + name in f.variables}   # files written before ELM1/ELM2 lack them
    finally:
        f.close()
    model.hdr.p.iint0 = iint
    return iint, arrays


def put(dst, arrays):
    for name, a in arrays.items():
        dst.put(name, a)


def main(argv=None):
    from . import DeviceModel, Model
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["save", "check"])
    ap.add_argument("file")
    ap.add_argument("--config", type=int, default=0)
    ap.add_argument("--steps", type=int, default=10)
    args = ap.parse_args(argv)
    if args.cmd == "save":
        m = Model.build(args.config)
        dev = DeviceModel(m)
        dev.run(args.steps)
        save(args.file, m, dev, dev.diagnostics().iint)
        print(json.dumps({"file": args.file, "iint": int(dev.diagnostics().iint)}))
    else:
        # resume the file for `steps` more steps and compare with one run of
        # the same total length
        m = Model.build(args.config)
        iint, arrays = load(args.file, m)
        dev = DeviceModel(m)
        put(dev, arrays)
        dev.run(args.steps)
        ref = DeviceModel(Model.build(args.config))
        ref.run(iint + args.steps)
        diff = {nm: float(np.max(np.abs(dev.get(nm) - ref.get(nm)))) for nm in STATE}
        print(json.dumps({"iint": iint + args.steps, "max_abs_diff": max(diff.values())}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
