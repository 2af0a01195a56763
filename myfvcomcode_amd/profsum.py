"""Text summaries of rocprofv3 databases (the rocpd SQLite files that
`rocprofv3 --kernel-trace` and `rocprofv3 --pmc` write).

    python -m myfvcomcode_amd.profsum DB              per-kernel calls / total / average (us)
    python -m myfvcomcode_amd.profsum DB --pmc        per-kernel average of each counter
    python -m myfvcomcode_amd.profsum --traffic STEP_FETCH STEP_WRITE COPY_FETCH COPY_WRITE OUT.json
        HBM bytes per launch of each kernel family of the 1001x501x41 step:
        FETCH_SIZE and WRITE_SIZE (KiB per dispatch) scaled by what they read
        for the 2^25-double device copy of `bench.py --copy-calibration`,
        beside the algorithmic bytes of roofline.py; prints the table

Kernel names are shortened to the function name with its template
arguments.  Only the standard library is used.
"""
import argparse
import re
import sqlite3
import sys


def short(name):
    m = re.match(r"(?:void )?([A-Za-z_:()0-9 ]*?[A-Za-z_0-9]+(?:<[^()]*>)?)\(", name)
    s = m.group(1) if m else name
    return s.replace("(anonymous namespace)::", "")


def kernel_stats(db):
    c = sqlite3.connect(db)
    rows = c.execute("select name, count(*), sum(duration), avg(duration) from kernels group by name").fetchall()
    return sorted(((short(n), cnt, tot / 1000.0, avg / 1000.0) for n, cnt, tot, avg in rows),
                  key=lambda r: -r[2])


def pmc_stats(db):
    c = sqlite3.connect(db)
    rows = c.execute("select kernel_name, counter_name, count(*), avg(value) from counters_collection "
                     "group by kernel_name, counter_name").fetchall()
    return sorted((short(k), cn, cnt, v) for k, cn, cnt, v in rows)


# kernel (short name) -> family of the step's timing events
FAMILY_OF = [
    (r"k_advave_edge<", "advave_edge"), (r"k_ext_elevation\b", "ext_elevation"),
    (r"k_ext_momentum<", "ext_momentum"), (r"k_cvflux_ts<", "cvflux_ts"), (r"k_cvflux_q<", "cvflux_q"),
    (r"k_cvrec<0,", "cvrec_ts"), (r"k_cvrec<1,", "cvrec_q"), (r"k_adv_ts_gather<", "adv_ts"),
    (r"k_adv_q_gather<", "adv_q"), (r"k_vertvl_tile<", "vertvl"),
    (r"k_viscofh_r<", "viscofh"), (r"k_vdif_q_r<", "vdif_q"), (r"k_vdif_uv_r<", "vdif_uv"),
    (r"k_vdif_ts_r<", "vdif_ts"), (r"k_n2e3d\b", "n2e3d"),
    (r"k_cvrecflux<0,", "cvrecflux_ts"), (r"k_cvrecflux<1,", "cvrecflux_q"),
    (r"k_uv_blk<0,", "adv3d_blk"), (r"k_uv_blk<1,", "advuv_blk"),
]


def family(kernel):
    for pat, fam in FAMILY_OF:
        if re.match(pat, kernel):
            return fam
    return None


def counter_avg(db, counter):
    return {k: v for k, cn, cnt, v in pmc_stats(db) if cn == counter}


def traffic(step_fetch, step_write, copy_fetch, copy_write, out_json, nx=1001, ny=501, kb=41):
    import json
    from . import Model
    from . import roofline
    n = 1 << 25
    cf = counter_avg(copy_fetch, "FETCH_SIZE").get("k_copy")
    cw = counter_avg(copy_write, "WRITE_SIZE").get("k_copy")
    if not cf or not cw:
        raise SystemExit("no k_copy dispatch in the calibration databases")
    kib = 8.0 * n / 1024.0
    fs, ws = kib / cf, kib / cw
    fetch, write = counter_avg(step_fetch, "FETCH_SIZE"), counter_avg(step_write, "WRITE_SIZE")
    model = Model.build(1, nx=nx, ny=ny, kb=kb)
    fams = {}
    for k in fetch:
        f = family(k)
        if f is None or k not in write:
            continue
        b = (fetch[k] * fs + write[k] * ws) * 1024.0
        alg = roofline.bytes_per_launch(f, model)
        fams[f] = {"kernel": k, "fetch_kib": fetch[k], "write_kib": write[k], "bytes_per_launch": b,
                   "algorithmic_bytes_per_launch": alg, "traffic_over_algorithmic": b / alg}
    doc = {"workload": "%dx%d-node jittered mesh, %d elements x %d layers, FP64" % (nx, ny, model.n, kb - 1),
           "calibration": {"copy_doubles": n, "copy_fetch_kib": cf, "copy_write_kib": cw,
                           "fetch_scale": fs, "write_scale": ws},
           "families": fams}
    with open(out_json, "w") as fh:
        json.dump(doc, fh, indent=1)
    print("Calibration: `k_copy` of 2^25 doubles (%.0f KiB read and written) reads FETCH_SIZE = %.0f KiB "
          "and WRITE_SIZE = %.0f KiB per dispatch: FETCH_SIZE is scaled by %.3f, WRITE_SIZE by %.3f.\n"
          % (kib, cf, cw, fs, ws))
    print("| family | kernel | FETCH (KiB) | WRITE (KiB) | traffic (MB) | algorithmic (MB) | ratio |")
    print("|---|---|---:|---:|---:|---:|---:|")
    order = [f for _, f in FAMILY_OF]
    for f in sorted(fams, key=order.index):
        r = fams[f]
        print("| %s | `%s` | %.0f | %.0f | %.1f | %.1f | %.2f |" % (
            f, r["kernel"], r["fetch_kib"], r["write_kib"], r["bytes_per_launch"] / 1e6,
            r["algorithmic_bytes_per_launch"] / 1e6, r["traffic_over_algorithmic"]))


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if argv and argv[0] == "--traffic":
        if len(argv) != 6:
            raise SystemExit("--traffic STEP_FETCH STEP_WRITE COPY_FETCH COPY_WRITE OUT.json")
        traffic(*argv[1:])
        return 0
    ap = argparse.ArgumentParser()
    ap.add_argument("db")
    ap.add_argument("--pmc", action="store_true")
    args = ap.parse_args(argv)
    if args.pmc:
        print("| kernel | counter | dispatches | average value |")
        print("|---|---|---:|---:|")
        for k, cn, cnt, v in pmc_stats(args.db):
            print("| `%s` | %s | %d | %.1f |" % (k, cn, cnt, v))
    else:
        rows = kernel_stats(args.db)
        total = sum(r[2] for r in rows)
        print("| kernel | calls | total (us) | average (us) | share |")
        print("|---|---:|---:|---:|---:|")
        for k, cnt, tot, avg in rows:
            print("| `%s` | %d | %.1f | %.2f | %.1f%% |" % (k, cnt, tot, avg, 100.0 * tot / total))
    return 0


if __name__ == "__main__":
    sys.exit(main())
