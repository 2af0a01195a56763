"""Text summaries of rocprofv3 databases (the rocpd SQLite files that
`rocprofv3 --kernel-trace` and `rocprofv3 --pmc` write).

    python -m myfvcomcode_amd.profsum DB              per-kernel calls / total / average (us)
    python -m myfvcomcode_amd.profsum DB --pmc        per-kernel average of each counter

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


def main(argv=None):
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
