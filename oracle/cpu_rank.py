"""TEST INFRASTRUCTURE ONLY: one rank of the CPU baseline's timing.

Run as a child process by bench.py's cpu_baseline leg: builds one sample of
the benchmark's workload (the jittered seamount generator at NX x NY nodes
and KB levels), steps it once with the host-tuned oracle build, prints
"ready", waits for a line on stdin, then times STEPS internal steps and
prints one JSON line.  Several ranks started together stand in for the
reference's MPI domain decomposition (each rank steps its own subdomain;
the halo exchanges are left out, which favours the CPU).
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def main():
    nx, ny, kb, steps = (int(a) for a in sys.argv[1:5])
    import myfvcomcode_amd as F
    from oracle import Oracle
    m = F.Model.build(1, nx=nx, ny=ny, kb=kb)
    o = Oracle(m, native=True)
    o.internal_step()                  # first step outside the clock (page-in)
    print("ready", flush=True)
    sys.stdin.readline()
    t0 = time.perf_counter()
    o.run(steps)
    wall = time.perf_counter() - t0
    print(json.dumps({"wall_s": wall, "elements": m.n, "layers": m.kb - 1, "steps": steps,
                      "dti": m.hdr.p.dte * m.hdr.p.isplit}), flush=True)


if __name__ == "__main__":
    main()
