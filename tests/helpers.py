"""Shared helpers of the test suite (field lists, comparisons)."""
import numpy as np

NODE2 = ["el", "d", "et", "dt", "egf", "dtfa"]
ELEM2 = ["ua", "va", "el1", "d1", "et1", "dt1", "cbc", "wubot", "wvbot"]
ELEM3 = ["u", "v", "w", "ww", "rho", "t", "s", "km1", "drhox", "drhoy", "viscofm"]
NODE3 = ["t1", "s1", "rho1", "q2", "q2l", "l", "km", "kh", "kq", "wts", "viscofh"]


def interior(name, arr, model):
    """Drop the dummy entry 0 (and layer KB rows where unused) of an array."""
    kb = model.kb
    if name in NODE2:
        return arr[1:]
    if name in ELEM2:
        return arr[1:]
    if name in ELEM3:
        return arr.reshape(kb, -1)[:, 1:]
    return arr.reshape(kb, -1)[:, 1:]


def compare(dev, orc, model, names, tol, report=None):
    """Relative L2 differences of the named arrays; returns the worst."""
    worst = 0.0
    bad = []
    for n in names:
        a = interior(n, dev.get(n), model)
        b = interior(n, orc.get(n), model)
        nb = np.linalg.norm(b)
        err = float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))
        if nb == 0:
            err = float(np.abs(a - b).max())
        if report is not None:
            report[n] = err
        worst = max(worst, err)
        if not err <= tol:
            bad.append((n, err))
    return worst, bad


def wheel_mesh(n=12, kb=6, **pkw):
    """A closed basin whose centre node has n triangles around it (12 is the
    reference's limit): the node's ring holds 13 entries and its control
    volume 2n sub-edges, so the device takes its widest register lists
    (rings of 16, sub-edge lists of 24).  Two rings of nodes at 2 and 4 km,
    depth 15-20 m, a cosine tilt of the surface and a stratified T; keyword
    arguments override the case parameters."""
    import myfvcomcode_amd as F
    R = 2000.0
    pts = [(0.0, 0.0)]
    pts += [(R * np.cos(2 * np.pi * i / n), R * np.sin(2 * np.pi * i / n)) for i in range(n)]
    pts += [(2 * R * np.cos(np.pi * i / n), 2 * R * np.sin(np.pi * i / n)) for i in range(2 * n)]
    xy = np.array(pts)

    def r1(i):
        return 2 + i % n

    def r2(i):
        return 2 + n + i % (2 * n)

    tris = []
    for i in range(n):
        tris += [(1, r1(i), r1(i + 1)), (r1(i), r2(2 * i), r2(2 * i + 1)),
                 (r1(i), r2(2 * i + 1), r1(i + 1)), (r1(i + 1), r2(2 * i + 1), r2(2 * i + 2))]
    rows = []
    for t in tris:
        (x1, y1), (x2, y2), (x3, y3) = xy[t[0] - 1], xy[t[1] - 1], xy[t[2] - 1]
        ccw = (x2 - x1) * (y3 - y1) - (y2 - y1) * (x3 - x1) > 0
        rows.append((t[0], t[2], t[1]) if ccw else t)      # clockwise
    h = 15.0 + 5.0 * np.hypot(xy[:, 0], xy[:, 1]) / (2 * R)
    p = F.Model.case_defaults(0).p
    for k, v in pkw.items():
        setattr(p, k, v)
    m = F.Model.from_mesh(xy[:, 0], xy[:, 1], np.array(rows), h, kb, [], p)
    x = m.field("vx")
    m.field("el")[1:] = 0.05 * np.cos(np.pi * x[1:] / (4 * R))
    m.field3("t1")[:-1, 1:] = 15.0 + np.linspace(3.0, 0.0, kb - 1)[:, None]
    m.field3("s1")[:-1, 1:] = 30.0
    m.cold_start()
    return m
