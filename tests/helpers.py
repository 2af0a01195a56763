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
