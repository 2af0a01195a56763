"""EnKF analysis of a batched ensemble on the device (ENKF_ASS_EVE,
mod_enkfassim.F:308-753, with ANALYSIS in mode 11, mod_enkfassim.F:2746-3010).

The members of a batched union (`Model.union`, identical meshes) are the
ensemble.  The analysis is done in two parts:

* In ensemble space, on the host (observations x members, small): the
  observed values H*A of every member, their anomalies S, perturbed
  observations D = d + eps - H*A (perturbations with their ensemble mean
  removed, mod_enkfassim.F:515-545), C = S*S' + (N-1)*R, its eigenvalues
  inverted until `truncation` of their sum is kept (`eigsign`), X3 = W
  diag(1/lambda) W' D (`genX3`) and X5 = I + S'*X3.
* In state space, on the device: A <- A*X5 for every assimilated array, over
  all members' slices of the union in place (`fvgpu_ensemble_update`).

Not restated: localization (ENKF_LOCALIZED), inflation, the square-root
modes (2x), and the reference's observation files and random number
generator (observations and a seed are passed in).
"""
import numpy as np

from . import gpu_lib, _check


def eig_inverse_truncated(eig, truncation):
    """`eigsign`: eigenvalues in increasing order; the largest are inverted
    until their running sum reaches `truncation` of the total, the rest set
    to zero."""
    eig = np.array(eig, np.float64)
    total = float(np.sum(eig))
    acc = 0.0
    out = np.zeros_like(eig)
    for i in range(len(eig) - 1, -1, -1):
        if acc / total < truncation:
            acc += eig[i]
            out[i] = 1.0 / eig[i]
        else:
            break
    return out


def analysis_x5(S, D, R, truncation=0.99):
    """X5 of the mode-11 analysis: S and D are nobs x nens, R nobs x nobs."""
    S = np.asarray(S, np.float64)
    D = np.asarray(D, np.float64)
    nobs, nens = S.shape
    if nobs == 1:
        e = 1.0 / (float(S[0] @ S[0]) + (nens - 1) * float(np.asarray(R).reshape(1)[0]))
        X3 = D * e
    else:
        C = S @ S.T + (nens - 1) * np.asarray(R, np.float64)
        lam, W = np.linalg.eigh(C)                  # increasing order, as eigC
        inv = eig_inverse_truncated(lam, truncation)
        X3 = W @ (inv[:, None] * (W.T @ D))         # genX3
    X5 = S.T @ X3
    X5[np.diag_indices(nens)] += 1.0
    return X5


class Ensemble:
    """The first `nens` instances of a union (identical meshes): member j owns
    nodes om_j+1..om_j+m and elements on_j+1..on_j+n."""

    def __init__(self, union, offs, nens):
        if nens < 2 or nens > 64 or nens > len(offs):
            raise ValueError("an ensemble needs 2..64 members of the union")
        m, n = offs[0][2], offs[0][3]
        if any(o[2] != m or o[3] != n for o in offs[:nens]):
            raise ValueError("ensemble members must share one mesh")
        self.union, self.nens, self.m, self.n = union, nens, m, n
        self.node_off = np.array([o[0] for o in offs[:nens]], np.int64)
        self.elem_off = np.array([o[1] for o in offs[:nens]], np.int64)

    def index(self, name, size, local, layer=None):
        """Union indices (one per member) of entity `local` (1-based) at
        `layer` (1-based, None for 2-D arrays) of array `name` of `size`."""
        m1, n1, kb = self.union.m + 1, self.union.n + 1, self.union.kb
        if size in (m1, m1 * kb):
            base, ls, lim = self.node_off, m1, self.m
        elif size in (n1, n1 * kb):
            base, ls, lim = self.elem_off, n1, self.n
        else:
            raise ValueError("%s: unknown extent" % name)
        if not 1 <= local <= lim:
            raise ValueError("%s: entity %d outside the member mesh" % (name, local))
        k = 0 if layer is None else layer - 1
        return k * ls + base + local

    def update(self, dev, name, X5, noalloc=False):
        """A <- A*X5 for array `name` over every member, in place on the
        device (fvgpu_ensemble_update, or fvgpu_ensemble_update_noalloc,
        which stages X5 in the library's fixed device array)."""
        L = gpu_lib()
        x5 = np.asfortranarray(X5, np.float64)
        fn = L.fvgpu_ensemble_update_noalloc if noalloc else L.fvgpu_ensemble_update
        _check(L, fn(dev._h, name.encode(), self.nens, self.node_off.ctypes.data, self.elem_off.ctypes.data,
                     self.m, self.n, x5.ravel(order="F").ctypes.data), "ensemble update " + name)


def assimilate(dev, ens, obs, variables, truncation=0.99, seed=0):
    """One analysis.  `obs` is a list of (array name, entity, layer or None,
    observed value, observation std); `variables` the arrays updated.
    Returns the RMS of observation minus ensemble mean before and after."""
    nobs = len(obs)
    cache = {}
    HA = np.zeros((nobs, ens.nens))
    idx = []
    for o, (name, local, layer, _, _) in enumerate(obs):
        if name not in cache:
            cache[name] = dev.get(name)
        ix = ens.index(name, cache[name].size, local, layer)
        idx.append((name, ix))
        HA[o] = cache[name][ix]
    d = np.array([ob[3] for ob in obs], np.float64)
    std = np.array([ob[4] for ob in obs], np.float64)
    S = HA - HA.mean(axis=1, keepdims=True)
    eps = np.random.default_rng(seed).standard_normal((nobs, ens.nens)) * std[:, None]
    eps -= eps.mean(axis=1, keepdims=True)
    D = d[:, None] + eps - HA
    X5 = analysis_x5(S, D, np.diag(std * std), truncation)
    for name in variables:
        ens.update(dev, name, X5)
    before = float(np.sqrt(np.mean((d - HA.mean(axis=1)) ** 2)))
    after_vals = np.array([dev.get(nm)[ix].mean() for nm, ix in idx])
    after = float(np.sqrt(np.mean((d - after_vals) ** 2)))
    return {"rms_before": before, "rms_after": after, "X5": X5}
