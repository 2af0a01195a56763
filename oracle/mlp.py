"""TEST INFRASTRUCTURE ONLY: FP64 numpy restatement of the learned vertical
closure -- the input features, the 5-64-64-2 MLP, its loss and gradient --
used to check the HIP training and inference kernels.  Only tests/ may use
it; the product never does.

Features follow the quantities VDIF_Q forms (FVCOM_source/vdif_q.F):
N^2 = -BOYGR (:256-268), S^2 from node-averaged U/V (:274-303), the surface
stress averaged over the node's elements (:172-190), z/h = Z(k), D.
"""
import numpy as np

H = 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * 5
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + 2 * H
NPARAM = OFF_B3 + 2
KEPS = 1e-7


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(H, 5), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3].reshape(2, H), p[OFF_B3:OFF_B3 + 2])


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def _softplus(z):
    return np.where(z > 20.0, z, np.log1p(np.exp(np.minimum(z, 20.0))))


def forward(p, x):
    """(KM, KH) predictions for samples x[n, 8] (the first 5 are inputs)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x5 = np.asarray(x, np.float64)[:, :5]
    z1 = x5 @ W1.T + b1
    h1 = z1 * _sig(z1)
    z2 = h1 @ W2.T + b2
    h2 = z2 * _sig(z2)
    o = h2 @ W3.T + b3
    return _softplus(o) + KEPS


def loss_grad(p, x, y):
    """Mean over samples of sum_c (ln K_c - ln max(y_c, eps))^2, and its gradient."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x5 = np.asarray(x, np.float64)[:, :5]
    y = np.maximum(np.asarray(y, np.float64), KEPS)
    n = len(x5)
    z1 = x5 @ W1.T + b1
    s1 = _sig(z1)
    h1 = z1 * s1
    z2 = h1 @ W2.T + b2
    s2 = _sig(z2)
    h2 = z2 * s2
    o = h2 @ W3.T + b3
    k = _softplus(o) + KEPS
    r = np.log(k) - np.log(y)
    loss = float(np.mean(np.sum(r * r, axis=1)))
    do = 2.0 * r / k * _sig(o) / n
    d2 = (do @ W3) * (s2 * (1.0 + z2 * (1.0 - s2)))
    d1 = (d2 @ W2) * (s1 * (1.0 + z1 * (1.0 - s1)))
    g = np.zeros(NPARAM)
    g[OFF_W1:OFF_B1] = (d1.T @ x5).ravel()
    g[OFF_B1:OFF_W2] = d1.sum(0)
    g[OFF_W2:OFF_B2] = (d2.T @ h1).ravel()
    g[OFF_B2:OFF_W3] = d2.sum(0)
    g[OFF_W3:OFF_B3] = (do.T @ h2).ravel()
    g[OFF_B3:] = do.sum(0)
    return loss, g


def wind_stress(model):
    """Element wind stress of the constant wind speed: the simple drag law
    of FVCOM_source/mod_bulk.F:58-90 (CD = 1.2e-3, rising above 11 m/s,
    capped at 25 m/s; air density 1.2)."""
    ux, uy = model.field("uuwind"), model.field("vvwind")
    w = np.sqrt(ux * ux + uy * uy)
    cd = np.where(w >= 25.0, (0.49 + 0.065 * 25.0) * 1e-3,
                  np.where(w >= 11.0, (0.49 + 0.065 * w) * 1e-3, float(np.float32(1.2e-3))))
    sx, sy = 1.2 * cd * ux * w, 1.2 * cd * uy * w
    sx[0] = sy[0] = 0.0
    return sx, sy


def features(model, get, ramp):
    """Normalised features x[n, 8] and MY2.5 targets y[n, 2] of a state.

    `get(name)` returns a state array (as Oracle.get / DeviceModel.get do),
    `ramp` is RAMP of the step the state belongs to; samples are ordered
    s = (k-2)*M + (i-1) for interfaces k = 2..KBM1."""
    M, kb = model.m, model.kb
    kbm1 = kb - 1
    nbve, off, ntve = model.field("nbve"), model.field("nbve_off"), model.field("ntve")
    d = get("d")
    rho1 = get("rho1").reshape(kb, M + 1)
    u = get("u").reshape(kb, -1)
    v = get("v").reshape(kb, -1)
    km = get("km").reshape(kb, M + 1)
    kh = get("kh").reshape(kb, M + 1)
    grav_n = model.field("grav_n")
    z = model.field("z").reshape(kb, M + 1)
    dzz = model.field("dzz").reshape(kb, M + 1)
    sx, sy = wind_stress(model) if model.hdr.p.wind_on else (np.zeros(model.n + 1), np.zeros(model.n + 1))
    x = np.zeros(((kbm1 - 1) * M, 8))
    y = np.zeros(((kbm1 - 1) * M, 2))
    for i in range(1, M + 1):
        ring = nbve[off[i]:off[i] + ntve[i]]
        nt = float(len(ring))
        wus = np.sum(-sx[ring] * ramp * 0.001) / nt
        wvs = np.sum(-sy[ring] * ramp * 0.001) / nt
        ustar = np.sqrt(np.sqrt(wus * wus + wvs * wvs))
        ubar = u[:, ring].sum(1) / nt
        vbar = v[:, ring].sum(1) / nt
        for k in range(2, kbm1 + 1):
            dzd = dzz[k - 2, i] * d[i]
            n2 = -grav_n[i] * (rho1[k - 2, i] - rho1[k - 1, i]) / dzd
            s2 = ((ubar[k - 1] - ubar[k - 2]) ** 2 + (vbar[k - 1] - vbar[k - 2]) ** 2) / (dzd * dzd)
            s = (k - 2) * M + (i - 1)
            x[s] = [np.arcsinh(n2 / 1e-5), np.log10(s2 + 1e-10) * 0.25 + 1.0, z[k - 1, i], ustar / 0.01,
                    np.log10(max(d[i], 1e-3)) - 1.0, 1.0, 0.0, 0.0]
            y[s] = [km[k - 1, i], kh[k - 1, i]]
    return x, y
