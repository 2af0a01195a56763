"""EnKF analysis of a batched ensemble: the ensemble-space algebra of mode
11 against the Kalman update with the ensemble covariance (a closed form),
its limits, and the device's in-place A*X5 over the union's members."""
import numpy as np
import pytest

import myfvcomcode_amd as F
from myfvcomcode_amd import enkf


def _kalman(A, H, R, D):
    """A + P H' (H P H' + R)^-1 D with P the ensemble covariance."""
    Ap = A - A.mean(axis=1, keepdims=True)
    P = Ap @ Ap.T / (A.shape[1] - 1)
    K = P @ H.T @ np.linalg.inv(H @ P @ H.T + R)
    return A + K @ D


@pytest.mark.parametrize("nobs", [1, 6])
def test_x5_equals_kalman_update(nobs):
    rng = np.random.default_rng(3)
    nd, ne = 40, 12
    A = rng.standard_normal((nd, 1)) + rng.standard_normal((nd, ne)) @ np.diag(rng.uniform(0.5, 2, ne))
    sel = rng.choice(nd, nobs, replace=False)
    H = np.zeros((nobs, nd))
    H[np.arange(nobs), sel] = 1.0
    R = np.diag(rng.uniform(0.1, 0.5, nobs))
    d = rng.standard_normal(nobs)
    eps = rng.standard_normal((nobs, ne))
    eps -= eps.mean(axis=1, keepdims=True)
    D = d[:, None] + eps - H @ A
    S = H @ (A - A.mean(axis=1, keepdims=True))
    X5 = enkf.analysis_x5(S, D, R, truncation=1.0)
    np.testing.assert_allclose(A @ X5, _kalman(A, H, R, D), rtol=1e-10, atol=1e-10)


def test_huge_observation_error_leaves_ensemble():
    rng = np.random.default_rng(4)
    S = rng.standard_normal((5, 8))
    S -= S.mean(axis=1, keepdims=True)
    X5 = enkf.analysis_x5(S, rng.standard_normal((5, 8)), 1e14 * np.eye(5), truncation=1.0)
    np.testing.assert_allclose(X5, np.eye(8), atol=1e-9)


def test_truncation_keeps_the_dominant_eigenvalues():
    np.testing.assert_allclose(enkf.eig_inverse_truncated([1.0, 2.0, 100.0], 0.9), [0.0, 0.0, 0.01])
    np.testing.assert_allclose(enkf.eig_inverse_truncated([1.0, 2.0, 4.0], 1.0), [1.0, 0.5, 0.25])
    # 4 of 7 reached: the third largest is still taken, the rest dropped
    np.testing.assert_allclose(enkf.eig_inverse_truncated([1.0, 2.0, 4.0], 0.6), [0.0, 0.5, 0.25])


def _ensemble(nens, **kw):
    parts = [F.Model.build(0, nx=21, ny=7, kb=6, seed=100 + j, **kw) for j in range(nens)]
    u, offs = F.Model.union(parts)
    return parts, u, offs


@pytest.mark.gpu
def test_device_update_is_the_plain_product(built):
    """fvgpu_ensemble_update equals sum_i A(:,i)*X5(i,j) in increasing i, bit
    for bit, for node and element, 2-D and 3-D arrays, and leaves entity 0."""
    nens = 5
    parts, u, offs = _ensemble(nens)
    dev = F.DeviceModel(u)
    print("case: 5 channel members 21x7 KB=6, random X5, update el/ua/t1/u", flush=True)
    dev.run(2, graph=False)
    ens = enkf.Ensemble(u, offs, nens)
    X5 = np.eye(nens) + 0.3 * np.random.default_rng(9).standard_normal((nens, nens))
    for name in ("el", "ua", "t1", "u"):
        before = dev.get(name)
        ens.update(dev, name, X5)
        after = dev.get(name)
        m1, n1, kb = u.m + 1, u.n + 1, u.kb
        node = before.size in (m1, m1 * kb)
        ls, off, ln = (m1, ens.node_off, ens.m) if node else (n1, ens.elem_off, ens.n)
        nl = before.size // ls
        expect = before.copy()
        for k in range(nl):
            A = np.stack([before[k * ls + off[i] + 1: k * ls + off[i] + 1 + ln] for i in range(nens)], axis=1)
            for j in range(nens):
                acc = np.zeros(ln)
                for i in range(nens):
                    acc = acc + A[:, i] * X5[i, j]
                expect[k * ls + off[j] + 1: k * ls + off[j] + 1 + ln] = acc
        assert np.array_equal(after, expect), name
        assert after[0] == before[0]


@pytest.mark.gpu
def test_assimilation_pulls_the_mean_to_the_observations(built):
    nens = 8
    parts, u, offs = _ensemble(nens)
    dev = F.DeviceModel(u)
    print("case: 8 channel members 21x7 KB=6, 4 elevation observations", flush=True)
    dev.run(3, graph=True)
    ens = enkf.Ensemble(u, offs, nens)
    # spread the members' elevations so the ensemble has a covariance
    el = dev.get("el")
    rng = np.random.default_rng(1)
    for j in range(nens):
        el[ens.node_off[j] + 1: ens.node_off[j] + 1 + ens.m] += 0.05 * rng.standard_normal()
    dev.put("el", el)
    obs = [("el", node, None, 0.2, 0.01) for node in (30, 60, 90, 120)]
    rep = enkf.assimilate(dev, ens, obs, ["el"], truncation=0.99, seed=2)
    print(rep["rms_before"], rep["rms_after"], flush=True)
    assert rep["rms_after"] < 0.2 * rep["rms_before"]
    dev.run(2, graph=True)
    assert not dev.diagnostics().nonfinite
