"""The noise tables generated on the device (pump_make_noise_device) against
the host generator (pump_make_noise) and against the mathematics: the FP64
Cholesky factor of the 6x6 process covariance, the FP32-propagated position
covariance, and the sample covariance of the tables."""
import math

import numpy as np
import pytest

import pump_amd as P

pytestmark = pytest.mark.gpu


def system(ns):
    """A_cl and Q of the closed-loop double integrator in FP64, host order."""
    dt = float(np.float32(ns.dt))
    a11 = 1 - ns.kp * dt * dt / 2
    a12 = dt - ns.kd * dt * dt / 2
    a21 = -ns.kp * dt
    a22 = 1 - ns.kd * dt
    A = np.zeros((6, 6))
    Q = np.zeros((6, 6))
    for d in range(3):
        A[d, d], A[d, d + 3], A[d + 3, d], A[d + 3, d + 3] = a11, a12, a21, a22
        Q[d, d] = ns.q * dt * dt * dt / 3
        Q[d, d + 3] = Q[d + 3, d] = ns.q * dt * dt / 2
        Q[d + 3, d + 3] = ns.q * dt
    return A, Q


def cholesky(Q):
    """Cholesky-Banachiewicz in FP64, the generator's operation order."""
    L = np.zeros((6, 6))
    for i in range(6):
        for j in range(i + 1):
            s = Q[i, j]
            for k in range(j):
                s -= L[i, k] * L[j, k]
            L[i, j] = math.sqrt(s) if i == j else s / L[j, j]
    return L


@pytest.mark.parametrize("n_mc,T", [(2048, 2001), (333, 97)])
def test_noise_device_matches_host_and_theory(n_mc, T):
    ns = P.noise_spec(n_mc=n_mc, tmax_steps=T)
    ot, xc, sg, L = P.make_noise_device(ns)
    hot, hxc, hsg = P.make_noise(ns, sigma=True)
    A, Q = system(ns)
    Lr = cholesky(Q)
    assert np.array_equal(L.view(np.uint64), Lr.view(np.uint64))
    assert np.allclose(L @ L.T, Q, rtol=1e-14, atol=1e-20)
    assert np.array_equal(sg.view(np.uint32), hsg.view(np.uint32))
    # both layouts hold the same values
    assert np.array_equal(ot.transpose(1, 0, 2), xc)
    # against the host tables: the only differences come from FP64 log/sin/cos
    # rounding, far below the float spacing of the offsets
    same = (ot.view(np.uint32) == hot.view(np.uint32)).mean()
    diff = np.abs(ot.astype(np.float64) - hot).max()
    print(f"device vs host tables: {same * 100:.4f}% bit-equal, max |diff| {diff:.2e}",
          flush=True)
    assert diff <= 1e-6 and same > 0.99
    # sample covariance of the positions at late steps against the propagated one
    for t in (T // 2, T - 1):
        x = ot[t].astype(np.float64)
        S = np.cov(x.T)
        sig = sg[t].astype(np.float64)
        se = np.sqrt((sig ** 2 + np.outer(np.diag(sig), np.diag(sig))) / (n_mc - 1))
        assert np.all(np.abs(S - sig) <= 5 * se + 1e-12), (t, S, sig)
