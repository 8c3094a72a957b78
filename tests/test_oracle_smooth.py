"""Pins of the oracle's smoothing and controller discretisation
(po_smooth, PUMP.cu:406-568).  PUMP.cu cannot be compiled here (it needs
CUDA and Thrust), so the loop is pinned by what fixes it: an obstacle-free
problem (every CP is 0, alpha stays 1, the kept cost is the smaller of the
solution's and the alpha = 1 iterate's), and every output recomputed from
the golden-pinned primitives (findOptimalPath, toptBisection, cost,
findDiscretizedPath, the full-MC hit count)."""
import numpy as np

import oracle
from oracle import fp, ip
from oracle import pipeline as OP

O = oracle.load()


class Prm:
    offset_mult = np.float32(0.5)
    dt = np.float32(0.05)
    cp_target = np.float32(0.05)


def chain(seed, n=5):
    rng = np.random.default_rng(seed)
    xs = np.concatenate([rng.uniform(0, 1, (n, 3)), rng.uniform(-0.5, 0.5, (n, 3))], 1)
    xs[0, 3:] = 0
    xs[-1, 3:] = 0
    return xs.astype(np.float32)


def noise(n_mc=128, T=1500, seed=3):
    rng = np.random.default_rng(seed)
    return rng.normal(0, 0.05, (n_mc, T, 3)).astype(np.float32)


def optimal_path(xs, dt, cap=4096):
    out = np.zeros((cap, 6), np.float32)
    n = np.zeros(1, np.int32)
    O.po_find_optimal_path(np.float32(dt), fp(out), fp(np.ascontiguousarray(xs)), xs.shape[0],
                           ip(n), cap)
    return out[: n[0]]


def path_cost(xs):
    c = np.float32(0)
    for i in range(xs.shape[0] - 1):
        a, b = np.ascontiguousarray(xs[i]), np.ascontiguousarray(xs[i + 1])
        tau = O.po_topt_bisection(fp(a), fp(b), np.float32(2))
        c = np.float32(c + np.float32(O.po_cost(np.float32(tau), fp(a), fp(b))))
    return c


def discretized(xs):
    rows = 1
    segs = []
    for i in range(xs.shape[0] - 1):
        a, b = np.ascontiguousarray(xs[i]), np.ascontiguousarray(xs[i + 1])
        tau = np.float32(O.po_topt_bisection(fp(a), fp(b), np.float32(2)))
        rows = int(np.float32(rows) + (tau / np.float32(0.05) + np.float32(1)))
        nw = int(tau / np.float32(0.05) + np.float32(1))
        wp = np.zeros((nw + 1, 6), np.float32)
        O.po_find_discretized_path(fp(wp), fp(a), fp(b), nw)
        segs.append(wp)
    out = np.zeros((rows, 6), np.float32)
    cur = 0
    for wp in segs:
        out[cur: cur + wp.shape[0]] = wp
        cur += wp.shape[0] - 1
    return out


def test_smooth_without_obstacles():
    xs = chain(1)
    xc = noise()
    soln_cost = path_cost(xs) * np.float32(1.5)  # a solution cost above the smoothed one
    res = dict(status=0, soln_nodes=list(range(xs.shape[0])), soln_cost=soln_cost, soln_cp=0.0)
    out = OP.smooth(res, xs, np.zeros((0, 12), np.float32), xc, Prm)
    assert out["status"] == 0 and out["itrs"] == 30
    assert np.all(out["trace"][:, 0] == 1) and np.all(out["trace"][:, 1] == 0)
    c1 = out["trace"][0, 2]
    assert np.all(out["trace"][:, 2] == c1)
    assert out["cp"] == 0 and np.float32(out["cost"]) == min(np.float32(soln_cost), c1)
    # alpha = 1: the iterate is the obstacle-free path's states
    assert np.float32(c1) == path_cost(out["xs_smooth"])
    assert np.array_equal(out["path"], optimal_path(out["xs_smooth"], Prm.dt))
    assert np.array_equal(out["discretized"], discretized(out["xs_smooth"]))
    # smoothing towards the unconstrained optimum lowers the cost
    assert c1 < path_cost(xs)


def test_smooth_bisection_against_primitives():
    # a detour over a box that the straight path from start to goal crosses
    xs = np.zeros((4, 6), np.float32)
    xs[:, :3] = [[0.1, 0.5, 0.5], [0.35, 0.85, 0.5], [0.65, 0.85, 0.5], [0.9, 0.5, 0.5]]
    xc = noise(n_mc=256)
    obs = np.zeros((1, 12), np.float32)
    obs[0, :3], obs[0, 6:9] = [0.4, 0.35, 0.35], [0.6, 0.65, 0.65]
    obs[0, 3:6], obs[0, 9:12] = -3, 3
    res = dict(status=0, soln_nodes=list(range(xs.shape[0])), soln_cost=path_cost(xs),
               soln_cp=0.0)
    out = OP.smooth(res, xs, obs, xc, Prm)
    assert out["status"] == 0 and out["itrs"] >= 1
    # the opt-path states at the solution's time fractions, and the blend
    ends = np.ascontiguousarray(np.stack([xs[0], xs[-1]]))
    opt = optimal_path(ends, np.float32(Prm.dt) / np.float32(10), cap=40000)
    cum = [np.float32(0)]
    for i in range(xs.shape[0] - 1):
        tau = O.po_topt_bisection(fp(np.ascontiguousarray(xs[i])),
                                  fp(np.ascontiguousarray(xs[i + 1])), np.float32(2))
        cum.append(np.float32(cum[-1] + np.float32(tau)))
    idx = [int(np.float32(opt.shape[0] - 1) * c / cum[-1]) for c in cum]
    xs_opt = opt[idx]
    lo, hi, alpha = np.float32(0), np.float32(1), np.float32(1)
    best = (np.float32(res["soln_cost"]), np.float32(0.0))
    for k, (a, cp, cost) in enumerate(out["trace"]):
        assert a == alpha
        xsm = ((np.float32(1) - alpha) * xs + alpha * xs_opt).astype(np.float32)
        path = optimal_path(xsm, Prm.dt)
        hits = O.po_mccp_hits(fp(obs), obs.shape[0], fp(xc), xc.shape[1], xc.shape[0],
                              Prm.offset_mult, fp(path), path.shape[0], None)
        assert cp == np.float32(hits) / np.float32(xc.shape[0])
        assert cost == path_cost(xsm)
        if cp < Prm.cp_target:
            lo = alpha
        else:
            hi = alpha
        alpha = np.float32((lo + hi) / np.float32(2))
        if cp < Prm.cp_target and (cp > best[1] or (cp == best[1] and cost < best[0])):
            best = (cost, cp)
    assert (np.float32(out["cost"]), np.float32(out["cp"])) == best
    assert np.array_equal(out["discretized"], discretized(out["xs_smooth"]))
    assert len(set(np.round(out["trace"][:, 1], 6))) > 1  # the CP varies along the bisection
