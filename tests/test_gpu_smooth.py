"""Smoothing and controller discretisation (PUMP.cu:406-568) on the device
(speculatively batched full-MC launches) against the oracle's sequential
restatement: the alpha/CP/cost trace, the kept path, its cost and CP, the
last iterate and the discretised rows, bit for bit."""
import numpy as np
import pytest

import pump_amd as P
from oracle import pipeline as OP
from pump_amd import configs

pytestmark = pytest.mark.gpu

CASES = [
    # (num, n_obs, w_fixed, offset_mult, n_p, k, seed, cp_target)
    (192, 10, 20, 0.5, 256, 30, 0, 0.01),
    (256, 12, 0, 1.0, 300, 25, 1, 0.05),
    (300, 0, 9, 0.5, 128, 30, 2, 0.01),   # no obstacles: every CP is 0
    (200, 40, 0, 0.7, 200, 40, 3, 0.02),
    (400, 25, 0, 2.0, 512, 30, 4, 0.1),
]


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("case", CASES)
def test_smooth_parity(case):
    num, n_obs, w_fixed, om, n_p, k, seed, target = case
    cfg = configs.small(num=num, n_obs=n_obs, seed=seed, w_fixed=w_fixed, n_p=n_p,
                        offset_mult=om, k=k, n_mc=512, tmax_steps=1200, cp_target=target)
    s, o = P.gen_problem(cfg["spec"])
    ot, xc = P.make_noise(cfg["noise"])
    prm = cfg["params"]
    print("case", case, flush=True)
    rm = P.Roadmap(s, o, s[-1], ot, xc, prm)
    res = rm.search()
    sm, path, xs_s, disc = rm.smooth(res)
    ref = OP.smooth(P.search_result_dict(res), s, o, xc, prm)
    print("device", sm.status, sm.itrs, sm.mccp_launches, sm.cost_before, sm.cost, sm.cp,
          "oracle", ref.get("itrs"), ref.get("cost"), ref.get("cp"), flush=True)
    assert sm.status == ref["status"]
    if sm.status != 0:
        rm.close()
        return
    assert sm.itrs == ref["itrs"]
    trace = np.array([[sm.trace[i][j] for j in range(3)] for i in range(sm.itrs)], np.float32)
    assert np.array_equal(bits(trace), bits(ref["trace"]))
    assert np.float32(sm.cost) == np.float32(ref["cost"])
    assert np.float32(sm.cp) == np.float32(ref["cp"])
    assert np.array_equal(bits(path), bits(ref["path"]))
    assert np.array_equal(bits(xs_s), bits(ref["xs_smooth"]))
    assert np.array_equal(bits(disc), bits(ref["discretized"]))
    assert sm.mccp_launches <= (sm.itrs + 3) // 4
    rm.close()


def test_smooth_detour_bisection():
    """A solution that detours over a box the straight path crosses: the CP
    runs from 1 (alpha = 1) through partial values to just below the target,
    so every branch of the bisection and the kept-path rule are exercised."""
    xs = np.zeros((4, 6), np.float32)
    xs[:, :3] = [[0.1, 0.5, 0.5], [0.35, 0.85, 0.5], [0.65, 0.85, 0.5], [0.9, 0.5, 0.5]]
    obs = np.zeros((1, 12), np.float32)
    obs[0, :3], obs[0, 6:9] = [0.4, 0.35, 0.35], [0.6, 0.65, 0.65]
    obs[0, 3:6], obs[0, 9:12] = -3, 3
    ot, xc = P.make_noise(P.noise_spec(n_mc=256, tmax_steps=1500))
    prm = P.params(rn_perc=0.9, w_fixed=9, offset_mult=0.5, n_p=128, cp_target=0.05)
    rm = P.Roadmap(xs, obs, xs[-1], ot, xc, prm)
    res = P.SearchResult()
    res.status, res.soln_len = 0, 4
    for i in range(4):
        res.soln_nodes[i] = i
    res.soln_cost, res.soln_cp = np.float32(150.0), 0.0
    sm, path, xs_s, disc = rm.smooth(res)
    ref = OP.smooth(P.search_result_dict(res), xs, obs, xc, prm)
    assert sm.status == 0 == ref["status"]
    trace = np.array([[sm.trace[i][j] for j in range(3)] for i in range(sm.itrs)], np.float32)
    assert sm.itrs == ref["itrs"] and np.array_equal(bits(trace), bits(ref["trace"]))
    assert len(set(trace[:, 1].tolist())) > 3
    assert np.float32(sm.cost) == np.float32(ref["cost"]) and np.float32(sm.cp) == ref["cp"]
    assert np.array_equal(bits(path), bits(ref["path"]))
    assert np.array_equal(bits(xs_s), bits(ref["xs_smooth"]))
    assert np.array_equal(bits(disc), bits(ref["discretized"]))
    rm.close()
