"""Parity at the benchmark's full sizes.

* NUM=4000 (10 boxes, 20 waypoints per edge, 1000 particles): the whole
  pipeline against the CPU oracle, every table and the search result, in both
  validity forms.
* Three of the 256 uneven batch instances (NUM 1000-8000, 5-50 boxes,
  random start/goal corners): the same, against the oracle.
* NUM=16000 (30 boxes, W <= 64, 2000 particles), the NUM=32000 wall world
  (200 thin boxes) and NUM=64000 (100 boxes, W <= 128, 8192 particles):
  sampled rows of the pair tables against the oracle, the
  neighbour radius as an exact order statistic of the full table, the edge
  sets of the sampled rows, waypoints, flags and half-spaces of sampled edges
  against the oracle, and the returned path re-derived (chain of free edges,
  cost and time sums, full-MC CP and HSMC CP recomputed by the oracle).
* Against the FP64 oracle (every returned path, NUM=4000 included): the same
  node sequence sampled from FP64 arrival times in FP64, its full-MC CP with
  the same noise draws within one MC standard error of the returned CP.
"""
import os

import numpy as np
import pytest

import oracle
import pump_amd as P
from oracle import pipeline as OP
from pump_amd import configs

pytestmark = pytest.mark.gpu

O = oracle.load()
NT = min(16, os.cpu_count() or 1)


def assert_close_f32(got, ref, rel=1e-5, max_mismatch_frac=1e-5):
    got = np.asarray(got, np.float32)
    ref = np.asarray(ref, np.float32)
    assert got.shape == ref.shape
    same = got.view(np.uint32) == ref.view(np.uint32)
    assert 1 - same.mean() <= max_mismatch_frac, f"{(~same).sum()} of {same.size} differ"
    if not same.all():
        d = np.abs(got[~same].astype(np.float64) - ref[~same])
        assert np.all(d <= rel * np.abs(ref[~same]) + 1e-30)


def build(i, **over):
    cfg = configs.config(i, keep_pairs=1, **over)
    s, o = P.gen_problem(cfg["spec"])
    ot, xc = P.make_noise(cfg["noise"])
    print("workload", cfg["name"], flush=True)
    rm = P.Roadmap(s, o, s[-1], ot, xc, cfg["params"])
    return cfg, s, o, ot, xc, rm


@pytest.mark.parametrize("bits", [0, 1])
def test_full_num4000_parity(bits):
    cfg, s, o, ot, xc, rm = build(0, validity_bits=bits)
    prm = cfg["params"]
    orc = OP.roadmap(s, o, s[-1], prm, nthreads=NT)
    c, t = rm.pairs()
    assert_close_f32(t, orc["topts"])
    assert_close_f32(c, orc["copts"])
    assert rm.info.rn_idx == orc["rn_idx"] and np.float32(rm.info.rn) == orc["rn"]
    e = rm.edges()
    for k in ("row_start", "edge_src", "edge_dst"):
        assert np.array_equal(e[k], orc[k]), k
    for k in ("copts_edge", "topts_edge", "h"):
        assert_close_f32(e[k], orc[k])
    f = rm.flags()
    for k in ("is_free_samples", "wp_offsets", "is_free_edges"):
        assert np.array_equal(f[k], orc[k]), k
    assert_close_f32(rm.waypoints(), orc["disc"])
    hs = rm.halfspaces()
    assert np.array_equal(hs["count_stored"], orc["count_stored"])
    assert_close_f32(hs["bs"], orc["bs"])
    assert_close_f32(hs["as_"], orc["as_"])
    got = P.search_result_dict(rm.search())
    ref = OP.search(orc, s, s[-1], o, ot, xc, prm, nthreads=NT, path_cap=prm.path_cap)
    print("gpu", got, "\noracle", ref, flush=True)
    for k in ("status", "itrs", "num_paths", "num_considered", "num_goal_paths", "soln_path_idx",
              "soln_nodes", "particle_steps", "clamped_time_idx"):
        assert got[k] == ref[k], k
    for k in ("soln_cost", "soln_cp", "soln_time", "soln_hsmc_cp"):
        assert np.float32(got[k]) == np.float32(ref[k]), k
    if got["soln_len"] > 0:
        rederive_path(cfg, s, o, ot, xc, rm, got, e, f, hs)
    rm.close()


def sub_edges(s, src, dst, W):
    """Arrays that present sampled edges to the oracle as their own roadmap:
    sample 2k is edge k's source, 2k+1 its destination."""
    n = src.shape[0]
    samples = np.empty((2 * n, 6), np.float32)
    samples[0::2], samples[1::2] = s[src], s[dst]
    row_start = np.repeat(np.arange(n + 1, dtype=np.int64), 2)[1:2 * n + 2]
    edge_dst = (2 * np.arange(n) + 1).astype(np.int32)
    off = np.zeros(n + 1, np.int64)
    np.cumsum(W, out=off[1:])
    return samples, row_start, edge_dst, off


def rederive_path(cfg, s, o, ot, xc, rm, got, e, f, hs):
    """The returned path (or, with no path below the CP target, the last goal
    path the selection checked) is a chain of free edges from sample 0 to the
    goal; its cost, time, full-MC CP and HSMC CP follow from the tables."""
    prm = cfg["params"]
    num = s.shape[0]
    nodes = list(got["soln_nodes"])
    assert nodes[0] == 0 and nodes[-1] == num - 1
    rs, dst = e["row_start"], e["edge_dst"]
    eids = []
    for a, b in zip(nodes[:-1], nodes[1:]):
        row = dst[rs[a]:rs[a + 1]]
        k = np.searchsorted(row, b)
        assert k < row.size and row[k] == b, (a, b)
        eid = int(rs[a] + k)
        assert f["is_free_edges"][eid] == 1
        eids.append(eid)
    cost = np.float32(0)
    tm = np.float32(0)
    for eid in eids:
        cost = np.float32(cost + e["copts_edge"][eid])
        tm = np.float32(tm + e["topts_edge"][eid])
    assert np.float32(got["soln_cost"]) == cost
    assert np.float32(got["soln_time"]) == tm
    # full-MC CP of the smoothed-free path (findOptimalPath + collisionProbability)
    xs = np.ascontiguousarray(s[np.asarray(nodes)], np.float32)
    cap = prm.path_cap
    split = np.zeros((cap, 6), np.float32)
    ln = np.zeros(1, np.int32)
    O.po_find_optimal_path(prm.dt, oracle.fp(split), oracle.fp(xs), len(nodes), oracle.ip(ln), cap)
    assert ln[0] > 0
    n_mc, T = xc.shape[0], xc.shape[1]
    cp = O.po_collision_probability(oracle.fp(o), o.shape[0], oracle.fp(xc), T, n_mc,
                                    prm.offset_mult, oracle.fp(split), int(ln[0]))
    assert np.float32(got["soln_cp"]) == np.float32(cp)
    assert (got["soln_cp"] < prm.cp_target) == (got["status"] == 0)
    # against the FP64 oracle: the same nodes' path sampled from FP64 arrival
    # times in FP64 and its full-MC CP with the same noise draws must lie
    # within one MC standard error
    split64 = np.zeros((cap, 6), np.float64)
    rows64 = O.po64_find_optimal_path(prm.dt, oracle.dp(split64), oracle.fp(xs),
                                      len(nodes), cap)
    assert rows64 > 0
    hits64 = O.po64_mccp_hits(oracle.fp(o), o.shape[0], oracle.fp(xc), T, n_mc,
                              float(np.float32(prm.offset_mult)),
                              oracle.dp(split64), rows64)
    cp64 = hits64 / n_mc
    se = max(np.sqrt(cp64 * (1 - cp64) / n_mc), 1.0 / n_mc)
    cost64 = sum(O.po64_cost(O.po64_topt(oracle.fp(xs[k]), oracle.fp(xs[k + 1]), 5.0),
                             oracle.fp(xs[k]), oracle.fp(xs[k + 1])) for k in range(len(nodes) - 1))
    print(f"FP64 oracle: path CP {cp:.6f} vs {cp64:.6f} ({abs(cp - cp64) / se:.2f} MC standard "
          f"errors), path cost rel err {abs(float(got['soln_cost']) - cost64) / cost64:.2e}",
          flush=True)
    assert abs(cp - cp64) <= se
    # the reference's FP32 bisection of a finite-difference derivative stops
    # short of the FP64 minimum (per-edge errors up to tens of percent in tau,
    # well under that in cost); the path cost stays within a few percent
    assert abs(float(got["soln_cost"]) - cost64) <= 0.05 * cost64
    # HSMC CP along the path: one propagate step per edge, oracle arithmetic
    n_p = prm.n_p
    wo = f["wp_offsets"]
    valid = np.ones((len(eids) + 1, n_p), np.uint8)
    times = np.zeros(len(eids) + 1, np.float32)
    for k, eid in enumerate(eids):
        O.po_propagate_valid_particles(
            1, oracle.lp(wo), oracle.ip(np.array([eid], np.int32)),
            oracle.ip(np.array([k], np.int32)), k + 1, oracle.fp(hs["as_"]),
            oracle.fp(hs["bs"]), oracle.fp(times), oracle.fp(e["topts_edge"]), prm.dt,
            rm.info.max_hs, oracle.fp(ot), ot.shape[0], np.float32(prm.offset_mult), n_mc,
            oracle.bp(valid), n_p, 1)
        times[k + 1] = np.float32(times[k] + e["topts_edge"][eid])
    hcp = np.float32(np.count_nonzero(valid[-1] == 0)) / np.float32(n_p)
    assert np.float32(got["soln_hsmc_cp"]) == hcp


@pytest.mark.parametrize("i", [1, 3, 4])
def test_full_size_sampled(i):
    # NUM=64000 keeps validity as bit masks: bytes would need 265 GB
    cfg, s, o, ot, xc, rm = build(i, **({"validity_bits": 1} if i == 4 else {}))
    prm = cfg["params"]
    num = s.shape[0]
    rng = np.random.default_rng(100 + i)
    c, t = rm.pairs()
    # sampled rows of the pair tables
    rows = np.sort(rng.choice(num, 24, replace=False)).astype(np.int32)
    rc = np.empty((rows.size, num - 1), np.float32)
    rt = np.empty((rows.size, num - 1), np.float32)
    O.po_fill_rows(oracle.fp(s), num, np.float32(prm.tmax), oracle.ip(rows), rows.size,
                   oracle.fp(rc), oracle.fp(rt), NT)
    for k, r in enumerate(rows.tolist()):
        assert_close_f32(t[r * (num - 1):(r + 1) * (num - 1)], rt[k])
        assert_close_f32(c[r * (num - 1):(r + 1) * (num - 1)], rc[k])
    # the neighbour radius is the rn_idx-th smallest cost of the whole table
    rn_idx = O.po_rn_index(num, np.float32(prm.rn_perc))
    assert rm.info.rn_idx == rn_idx
    rn = np.float32(rm.info.rn)
    assert np.count_nonzero(c < rn) <= rn_idx < np.count_nonzero(c <= rn)
    # edge sets of the sampled rows: costs below rn, ascending destination
    e = rm.edges()
    assert e["row_start"][-1] == rm.info.n_edges == np.count_nonzero(c < rn)
    jj = np.arange(num - 1)
    for r in rows.tolist():
        seg = c[r * (num - 1):(r + 1) * (num - 1)]
        sel = seg < rn
        dst = (jj + (jj >= r))[sel]
        lo, hi = e["row_start"][r], e["row_start"][r + 1]
        assert np.array_equal(e["edge_dst"][lo:hi], dst)
        assert np.array_equal(e["copts_edge"][lo:hi], seg[sel])
        assert np.array_equal(e["topts_edge"][lo:hi], t[r * (num - 1):(r + 1) * (num - 1)][sel])
    del c, t
    # waypoints, flags and half-spaces of sampled edges
    f = rm.flags()
    wo = f["wp_offsets"]
    eid = np.sort(rng.choice(rm.info.n_edges, 4000, replace=False))
    src, dst = e["edge_src"][eid], e["edge_dst"][eid]
    W = (wo[eid + 1] - wo[eid]).astype(np.int64)
    for k in range(eid.size):
        assert O.po_waypoints_for(e["topts_edge"][eid[k]], prm.w_fixed, np.float32(prm.wp_dt),
                                  prm.w_cap) == W[k]
    samples, row_start, edge_dst, off = sub_edges(s, src, dst, W)
    disc = np.zeros((int(off[-1]), 6), np.float32)
    O.po_discretize_edges(oracle.fp(samples), oracle.lp(row_start), oracle.ip(edge_dst),
                          samples.shape[0], oracle.lp(off), oracle.fp(disc), NT)
    gw = rm.waypoints()
    take = np.concatenate([np.arange(wo[x], wo[x + 1]) for x in eid])
    assert_close_f32(gw[take], disc)
    flags = np.zeros(eid.size, np.uint8)
    O.po_free_edges(oracle.fp(o), o.shape[0], oracle.fp(disc), oracle.lp(off), eid.size,
                    oracle.bp(flags), NT)
    assert np.array_equal(f["is_free_edges"][eid], flags)
    assert flags.min() == 0 and flags.max() == 1
    mh = rm.info.max_hs
    n_wp = int(off[-1])
    as_ = np.zeros((n_wp * mh, 3), np.float32)
    bs = np.zeros(n_wp * mh, np.float32)
    cnt = np.zeros(n_wp, np.int32)
    O.po_setup_as_bs(n_wp * mh, oracle.fp(bs), oracle.fp(as_))
    O.po_cache_halfspaces(eid.size, oracle.fp(disc), oracle.lp(off), oracle.bp(flags),
                          o.shape[0], mh, oracle.fp(o), oracle.fp(as_), oracle.fp(bs),
                          oracle.ip(cnt), O.po_maxd2(np.float32(prm.offset_mult)), NT)
    hs = rm.halfspaces()
    assert np.array_equal(hs["count_stored"][take], cnt)
    slots = (take[:, None] * mh + np.arange(mh)[None, :]).ravel()
    assert_close_f32(hs["bs"][slots], bs)
    assert_close_f32(hs["as_"][slots], as_)
    assert cnt.max() > 0
    # the search and its returned path
    got = P.search_result_dict(rm.search())
    print(cfg["name"], {k: got[k] for k in ("status", "itrs", "num_paths", "soln_cost",
                                            "soln_cp", "soln_hsmc_cp")}, flush=True)
    assert got["status"] in (0, 1)
    if got["soln_len"] > 0:
        rederive_path(cfg, s, o, ot, xc, rm, got, e, f, hs)
    rm.close()


def test_batch_instances_parity():
    """Three of the 256 uneven batch instances (the smallest NUM, the
    fewest and the most boxes): whole pipeline and search against the oracle."""
    batch = configs.batch_config()
    nums = np.array([b["spec"].num for b in batch])
    nobs = np.array([b["spec"].n_obs for b in batch])
    picks = sorted({int(np.argmin(nums)), int(np.argmin(nobs + nums / 1e5)),
                    int(np.argmax(nobs - nums / 1e5))})
    for k in picks:
        cfg = batch[k]
        prm = cfg["params"]
        prm.keep_pairs = 1
        s, o = P.gen_problem(cfg["spec"])
        ot, xc = P.make_noise(cfg["noise"])
        print("instance", cfg["name"], "boxes", o.shape[0], "n_p", prm.n_p, flush=True)
        rm = P.Roadmap(s, o, s[-1], ot, xc, prm)
        orc = OP.roadmap(s, o, s[-1], prm, nthreads=NT)
        c, t = rm.pairs()
        assert_close_f32(t, orc["topts"])
        assert_close_f32(c, orc["copts"])
        assert np.float32(rm.info.rn) == orc["rn"]
        e = rm.edges()
        for key in ("row_start", "edge_dst"):
            assert np.array_equal(e[key], orc[key]), key
        f = rm.flags()
        for key in ("wp_offsets", "is_free_edges", "is_free_samples"):
            assert np.array_equal(f[key], orc[key]), key
        hs = rm.halfspaces()
        assert np.array_equal(hs["count_stored"], orc["count_stored"])
        got = P.search_result_dict(rm.search())
        ref = OP.search(orc, s, s[-1], o, ot, xc, prm, nthreads=NT, path_cap=prm.path_cap)
        for key in ("status", "itrs", "num_paths", "num_goal_paths", "soln_nodes"):
            assert got[key] == ref[key], key
        for key in ("soln_cost", "soln_cp", "soln_hsmc_cp"):
            assert np.float32(got[key]) == np.float32(ref[key]), key
        rm.close()

