"""Parity at the benchmark's full sizes.

* NUM=4000 (10 boxes, 20 waypoints per edge, 1000 particles) and NUM=16000
  (30 boxes, W <= 64, 2000 particles): the whole pipeline against the CPU
  oracle -- every pair of the all-pairs tables, rn, every edge, waypoint,
  flag, half-space -- and the search: its result and its whole stored path
  table against the oracle's search record.
* Three of the 256 uneven batch instances (NUM 1000-8000, 5-50 boxes,
  random start/goal corners): the same, against the oracle.
* The NUM=32000 wall world (200 thin boxes, 4096 particles) and NUM=64000
  (100 boxes, W <= 128, 8192 particles): sampled rows of the pair tables,
  the neighbour radius as an exact order statistic of the full table, the
  edge sets of the sampled rows; then the whole edge pass over the device's
  edge set (every waypoint, flag and half-space) and the search on the
  device's tables (result and stored path table) against the oracle; the
  split-BF16 form of the search against the FP32 one (identical table, band
  count reported).  NUM=32000 also runs the batched full MC on 64 candidate
  paths against the oracle; NUM=64000 reports the cost error of the
  FP64-polished arrival times against the FP64 oracle.
* Against the FP64 oracle (every returned path): the same node sequence
  sampled from FP64 arrival times in FP64, its full-MC CP with the same noise
  draws within one MC standard error of the returned CP.
"""
import os

import numpy as np
import pytest

import oracle
import pump_amd as P
from oracle import pipeline as OP
from pump_amd import configs
from test_gpu_edges import check_fp32_waypoints, fp32_bound

pytestmark = pytest.mark.gpu

O = oracle.load()
NT = min(16, os.cpu_count() or 1)


def assert_close_f32(got, ref, rel=1e-5, max_mismatch_frac=1e-5):
    got = np.asarray(got, np.float32)
    ref = np.asarray(ref, np.float32)
    assert got.shape == ref.shape
    if got.size == 0:
        return
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


def gpu_tables(rm, e, f, hs):
    """The device's roadmap arrays as the oracle search reads them."""
    return dict(row_start=e["row_start"], edge_dst=e["edge_dst"], copts_edge=e["copts_edge"],
                topts_edge=e["topts_edge"], h=e["h"], is_free_edges=f["is_free_edges"],
                is_free_samples=f["is_free_samples"], wp_offsets=f["wp_offsets"],
                as_=hs["as_"], bs=hs["bs"], max_hs=rm.info.max_hs, rn=np.float32(rm.info.rn),
                n_edges=rm.info.n_edges)


SEARCH_KEYS = ("status", "itrs", "num_paths", "num_considered", "num_goal_paths",
               "soln_path_idx", "soln_nodes", "particle_steps", "clamped_time_idx", "mccp_calls")
TABLE_KEYS = ("path_prev", "path_node", "path_cost", "path_time", "path_cp")


def search_parity(rm, prm, tables, s, o, ot, xc):
    """The device search against the oracle's search over the same tables:
    every result field and every row of the stored path table, bit for bit."""
    got = P.search_result_dict(rm.search(prm))
    ref = OP.search(tables, s, s[-1], o, ot, xc, prm, nthreads=NT, path_cap=prm.path_cap,
                    trace=True, wf_cap=0)
    print("gpu", {k: got[k] for k in SEARCH_KEYS + ("soln_cost", "soln_cp", "band_particles")},
          "\noracle", {k: ref[k] for k in SEARCH_KEYS + ("soln_cost", "soln_cp")}, flush=True)
    for k in SEARCH_KEYS:
        assert got[k] == ref[k], k
    for k in ("soln_cost", "soln_cp", "soln_time", "soln_hsmc_cp"):
        assert np.float32(got[k]) == np.float32(ref[k]), k
    tab = rm.paths(got["num_paths"])
    for k in TABLE_KEYS:
        assert np.array_equal(tab[k], ref[k]), k
    return got, tab


@pytest.mark.timeout(900)
@pytest.mark.parametrize("i,bits", [(0, 0), (0, 1), (1, 0)])
def test_full_pipeline_parity(i, bits):
    cfg, s, o, ot, xc, rm = build(i, validity_bits=bits)
    prm = cfg["params"]
    orc = OP.roadmap(s, o, s[-1], prm, nthreads=NT)
    c, t = rm.pairs()
    assert_close_f32(t, orc["topts"])
    assert_close_f32(c, orc["copts"])
    del c, t
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
    print(cfg["name"], "edges", rm.info.n_edges, "free", int(f["is_free_edges"].sum()), flush=True)
    # the FP32-waypoint edge pass on the same edges: the same flags bit for
    # bit, waypoints within its stated bound of the oracle's
    dv = rm.device()
    E, n_wp = rm.info.n_edges, rm.info.n_waypoints
    disc32 = P.DeviceArray(shape=(n_wp, 6), dtype=np.float32, fill=0)
    flags32 = P.DeviceArray(shape=(E,), dtype=np.uint8, fill=9)
    cnt = P.DeviceArray(shape=(1,), dtype=np.uint64, fill=0)
    P.check("pump_edges_discretize_check_fp32", P.lib.pump_edges_discretize_check_fp32(
        dv.samples, dv.edge_src, dv.edge_dst, dv.topts_edge, E, dv.wp_offsets, 0, dv.obstacles,
        o.shape[0], disc32.p, flags32.p, cnt.p, None))
    assert np.array_equal(flags32.numpy(), orc["is_free_edges"])
    worst = check_fp32_waypoints(disc32.numpy(), orc["disc"], fp32_bound(
        s, e["edge_src"], e["edge_dst"], e["topts_edge"], f["wp_offsets"]))
    print(cfg["name"], "FP32 waypoints: flags equal, exact re-tests", int(cnt.numpy()[0]),
          "largest error / bound", worst, flush=True)
    # the search on the oracle's own roadmap (equal to the device's above)
    got, _ = search_parity(rm, prm, orc, s, o, ot, xc)
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


def full_edge_pass(cfg, s, o, rm, e, f, hs):
    """Every waypoint, flag and half-space of the device's edge set against
    the oracle (findDiscretizedPath re-bisects each edge, as the reference's
    host loop does; then freeEdges and cacheHalfspaces)."""
    prm = cfg["params"]
    num = s.shape[0]
    wo = f["wp_offsets"]
    E = rm.info.n_edges
    W = np.diff(wo)
    assert np.array_equal(W, OP.waypoint_counts(e["topts_edge"], prm.w_fixed, prm.wp_dt,
                                                 prm.w_cap))
    n_wp = int(wo[-1])
    disc = np.zeros((n_wp, 6), np.float32)
    O.po_discretize_edges(oracle.fp(s), oracle.lp(e["row_start"]), oracle.ip(e["edge_dst"]), num,
                          oracle.lp(wo), oracle.fp(disc), NT)
    assert_close_f32(rm.waypoints(), disc)
    flags = np.zeros(E, np.uint8)
    O.po_free_edges(oracle.fp(o), o.shape[0], oracle.fp(disc), oracle.lp(wo), E,
                    oracle.bp(flags), NT)
    assert np.array_equal(f["is_free_edges"], flags)
    mh = rm.info.max_hs
    as_ = np.zeros((n_wp * mh, 3), np.float32)
    bs = np.zeros(n_wp * mh, np.float32)
    cnt = np.zeros(n_wp, np.int32)
    O.po_setup_as_bs(n_wp * mh, oracle.fp(bs), oracle.fp(as_))
    O.po_cache_halfspaces(E, oracle.fp(disc), oracle.lp(wo), oracle.bp(flags), o.shape[0], mh,
                          oracle.fp(o), oracle.fp(as_), oracle.fp(bs), oracle.ip(cnt),
                          O.po_maxd2(np.float32(prm.offset_mult)), NT)
    assert np.array_equal(hs["count_stored"], cnt)
    assert_close_f32(hs["bs"], bs)
    assert_close_f32(hs["as_"], as_)
    print(cfg["name"], "edges", E, "free", int(flags.sum()), "waypoints", n_wp,
          "stored half-spaces", int(cnt.sum()), flush=True)


def mccp_candidates(cfg, s, o, xc, rm, tab, n=64):
    """The batched full MC (pump_collision_probability_many) on n candidate
    paths -- the deepest stored paths of the search, traced back to the init
    and sampled by findOptimalPath -- against the oracle's per-path hits."""
    prm = cfg["params"]
    prev, node = tab["path_prev"], tab["path_node"]
    depth = np.zeros(prev.size, np.int64)
    for k in range(1, prev.size):
        depth[k] = depth[prev[k]] + 1
    pick = np.lexsort((np.arange(prev.size), -depth))[:n]
    cap = prm.path_cap
    paths = np.zeros((pick.size, cap, 6), np.float32)
    T = np.zeros(pick.size, np.int32)
    n_mc, tsteps = xc.shape[0], xc.shape[1]
    ref = np.zeros(pick.size, np.int64)
    for q, k in enumerate(pick.tolist()):
        chain = []
        while k != -2:
            chain.append(int(node[k]))
            k = int(prev[k])
        xs = np.ascontiguousarray(s[np.asarray(chain[::-1])], np.float32)
        ln = np.zeros(1, np.int32)
        O.po_find_optimal_path(prm.dt, oracle.fp(paths[q]), oracle.fp(xs), xs.shape[0],
                               oracle.ip(ln), cap)
        assert ln[0] > 0
        T[q] = ln[0]
        hits = np.zeros(n_mc, np.int32)
        ref[q] = O.po_mccp_hits(oracle.fp(o), o.shape[0], oracle.fp(xc), tsteps, n_mc,
                                np.float32(prm.offset_mult), oracle.fp(paths[q]), int(ln[0]),
                                oracle.ip(hits))
    g_o, g_x = P.DeviceArray(o), P.DeviceArray(xc)
    g_p, g_t = P.DeviceArray(paths), P.DeviceArray(T)
    g_h = P.DeviceArray(shape=(pick.size,), dtype=np.int32)
    P.check("pump_collision_probability_many", P.lib.pump_collision_probability_many(
        g_o.p, o.shape[0], g_x.p, tsteps, n_mc, np.float32(prm.offset_mult), g_p.p, cap, g_t.p,
        pick.size, g_h.p, 1, None))
    got = g_h.numpy()
    print(cfg["name"], f"full MC of {pick.size} candidate paths (depth {depth[pick].min()}-"
          f"{depth[pick].max()}, {T.min()}-{T.max()} rows): hits {got.min()}-{got.max()} of "
          f"{n_mc}", flush=True)
    assert np.array_equal(got, ref)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("i", [3, 4])
def test_full_size_sampled(i):
    # NUM=64000 keeps validity as bit masks (bytes would need 265 GB) and
    # also builds the FP64-polished arrival times
    over = {"validity_bits": 1, "polish": 1} if i == 4 else {}
    cfg, s, o, ot, xc, rm = build(i, **over)
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
    # edge sets of the sampled rows, from the oracle's rows: costs below rn,
    # ascending destination
    e = rm.edges()
    assert e["row_start"][-1] == rm.info.n_edges == np.count_nonzero(c < rn)
    jj = np.arange(num - 1)
    for k, r in enumerate(rows.tolist()):
        sel = rc[k] < rn
        dst = (jj + (jj >= r))[sel]
        lo, hi = e["row_start"][r], e["row_start"][r + 1]
        assert np.array_equal(e["edge_dst"][lo:hi], dst)
        assert_close_f32(e["copts_edge"][lo:hi], rc[k][sel])
        assert_close_f32(e["topts_edge"][lo:hi], rt[k][sel])
    del c, t
    if i == 4:
        # FP64-polished arrival times of the sampled rows against the FP64
        # oracle: cost errors of c64(polished tau), of the FP32 table, and of
        # c64(bisected tau), each relative to max(|c64(tau*)|, 0.01)
        pol = rm.polished()
        tp = np.concatenate([pol[r * (num - 1):(r + 1) * (num - 1)] for r in rows.tolist()])
        del pol
        st = np.zeros(6, np.float64)
        O.po64_rows_cost_error(oracle.fp(s), num, oracle.ip(rows), rows.size,
                               np.float32(prm.tmax), oracle.fp(rt), oracle.fp(rc),
                               oracle.dp(np.ascontiguousarray(tp)), oracle.dp(st), NT)
        print(f"FP64 oracle over {int(st[5])} pairs: max relative cost error, polished tau "
              f"{st[0]:.2e}, FP32 table {st[1]:.2e}, bisected tau {st[2]:.2e}; max relative "
              f"tau error polished {st[3]:.2e}, bisected {st[4]:.2e}", flush=True)
        assert st[0] <= 1e-6 and st[3] <= 1e-6
        assert st[0] < st[2]
    # every waypoint, flag and half-space of the device's edge set
    f = rm.flags()
    hs = rm.halfspaces()
    full_edge_pass(cfg, s, o, rm, e, f, hs)
    # the search on the device's tables, against the oracle's search record
    tables = gpu_tables(rm, e, f, hs)
    got, tab = search_parity(rm, prm, tables, s, o, ot, xc)
    assert got["status"] in (0, 1)
    if got["soln_len"] > 0:
        rederive_path(cfg, s, o, ot, xc, rm, got, e, f, hs)
    # split-BF16 half-space products: the same stored path table
    p1 = configs.config(i, keep_pairs=1, **dict(over, validity_bits=1, precision=1))["params"]
    got1 = P.search_result_dict(rm.search(p1))
    tab1 = rm.paths(got1["num_paths"])
    print(cfg["name"], "split-BF16 search: band re-tests", got1["band_particles"],
          "of particle-steps", got1["particle_steps"], flush=True)
    for k in SEARCH_KEYS:
        assert got1[k] == got[k], k
    for k in TABLE_KEYS:
        assert np.array_equal(tab1[k], tab[k]), k
    if i == 3:
        mccp_candidates(cfg, s, o, xc, rm, tab)
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

