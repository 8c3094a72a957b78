"""End-to-end CPU oracle run of the roadmap precompute (mainPUMP.cu:248-504)
and the PUMP search (PUMP.cu:3-405) for one problem instance.

TEST INFRASTRUCTURE ONLY (the checker): used by tests/, smoke() and the
cpu_baseline leg of bench.py.
"""
import ctypes
import time

import numpy as np

from . import (SearchIn, SearchOut, SearchTrace, bp, dp, fp, ip, load, lp)

DT_WP = np.float32(0.05)


def waypoint_counts(topts_edge, w_fixed, wp_dt=0.05, w_cap=64):
    """W_e per edge (fixed numDisc+1, or clamp(ceil(tau/wp_dt), 2, cap) in float)."""
    if w_fixed > 0:
        return np.full(topts_edge.shape[0], w_fixed, np.int64)
    w = np.ceil(topts_edge.astype(np.float32) / np.float32(wp_dt)).astype(np.int64)
    return np.clip(w, 2, w_cap)


def offsets_from_counts(w):
    off = np.zeros(w.shape[0] + 1, np.int64)
    np.cumsum(w, out=off[1:])
    return off


def roadmap(samples, obstacles, goal, prm, nthreads=0, pairs=None, timings=None):
    """Oracle roadmap arrays for a Params-like object prm (tmax, rn_perc,
    w_fixed, wp_dt, w_cap, offset_mult, max_hs)."""
    O = load()
    samples = np.ascontiguousarray(samples, np.float32)
    obstacles = np.ascontiguousarray(obstacles, np.float32).reshape(-1, 12)
    goal = np.ascontiguousarray(goal, np.float32)
    num = samples.shape[0]
    n_obs = obstacles.shape[0]
    obs = obstacles if n_obs else np.zeros((1, 12), np.float32)
    t = timings if timings is not None else {}
    n_pairs = num * (num - 1)
    t0 = time.perf_counter()
    if pairs is None:
        copts = np.empty(n_pairs, np.float32)
        topts = np.empty(n_pairs, np.float32)
        O.po_fill_copts_topts(fp(samples), num, prm.tmax, fp(copts), fp(topts), nthreads)
    else:
        copts, topts = pairs
    t["pairs"] = time.perf_counter() - t0
    rn_idx = O.po_rn_index(num, prm.rn_perc)
    rn = O.po_select_rn(fp(copts), n_pairs, rn_idx)
    row_start = np.empty(num + 1, np.int64)
    n_edges = O.po_count_edges(fp(copts), num, rn, lp(row_start))
    edge_dst = np.empty(max(n_edges, 1), np.int32)
    copts_edge = np.empty(max(n_edges, 1), np.float32)
    topts_edge = np.empty(max(n_edges, 1), np.float32)
    O.po_fill_edges(fp(copts), fp(topts), num, rn, lp(row_start), ip(edge_dst), fp(copts_edge),
                    fp(topts_edge))
    edge_dst, copts_edge, topts_edge = edge_dst[:n_edges], copts_edge[:n_edges], topts_edge[:n_edges]
    edge_src = np.repeat(np.arange(num, dtype=np.int32), np.diff(row_start))
    max_nn = O.po_max_nn_size(lp(row_start), ip(np.ascontiguousarray(edge_dst)), n_edges, num)
    h = np.empty(num, np.float32)
    O.po_heuristic(fp(samples), num, fp(goal), prm.tmax, fp(h))
    free_samples = np.empty(num, np.uint8)
    O.po_sample_free(fp(obs), n_obs, fp(samples), num, bp(free_samples))
    w = waypoint_counts(topts_edge, prm.w_fixed, prm.wp_dt, prm.w_cap)
    wp_offsets = offsets_from_counts(w)
    n_wp = int(wp_offsets[-1])
    disc = np.zeros((max(n_wp, 1), 6), np.float32)
    t1 = time.perf_counter()
    O.po_discretize_edges(fp(samples), lp(row_start), ip(np.ascontiguousarray(edge_dst)), num,
                          lp(wp_offsets), fp(disc), nthreads)
    free_edges = np.empty(max(n_edges, 1), np.uint8)
    O.po_free_edges(fp(obs), n_obs, fp(disc), lp(wp_offsets), n_edges, bp(free_edges), nthreads)
    t["edge_pass"] = time.perf_counter() - t1
    free_edges = free_edges[:n_edges]
    max_hs = min(n_obs, prm.max_hs)
    hs = max(n_wp * max(max_hs, 1), 1)
    as_ = np.empty((hs, 3), np.float32)
    bs = np.empty(hs, np.float32)
    count = np.zeros(max(n_wp, 1), np.int32)
    O.po_setup_as_bs(hs, fp(bs), fp(as_))
    maxd2 = O.po_maxd2(prm.offset_mult)
    O.po_cache_halfspaces(n_edges, fp(disc), lp(wp_offsets), bp(np.ascontiguousarray(free_edges)),
                          n_obs, max_hs, fp(obs), fp(as_), fp(bs), ip(count), maxd2, nthreads)
    return dict(copts=copts, topts=topts, rn_idx=rn_idx, rn=np.float32(rn), row_start=row_start,
                edge_src=edge_src, edge_dst=np.ascontiguousarray(edge_dst),
                copts_edge=np.ascontiguousarray(copts_edge),
                topts_edge=np.ascontiguousarray(topts_edge), max_nn=max_nn, h=h,
                is_free_samples=free_samples, wp_offsets=wp_offsets, disc=disc[:n_wp],
                is_free_edges=np.ascontiguousarray(free_edges), max_hs=max_hs,
                as_=as_, bs=bs, count_stored=count[:n_wp], maxd2=np.float32(maxd2),
                n_edges=n_edges, n_wp=n_wp)


def search(rm, samples, goal, obstacles, offsets_time, xcomb, prm, nthreads=0, path_cap=4096,
           trace=False, wf_cap=1 << 16, lib=None):
    """Oracle PUMP search over the oracle roadmap rm.

    With trace=True the result also holds the search record: the stored path
    table (path_prev, path_node, path_bucket, path_cost, path_time, path_cp
    over num_paths rows), per-iteration gidx, n_new, n_valid, size_g and
    threshold, and the first wf_cap wavefront entries (wf_prev, wf_edge,
    wf_fate: stored index, -1 CP cutoff, -2 dominated, -3 node 0)."""
    O = lib if lib is not None else load()
    samples = np.ascontiguousarray(samples, np.float32)
    goal = np.ascontiguousarray(goal, np.float32)
    obstacles = np.ascontiguousarray(obstacles, np.float32).reshape(-1, 12)
    obs = obstacles if obstacles.shape[0] else np.zeros((1, 12), np.float32)
    num = samples.shape[0]
    si = SearchIn()
    keep = []

    def keepp(a):
        keep.append(a)
        return a

    si.num = num
    si.samples = fp(samples)
    si.goal = fp(goal)
    si.init_idx = 0
    si.goal_idx = num - 1
    si.row_start = lp(keepp(rm["row_start"]))
    si.edge_dst = ip(keepp(rm["edge_dst"] if rm["n_edges"] else np.zeros(1, np.int32)))
    si.topts_edge = fp(keepp(rm["topts_edge"] if rm["n_edges"] else np.zeros(1, np.float32)))
    si.copts_edge = fp(keepp(rm["copts_edge"] if rm["n_edges"] else np.zeros(1, np.float32)))
    si.h = fp(keepp(rm["h"]))
    si.is_free_edges = bp(keepp(rm["is_free_edges"] if rm["n_edges"] else np.zeros(1, np.uint8)))
    si.is_free_samples = bp(keepp(rm["is_free_samples"]))
    si.wp_offsets = lp(keepp(rm["wp_offsets"]))
    si.as_ = fp(keepp(rm["as_"]))
    si.bs = fp(keepp(rm["bs"]))
    si.max_hs = rm["max_hs"]
    ot = keepp(np.ascontiguousarray(offsets_time, np.float32))
    xc = keepp(np.ascontiguousarray(xcomb, np.float32))
    si.offsets_time = fp(ot)
    si.xcomb = fp(xc)
    si.tmax_steps = ot.shape[0]
    si.n_mc = ot.shape[1]
    si.offset_mult = prm.offset_mult
    si.obstacles = fp(keepp(obs))
    si.n_obs = obstacles.shape[0]
    si.rn = float(rm["rn"])
    si.lambda_ = prm.lambda_
    si.dt = prm.dt
    si.cp_target = prm.cp_target
    si.cp_factor = prm.cp_factor
    si.eps_cost = prm.eps_cost
    si.eps_cp = prm.eps_cp
    si.num_buckets = prm.num_buckets
    si.n_p = prm.n_p
    si.max_itrs = prm.max_itrs
    si.max_path_count = prm.max_path_count if prm.max_path_count > 0 else int(
        (2 * num) * np.sqrt(float(num)))
    si.path_cap = path_cap
    si.nthreads = nthreads
    tr = None
    if trace:
        pcap = int(si.max_path_count)
        nb = int(prm.num_buckets)
        icap = max(int(prm.max_itrs), 1)
        arr = dict(path_prev=np.full(pcap, -9, np.int32), path_node=np.full(pcap, -9, np.int32),
                   path_bucket=np.full(pcap, -9, np.int32), path_cost=np.zeros(pcap, np.float32),
                   path_time=np.zeros(pcap, np.float32), path_cp=np.zeros(pcap, np.float32),
                   itr_gidx=np.zeros(icap, np.int32), itr_n_new=np.zeros(icap, np.int32),
                   itr_n_valid=np.zeros(icap, np.int32), itr_size_g=np.zeros((icap, nb), np.int32),
                   itr_threshold=np.zeros(icap, np.float32),
                   wf_prev=np.zeros(max(wf_cap, 1), np.int32),
                   wf_edge=np.zeros(max(wf_cap, 1), np.int32),
                   wf_fate=np.zeros(max(wf_cap, 1), np.int32))
        keep.append(arr)
        tr = SearchTrace()
        tr.path_cap = pcap
        tr.itr_cap = icap
        tr.wf_cap = wf_cap
        for k, a in arr.items():
            setattr(tr, k, fp(a) if a.dtype == np.float32 else ip(a))
        si.trace = ctypes.pointer(tr)
    so = SearchOut()
    O.po_search(ctypes.byref(si), ctypes.byref(so))
    d = {k: getattr(so, k) for k, _ in SearchOut._fields_ if k != "soln_nodes"}
    d["soln_nodes"] = list(so.soln_nodes)[: so.soln_len]
    if tr is not None:
        n = int(so.num_paths)
        # iterations that reached the compaction (one that found no new path stops before it)
        ni = int(np.count_nonzero(arr["itr_n_new"]))
        for k in ("path_prev", "path_node", "path_bucket", "path_cost", "path_time", "path_cp"):
            d[k] = arr[k][:n].copy()
        for k in ("itr_gidx", "itr_n_new", "itr_n_valid", "itr_size_g", "itr_threshold"):
            d[k] = arr[k][:ni].copy()
        d["wf_len"] = int(tr.wf_len)
        m = min(int(tr.wf_len), wf_cap)
        for k in ("wf_prev", "wf_edge", "wf_fate"):
            d[k] = arr[k][:m].copy()
    return d


def fp64_path_cp(nodes, samples, obstacles, xcomb, offset_mult, dt=0.05, cap=8192):
    """FP64 oracle CP of a node chain: FP64 tau*, FP64 path, FP64 MC with the
    same noise draws.  Returns (cp, hits, n_mc, T)."""
    O = load()
    xs = np.ascontiguousarray(samples[np.asarray(nodes)], np.float32)
    path = np.zeros((cap, 6), np.float64)
    T = O.po64_find_optimal_path(dt, dp(path), fp(xs), xs.shape[0], cap)
    obstacles = np.ascontiguousarray(obstacles, np.float32).reshape(-1, 12)
    obs = obstacles if obstacles.shape[0] else np.zeros((1, 12), np.float32)
    xc = np.ascontiguousarray(xcomb, np.float32)
    n_mc, tsteps = xc.shape[0], xc.shape[1]
    hits = O.po64_mccp_hits(fp(obs), obstacles.shape[0], fp(xc), tsteps, n_mc, offset_mult,
                            dp(path), T)
    return hits / n_mc, hits, n_mc, T


def smooth(search_res, samples, obstacles, xcomb, prm, path_cap=4096, disc_cap=1 << 16):
    """Oracle smoothing and controller discretisation (PUMP.cu:406-568) of a
    search result (a dict with status, soln_nodes, soln_cost, soln_cp)."""
    O = load()
    nodes = list(search_res["soln_nodes"])
    if search_res["status"] != 0 or len(nodes) < 2:
        return {"status": 1}
    samples = np.ascontiguousarray(samples, np.float32)
    xs = np.ascontiguousarray(samples[nodes], np.float32)
    obstacles = np.ascontiguousarray(obstacles, np.float32).reshape(-1, 12)
    obs = obstacles if obstacles.shape[0] else np.zeros((1, 12), np.float32)
    xc = np.ascontiguousarray(xcomb, np.float32)
    n_mc, tmax_steps = xc.shape[0], xc.shape[1]
    best = np.zeros((path_cap, 6), np.float32)
    rows = np.zeros(1, np.int32)
    cc = np.zeros(2, np.float32)
    xs_s = np.zeros_like(xs)
    disc = np.zeros((disc_cap, 6), np.float32)
    drows = np.zeros(1, np.int32)
    trace = np.zeros((30, 3), np.float32)
    itrs = np.zeros(1, np.int32)
    rc = O.po_smooth(fp(xs), xs.shape[0], np.float32(search_res["soln_cost"]),
                     np.float32(search_res["soln_cp"]), fp(obs), obstacles.shape[0], fp(xc),
                     tmax_steps, n_mc, prm.offset_mult, prm.dt, prm.cp_target, path_cap, fp(best),
                     ip(rows), fp(cc), fp(xs_s), disc_cap, fp(disc), ip(drows), fp(trace),
                     ip(itrs))
    if rc != 0:
        return {"status": rc}
    k = int(itrs[0])
    return {"status": 0, "itrs": k, "trace": trace[:k].copy(), "path": best[: rows[0]].copy(),
            "cost": float(cc[0]), "cp": float(cc[1]), "xs_smooth": xs_s,
            "discretized": disc[: drows[0]].copy()}
