"""Benchmark of the planning hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 1]

Workload (the default, --config 1 of pump_amd.configs): NUM=16000 samples,
30 boxes, ~80 neighbours per sample, W_e = clamp(ceil(tau/0.05), 2, 64)
waypoints per edge, 2000 particles, noise multiplier 0.03; synthetic seeded
inputs.

Setup (untimed): problem + noise tables, then the roadmap precompute on the
device (all-pairs 2pBVP, rn selection, edge numbering, heuristic,
sampleFree, waypoint offsets, fused edge pass, half-space cache).

Step (timed K times, after W warmups): one edge pass -- kernel (b), waypoint
generation fused with the segment/AABB test -- over every candidate edge.
    value = edge-checks/s = E / (seconds per step)
Also measured in the same run, each over its own timed repetitions:
    * the same edge pass through the reference-shaped pair of calls
      (pump_discretize_edges, then pump_free_edges over the stored
      waypoints), with its flags compared to the fused pass's;
    * kernel (c): one HSMC propagate + CP pass in which every free edge is
      expanded once from a synthetic parent path (particle-steps/s): the
      drop-in pair (byte validity, the reference layout), the batched byte
      and bit-mask forms and the copy-on-write forms, FP32 and split-BF16;
    * the FP32-waypoint form of the edge pass (flags compared);
    * the PUMP search with its bookkeeping on the device and on the host
      (same path), and the smoothing of its result;
    * full Monte Carlo CP (MCCP) of the search's 64 deepest paths
      (segment-checks/s);
    * kernel (a): the all-pairs 2pBVP solve (pair-solves/s), and its
      algorithmic flops (from the counted bisection steps) against the
      FP64 vector peak;
    * the uneven batch workload (256 problems, NUM log-uniform in
      [1000, 8000], 5-50 boxes): the batched build's pair pass and edge pass +
      half-space cache (one launch each for all instances, hipGraph replays)
      against the same passes launched instance by instance, and the HSMC
      step of every instance (one launch per stage, hipGraph replays);
    * cpu_baseline: the CPU oracle's edge pass on a bounded sample of the
      same edges with all host cores, beside it on one thread, and its pair
      solves, HSMC step and full MC on bounded samples, each checked against
      the GPU's results.
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK = 8.0e12          # B/s, MI355X
FP32_PEAK = 157.3e12       # FLOP/s, MI355X vector FP32
FP64_PEAK = 78.6e12        # FLOP/s, MI355X vector FP64
METRIC = "edge-checks/s + MC particle-steps/s, batched split-BF16 vs plain FP32, 1 MI355X"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--num", type=int, default=0, help="override NUM (smaller runs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--hsmc-reps", type=int, default=3)
    ap.add_argument("--skip-hsmc", action="store_true")
    ap.add_argument("--batch", type=int, default=256,
                    help="instances of the uneven batch workload (0: skip)")
    ap.add_argument("--hs-sweep", action="store_true",
                    help="only the half-space density sweep of the MC pass (FP32 against "
                         "split-BF16), one JSON line per density")
    return ap.parse_args()


def edge_pass(P, d, n_obs, E, stream):
    P.check("pump_edges_discretize_check", P.lib.pump_edges_discretize_check(
        d.samples, d.edge_src, d.edge_dst, d.topts_edge, E, d.wp_offsets, 0, d.obstacles,
        n_obs, d.disc, d.is_free_edges, stream.h))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the split of the "
              "work across GPUs is not run here.", file=sys.stderr)
        sys.exit(2)
    import pump_amd as P
    from pump_amd import configs

    assert P.lib.pump_dev_count() >= 1, "no GPU visible"
    if args.hs_sweep:
        hs_sweep(P, configs, args.hsmc_reps)
        return
    log("device:", P.device_name())
    over = {"num": args.num} if args.num else {}
    cfg = configs.config(args.config, **over)
    prm = cfg["params"]
    t0 = time.perf_counter()
    samples, obstacles = P.gen_problem(cfg["spec"])
    ot, xc = P.make_noise(cfg["noise"])
    log(f"problem + noise tables: {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    rm = P.Roadmap(samples, obstacles, samples[-1], ot, xc, prm)
    info = rm.info
    log(f"roadmap build {time.perf_counter() - t0:.1f} s: pairs {info.n_pairs} "
        f"({info.ms_pairs:.1f} ms), E={info.n_edges}, waypoints={info.n_waypoints}, "
        f"free={info.n_free_edges}, rn={info.rn:.4f}")
    d = rm.device()
    n_obs = int(obstacles.shape[0])
    E = int(info.n_edges)
    n_wp = int(info.n_waypoints)
    stream = P.Stream()
    timer = P.Timer(stream)

    # ---------------- step: kernel (b) edge pass over all candidate edges --
    for _ in range(args.warmup):
        edge_pass(P, d, n_obs, E, stream)
    stream.sync()
    P.sync()
    wall0 = time.perf_counter()
    timer.start()
    for _ in range(args.steps):
        edge_pass(P, d, n_obs, E, stream)
    ms_total = timer.stop()
    P.sync()
    wall = time.perf_counter() - wall0
    ms_step = ms_total / args.steps
    value = E / (ms_step / 1e3)
    alg_bytes = 24 * n_wp + 33 * E          # waypoints + per-edge states/tau/index/flag
    achieved = alg_bytes / (ms_step / 1e3)
    log(f"edge pass: {ms_step:.3f} ms/step, {value:.3e} edge-checks/s, "
        f"{achieved / 1e9:.1f} GB/s algorithmic, wall {wall:.3f} s for {args.steps} steps")

    # the same pass through the reference-shaped pair of calls (discretize
    # every edge, then freeEdges over the stored waypoints): the plain path
    fused_flags = rm.flags()["is_free_edges"]

    def plain_pass():
        P.check("pump_discretize_edges", P.lib.pump_discretize_edges(
            d.samples, d.edge_src, d.edge_dst, d.topts_edge, E, d.wp_offsets, 0, d.disc,
            stream.h))
        P.check("pump_free_edges_var", P.lib.pump_free_edges_var(
            d.obstacles, n_obs, d.wp_offsets, d.disc, d.is_free_edges, E, stream.h))

    plain_pass()
    stream.sync()
    timer.start()
    for _ in range(args.steps):
        plain_pass()
    ms_plain = timer.stop() / args.steps
    plain = {"ms": ms_plain, "edge_checks_per_s": E / (ms_plain / 1e3),
             "fused_vs_plain": ms_plain / ms_step,
             "flags_equal": bool(np.array_equal(rm.flags()["is_free_edges"], fused_flags))}
    log(f"plain edge pass (discretize + freeEdges): {ms_plain:.3f} ms, "
        f"fused is {plain['fused_vs_plain']:.2f}x, flags equal {plain['flags_equal']}")

    # chosen form: the fused pass with FP32 waypoints (within a stated bound
    # of the reference's) and exact flags, into separate output arrays
    disc32 = P.DeviceArray(shape=(n_wp, 6), dtype=np.float32, fill=0)
    flags32 = P.DeviceArray(shape=(E,), dtype=np.uint8, fill=9)
    retests = P.DeviceArray(shape=(1,), dtype=np.uint64, fill=0)

    def fp32_pass():
        P.check("pump_edges_discretize_check_fp32", P.lib.pump_edges_discretize_check_fp32(
            d.samples, d.edge_src, d.edge_dst, d.topts_edge, E, d.wp_offsets, 0, d.obstacles,
            n_obs, disc32.p, flags32.p, retests.p, stream.h))

    fp32_pass()
    stream.sync()
    n_retests = int(retests.numpy()[0])
    timer.start()
    for _ in range(args.steps):
        fp32_pass()
    ms_fp32 = timer.stop() / args.steps
    edges_fp32 = {"ms": ms_fp32, "edge_checks_per_s": E / (ms_fp32 / 1e3),
                  "achieved_gbs": alg_bytes / (ms_fp32 / 1e3) / 1e9,
                  "frac_hbm_peak": alg_bytes / (ms_fp32 / 1e3) / HBM_PEAK,
                  "vs_exact_fused": ms_step / ms_fp32, "vs_plain": ms_plain / ms_fp32,
                  "exact_retests_per_pass": n_retests,
                  "traffic": (edge_traffic("fp32_form")
                              if not args.num and args.config == 1 else None),
                  "flags_equal": bool(np.array_equal(flags32.numpy(), fused_flags))}
    log(f"FP32-waypoint edge pass: {ms_fp32:.3f} ms, {E / (ms_fp32 / 1e3):.3e} edge-checks/s, "
        f"{edges_fp32['frac_hbm_peak'] * 100:.1f}% of HBM peak, {n_retests} exact re-tests, "
        f"flags equal {edges_fp32['flags_equal']}")
    disc32.free()

    # a plain device copy of known size: calibrates the HBM byte counters of a
    # profiled run (FETCH_SIZE / WRITE_SIZE of the copy vs its 2 x 256 MiB)
    calib = P.DeviceArray(shape=(2, 256 << 20), dtype=np.uint8, fill=7)
    P.check("pump_dev_copy", P.lib.pump_dev_copy(calib.ptr + (256 << 20), calib.ptr, 256 << 20))
    P.sync()
    calib.free()

    # ---------------- kernel (c): HSMC pass over every free edge -----------
    hsmc, hsmc_cp = None, None
    if not args.skip_hsmc:
        hsmc, hsmc_cp = hsmc_pass(P, rm, d, prm, info, stream, args.hsmc_reps)

    # ---------------- the PUMP search, bookkeeping on the device and host --
    search = search_pass(P, rm, prm)
    mccp, mccp_data = (mccp_pass(P, rm, prm, samples, int(search["num_paths"]), 5)
                       if search["num_paths"] > 1 else (None, None))

    # ---------------- kernel (a): all-pairs 2pBVP --------------------------
    # The build's own pair pass is the process's first GPU work (clocks still
    # ramping); the pass is timed again here into separate tables, after the
    # passes above, and those tables are checked against the build's
    # (sampled rows when the build did not keep them).
    g_c = P.DeviceArray(shape=(int(info.n_pairs),), dtype=np.float32, fill=0)
    g_t = P.DeviceArray(shape=(int(info.n_pairs),), dtype=np.float32, fill=0)

    def pair_pass():
        P.check("pump_fill_copts_topts", P.lib.pump_fill_copts_topts(
            d.samples, int(info.num), np.float32(prm.tmax), g_c.p, g_t.p, stream.h))

    pair_pass()
    stream.sync()
    reps = max(2, min(args.steps, 5))
    timer.start()
    for _ in range(reps):
        pair_pass()
    ms_pairs = timer.stop() / reps
    pairs = {"pair_solves_per_s": info.n_pairs / (ms_pairs / 1e3), "ms": ms_pairs,
             "ms_in_build": info.ms_pairs, "n_pairs": int(info.n_pairs), "reps": reps}
    if info.pair_steps > 0:
        # algorithmic flops: a cost evaluation is 40 FP64 + 14 FP32 operations
        # (2pBVP.cu:97-107 as written: no fused multiply-adds), a bisection
        # step two evaluations and the dcost quotient (110); a pair's final
        # step needs one evaluation (-56).  Peak: FP64 vector, which is also
        # the unpacked FP32 rate.
        flops = 110.0 * info.pair_steps - 56.0 * info.n_pairs
        tflops = flops / (ms_pairs / 1e3) / 1e12
        pairs["roofline"] = {"bound": "valu", "achieved": tflops, "peak": FP64_PEAK / 1e12,
                             "unit": "TFLOP/s", "frac": tflops * 1e12 / FP64_PEAK,
                             "kernel": "k_fill_pairs_persistent (pump_fill_copts_topts)",
                             "steps_per_pair": info.pair_steps / info.n_pairs,
                             "flops_per_pair": flops / info.n_pairs}
    # the re-run tables against the build's: its edge list holds, row by row,
    # the table's entries below rn and their destinations
    e = rm.edges()
    rows = np.linspace(0, int(info.num) - 1, 8).astype(np.int64)
    nm1 = int(info.num) - 1
    same = True
    for r in rows.tolist():
        cr = np.empty(nm1, np.float32)
        P.check("pump_dev_download", P.lib.pump_dev_download(
            cr.ctypes.data, g_c.ptr + r * nm1 * 4, nm1 * 4))
        lo, hi = int(e["row_start"][r]), int(e["row_start"][r + 1])
        jj = np.arange(nm1)
        sel = cr < np.float32(info.rn)
        same &= bool(np.array_equal(cr[sel], e["copts_edge"][lo:hi]))
        same &= bool(np.array_equal((jj + (jj >= r))[sel], e["edge_dst"][lo:hi]))
    pairs["rows_match_build"] = same
    log(f"pair pass: {ms_pairs:.2f} ms ({info.ms_pairs:.2f} ms in the build), "
        f"{pairs['pair_solves_per_s']:.3e} pair-solves/s"
        + (f", {pairs['roofline']['steps_per_pair']:.2f} bisection steps per pair, "
           f"{pairs['roofline']['achieved']:.1f} TFLOP/s "
           f"({pairs['roofline']['frac'] * 100:.1f}% of FP64 peak)" if "roofline" in pairs else "")
        + f", rows match the build {same}")

    # ---------------- the uneven batch workload ---------------------------
    batch = None
    if args.batch > 0:
        batch = batch_pass(P, args.batch, args.hsmc_reps)

    # ---------------- CPU baseline (oracle, bounded sample) ----------------
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(rm, samples, obstacles, prm, args.cpu_seconds)
        cpu["one_thread"] = cpu_edges_one_thread(rm, samples, obstacles, args.cpu_seconds / 3)
        cpu["pair_solves"] = cpu_pairs(rm, samples, prm, g_c, g_t, args.cpu_seconds / 3)
        if hsmc_cp is not None:
            cpu["particle_steps"] = cpu_particles(rm, prm, ot, hsmc_cp, args.cpu_seconds / 3)
        if mccp_data is not None:
            cpu["mccp_segment_checks"] = cpu_mccp(prm, obstacles, xc, *mccp_data,
                                                  args.cpu_seconds / 3)

    g_c.free()
    g_t.free()
    traffic = edge_traffic(None) if not args.num and args.config == 1 else None

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "edge-checks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (2pBVP cost/waypoints in f64 as the reference's host build)",
        "data": "synthetic (seeded Philox draws)",
        "config": {"workload": cfg["name"], "num": int(info.num), "n_obs": int(info.n_obs),
                   "n_edges": E, "n_waypoints": n_wp, "n_free_edges": int(info.n_free_edges),
                   "n_p": int(prm.n_p), "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "kernel": "k_edges<true,true,128> (pump_edges_discretize_check)",
                     "alg_bytes_per_launch": alg_bytes,
                     # the two ceilings in edge-checks/s: HBM (algorithmic bytes at
                     # 8 TB/s) and FP64 VALU (~90 flops per interior waypoint of
                     # pathPoint, 2pBVP.cu:117-131, at 78.6 TFLOP/s; endpoints are copies)
                     "hbm_ceiling_edge_checks_per_s": E * HBM_PEAK / alg_bytes,
                     "valu_ceiling_edge_checks_per_s":
                         E * FP64_PEAK / (90.0 * max(n_wp - 2 * E, 1))},
        "cpu_baseline": cpu,
        "edges_plain": plain,
        "edges_fp32": edges_fp32,
        "hsmc": hsmc,
        "pairs": pairs,
        "search": search,
        "mccp": mccp,
        "batch": batch,
        "pipeline_ms": {k: getattr(info, k) for k in (
            "ms_pairs", "ms_select", "ms_edges_build", "ms_nn", "ms_sample_free",
            "ms_edge_pass", "ms_hs", "ms_total")},
    }
    print(json.dumps(line), flush=True)
    rm.close()


def synthetic_wavefront(rm):
    """The HSMC workload: every free edge (up to 1M) expanded once, child i
    from parent i % 4096; parent times uniform in [0, 2] s."""
    fl = rm.flags()
    free = np.nonzero(fl["is_free_edges"])[0].astype(np.int32)
    cap = int(os.environ.get("PUMP_BENCH_MAX_NEW", "1000000"))
    free = free[:cap]
    n_par = 4096
    rng = np.random.default_rng(0)
    times = rng.uniform(0, 2.0, n_par).astype(np.float32)
    prev = (np.arange(free.shape[0]) % n_par).astype(np.int32)
    return dict(free=free, prev=prev, times=times, n_par=n_par, wp_offsets=fl["wp_offsets"])


def hsmc_pass(P, rm, d, prm, info, stream, reps):
    """Every free edge expanded once from one of 4096 synthetic parent paths
    (times uniform in [0, 2] s, all particles valid).  Returns the figures
    and the drop-in form's CPs (the CPU baseline's check)."""
    n_p, n_mc, T = prm.n_p, info.n_mc, info.tmax_steps
    wf = synthetic_wavefront(rm)
    free, prev, times, n_par = wf["free"], wf["prev"], wf["times"], wf["n_par"]
    n_new = int(free.shape[0])
    wo = wf["wp_offsets"]
    W = (wo[free + 1] - wo[free]).astype(np.int64)
    steps = int(n_p) * int(W.sum())
    counts = rm.halfspaces()["count_stored"]
    # stored half-spaces over the expanded waypoints: 3 MACs = 6 flops each
    cs = np.concatenate([[0], np.cumsum(counts.astype(np.int64))])
    hs_total = int((cs[wo[free + 1]] - cs[wo[free]]).sum())
    flops = 6.0 * n_p * hs_total
    # waypoints that hold a half-space, and the paths whose edge has one
    marked = np.concatenate([[0], np.cumsum((counts > 0).astype(np.int64))])
    marked_per_path = marked[wo[free + 1]] - marked[wo[free]]
    n_marked = int(marked_per_path.sum())
    g_edge, g_prev = P.DeviceArray(free), P.DeviceArray(prev)
    g_times = P.DeviceArray(times)
    g_cp = P.DeviceArray(shape=(n_new,), dtype=np.float32, fill=0)
    # room for every marked (path, waypoint) pair: the binned form at any density
    ws_bytes = int(P.lib.pump_propagate_workspace_pairs(n_new, T, n_marked))
    g_ws = P.DeviceArray(shape=(ws_bytes,), dtype=np.uint8, fill=0)
    # H_w histogram: expanded waypoints by their number of stored
    # half-spaces (waypoint k of edge e counts once per child expanding e)
    edge_of_wp = np.repeat(np.arange(len(wo) - 1), np.diff(wo))
    mult = np.bincount(free, minlength=len(wo) - 1)
    hw_hist = np.bincount(np.minimum(counts, info.max_hs).astype(np.int64),
                          weights=mult[edge_of_wp], minlength=int(info.max_hs) + 1).astype(np.int64)
    out = {"n_new_paths": n_new, "particle_steps": steps, "stored_halfspace_tests": n_p * hs_total,
           "hw_histogram": [int(x) for x in hw_hist],
           "paths_with_halfspaces": int((marked_per_path > 0).sum()),
           "marked_waypoints": n_marked,
           # offsets read by the half-space tests: 12 B per (particle, marked waypoint)
           "offset_bytes": 12 * n_p * n_marked}
    timer = P.Timer(stream)
    cps = {}

    # drop-in: the reference's call pair (propagateValidParticles in place on
    # one byte array, parents first, then calculateCP_HSMC), FP32
    inplace = P.DeviceArray(shape=(n_par + n_new, n_p), dtype=np.uint8, fill=1)
    cp_in = P.DeviceArray(shape=(n_par + n_new,), dtype=np.float32, fill=0)

    def run_dropin():
        P.check("pump_propagate_valid_particles", P.lib.pump_propagate_valid_particles(
            n_new, 0, d.wp_offsets, g_edge.p, g_prev.p, n_par, d.as_, d.bs, d.count_stored,
            g_times.p, d.topts_edge, prm.dt, info.max_hs, d.offsets_time, T, prm.offset_mult,
            n_mc, inplace.p, n_p, None, stream.h))
        P.check("pump_calculate_cp_hsmc", P.lib.pump_calculate_cp_hsmc(
            n_new, n_par, cp_in.p, inplace.p, n_p, stream.h))

    # batched, byte validity (separate parent / child arrays), FP32
    src = P.DeviceArray(shape=(n_par, n_p), dtype=np.uint8, fill=1)
    dst = P.DeviceArray(shape=(n_new, n_p), dtype=np.uint8, fill=0)
    cp_bytes = P.DeviceArray(shape=(n_new,), dtype=np.float32, fill=0)

    def run_bytes():
        P.check("pump_propagate_cp", P.lib.pump_propagate_cp(
            n_new, 0, d.wp_offsets, g_edge.p, g_prev.p, d.as_, d.bs, d.edge_hs_mask, g_times.p,
            d.topts_edge, prm.dt, info.max_hs, d.offsets_time, T, prm.offset_mult, n_mc, src.p,
            dst.p, 0, n_p, cp_bytes.p, 0, None, g_ws.p, ws_bytes, stream.h))

    # batched, bit-mask validity, FP32
    nw = (n_p + 63) // 64
    bsrc = P.DeviceArray(shape=(n_par, nw), dtype=np.uint64, fill=0)
    P.check("pump_setup_valid_bits", P.lib.pump_setup_valid_bits(bsrc.p, n_p, n_par, None))
    bdst = P.DeviceArray(shape=(n_new, nw), dtype=np.uint64, fill=0)

    def run_bits():
        P.check("pump_propagate_cp_bits", P.lib.pump_propagate_cp_bits(
            n_new, 0, d.wp_offsets, g_edge.p, g_prev.p, d.as_, d.bs, d.edge_hs_mask, g_times.p,
            d.topts_edge, prm.dt, info.max_hs, d.offsets_time, T, prm.offset_mult, n_mc, bsrc.p,
            bdst.p, 0, n_p, g_cp.p, 0, 0, None, None, None, g_ws.p, ws_bytes, stream.h))

    # batched, bit-mask validity, split-BF16 MFMA half-space tests (FP32 band re-test)
    bdst2 = P.DeviceArray(shape=(n_new, nw), dtype=np.uint64, fill=0)
    cp_split = P.DeviceArray(shape=(n_new,), dtype=np.float32, fill=0)
    band = P.DeviceArray(shape=(1,), dtype=np.uint64, fill=0)

    def run_split():
        P.check("pump_propagate_cp_bits", P.lib.pump_propagate_cp_bits(
            n_new, 0, d.wp_offsets, g_edge.p, g_prev.p, d.as_, d.bs, d.edge_hs_mask, g_times.p,
            d.topts_edge, prm.dt, info.max_hs, d.offsets_time, T, prm.offset_mult, n_mc, bsrc.p,
            bdst2.p, 0, n_p, cp_split.p, 0, 1, d.offsets_split, band.p, None, g_ws.p, ws_bytes,
            stream.h))

    # copy-on-write forms: children that meet no half-space keep their
    # parent's row and CP (parents all valid: CP 0); only the listed children's
    # rows are written, the rows a caller keeps are gathered afterwards
    g_pcp = P.DeviceArray(shape=(n_par,), dtype=np.float32, fill=0)
    row_src = P.DeviceArray(shape=(n_new,), dtype=np.int32, fill=0)
    cdst = P.DeviceArray(shape=(n_new, nw), dtype=np.uint64, fill=0)
    cp_cow = P.DeviceArray(shape=(n_new,), dtype=np.float32, fill=0)
    cp_cow_split = P.DeviceArray(shape=(n_new,), dtype=np.float32, fill=0)
    cdst_b = P.DeviceArray(shape=(n_new, n_p), dtype=np.uint8, fill=0)
    cp_cow_b = P.DeviceArray(shape=(n_new,), dtype=np.float32, fill=0)

    def cow(bits, precision, srcp, dstp, cpp):
        P.check("pump_propagate_cp_cow", P.lib.pump_propagate_cp_cow(
            bits, n_new, 0, d.wp_offsets, g_edge.p, g_prev.p, d.as_, d.bs, d.edge_hs_mask,
            g_times.p, d.topts_edge, prm.dt, info.max_hs, d.offsets_time, T, prm.offset_mult,
            n_mc, srcp, g_pcp.p, dstp, 0, row_src.p, n_p, cpp, 0, precision,
            d.offsets_split if precision else None, None, None, g_ws.p, ws_bytes, stream.h))

    # particles per wave of the binned half-space test, tuned per
    # configuration on this workload (the tuned shape then serves every
    # batched form of the configuration below)
    import ctypes
    ppw = {}
    for name, bits, prec, srcp, dstp, cpp in (
            ("fp32_bits", 1, 0, bsrc.p, cdst.p, cp_cow.p),
            ("split_bf16_bits", 1, 1, bsrc.p, cdst.p, cp_cow_split.p),
            ("fp32_bytes", 0, 0, src.p, cdst_b.p, cp_cow_b.p)):
        chosen = ctypes.c_int(0)
        P.check("pump_hsmc_autotune_cow", P.lib.pump_hsmc_autotune_cow(
            bits, n_new, 0, d.wp_offsets, g_edge.p, g_prev.p, d.as_, d.bs, d.edge_hs_mask,
            g_times.p, d.topts_edge, prm.dt, info.max_hs, d.offsets_time, T, prm.offset_mult,
            n_mc, srcp, g_pcp.p, dstp, 0, row_src.p, n_p, cpp, 0, prec,
            d.offsets_split if prec else None, None, None, g_ws.p, ws_bytes, stream.h,
            ctypes.byref(chosen)))
        ppw[name] = chosen.value
    out["particles_per_wave"] = ppw
    log(f"hsmc autotune, particles per wave: {ppw}")

    forms = (("dropin_fp32_bytes", run_dropin, lambda: cp_in.numpy()[n_par:]),
             ("batched_fp32_bytes", run_bytes, lambda: cp_bytes.numpy()),
             ("batched_fp32_bits", run_bits, lambda: g_cp.numpy()),
             ("batched_split_bf16_bits", run_split, lambda: cp_split.numpy()),
             ("cow_fp32_bytes", lambda: cow(0, 0, src.p, cdst_b.p, cp_cow_b.p),
              lambda: cp_cow_b.numpy()),
             ("cow_fp32_bits", lambda: cow(1, 0, bsrc.p, cdst.p, cp_cow.p),
              lambda: cp_cow.numpy()),
             ("cow_split_bf16_bits", lambda: cow(1, 1, bsrc.p, cdst.p, cp_cow_split.p),
              lambda: cp_cow_split.numpy()))
    for name, fn, result in forms:
        fn()
        stream.sync()
        cps[name] = result()
        timer.start()
        for _ in range(reps):
            fn()
        ms = timer.stop() / reps
        out[name] = {"ms": ms, "particle_steps_per_s": steps / (ms / 1e3),
                     "achieved_tflops": flops / (ms / 1e3) / 1e12,
                     "frac_fp32_peak": flops / (ms / 1e3) / FP32_PEAK}
        log(f"hsmc {name}: {ms:.3f} ms, {steps / (ms / 1e3):.3e} particle-steps/s")
    ref = cps["dropin_fp32_bytes"]
    out["forms_agree"] = bool(all(np.array_equal(ref, v) for v in cps.values()))
    # the copy-on-write rows: every child gathered (the upper bound of what a
    # search keeps) must equal the bit form's rows; the gather is timed too
    full = P.DeviceArray(shape=(n_new, nw), dtype=np.uint64, fill=0)

    def gather_all():
        P.check("pump_gather_rows_cow", P.lib.pump_gather_rows_cow(
            cdst.p, bsrc.p, row_src.p, None, n_new, full.p, 0, nw * 8, 1, stream.h))

    cow(1, 0, bsrc.p, cdst.p, cp_cow.p)
    gather_all()
    stream.sync()
    out["cow_rows_agree"] = bool(np.array_equal(full.numpy(), bdst.numpy()))
    timer.start()
    for _ in range(reps):
        gather_all()
    out["cow_gather_all_bits_ms"] = timer.stop() / reps
    rsrc = row_src.numpy()
    out["cow_rows_written"] = int((rsrc >= 0).sum())
    out["mean_cp"] = float(ref.mean())
    band.upload(np.zeros(1, np.uint64))
    run_split()
    stream.sync()
    out["band_retests_per_pass"] = int(band.numpy()[0])
    out["split_bf16_vs_dropin"] = (out["dropin_fp32_bytes"]["ms"]
                                   / out["batched_split_bf16_bits"]["ms"])
    out["fp32_bits_vs_dropin"] = (out["dropin_fp32_bytes"]["ms"]
                                  / out["batched_fp32_bits"]["ms"])
    out["cow_fp32_bits_vs_dropin"] = (out["dropin_fp32_bytes"]["ms"]
                                      / out["cow_fp32_bits"]["ms"])
    out["cow_split_bf16_vs_dropin"] = (out["dropin_fp32_bytes"]["ms"]
                                       / out["cow_split_bf16_bits"]["ms"])
    best = min(("batched_fp32_bits", "batched_split_bf16_bits", "cow_fp32_bits",
                "cow_split_bf16_bits"), key=lambda k: out[k]["ms"])
    out["best_form"] = best
    out["value"] = out[best]["particle_steps_per_s"]
    out["unit"] = "particle-steps/s"
    ms_best = out[best]["ms"]
    out["stored_halfspace_tests_per_s"] = n_p * hs_total / (ms_best / 1e3)
    # the MC pass's HBM bytes in the chosen form: per listed child its parent
    # row read and its own row written, per child the CP and row source
    # written and its edge, parent and time read, and each offsets row the
    # tests need once (12 bytes per particle and time index; at most the
    # table's T rows, whatever the number of marked waypoints)
    row_bytes = nw * 8 if "bits" in best else n_p
    listed = out["cow_rows_written"] if best.startswith("cow") else n_new
    offset_rows = min(T, n_marked)
    mc_bytes = 2 * listed * row_bytes + 8 * n_new + 12 * n_new + 12 * n_p * offset_rows
    out["mc"] = {"form": best, "alg_bytes_per_pass": mc_bytes,
                 "achieved_gbs": mc_bytes / (ms_best / 1e3) / 1e9,
                 "frac_hbm_peak": mc_bytes / (ms_best / 1e3) / HBM_PEAK,
                 "frac_fp32_peak": out[best]["frac_fp32_peak"],
                 "offset_rows_bound": offset_rows,
                 "counters": mc_counters()}
    return out, ref


def hs_sweep(P, configs, reps):
    """The batched MC pass (bit validity) in FP32 and in split-BF16 over
    synthetic half-space tables of rising density: 200,000 edges of 8
    waypoints, each expanded once from one of 4096 all-valid parents, 2000
    particles, the noise table of the NUM=16000 workload; a fraction f of
    the waypoints holds k live half-spaces (a ~ N(0, 0.05^2) per component,
    b = a.a / 2 rescaled as the cache stores it).  The workspace holds every
    (path, waypoint) pair, so the binned form runs at every density; each
    form is timed at each particles-per-wave choice and reported at its
    best.  Both forms must give the same CPs."""
    cfg = configs.config(1)
    prm = cfg["params"]
    ot, _ = P.make_noise(cfg["noise"])
    T, n_mc = ot.shape[0], ot.shape[1]
    n_edges, W, max_hs, n_p, n_par = 200000, 8, 8, 2000, 4096
    off = (np.arange(n_edges + 1, dtype=np.int64) * W)
    rng = np.random.default_rng(5)
    topts = rng.uniform(0.3, 1.2, n_edges).astype(np.float32)
    times = rng.uniform(0, 2.0, n_par).astype(np.float32)
    edge = np.arange(n_edges, dtype=np.int32)
    prev = (np.arange(n_edges) % n_par).astype(np.int32)
    g = {"off": P.DeviceArray(off), "topts": P.DeviceArray(topts), "times": P.DeviceArray(times),
         "edge": P.DeviceArray(edge), "prev": P.DeviceArray(prev), "ot": P.DeviceArray(ot)}
    split_bytes = int(P.lib.pump_split_offsets_bytes(T, n_mc))
    split = P.DeviceArray(shape=(split_bytes,), dtype=np.uint8, fill=0)
    P.check("pump_split_offsets", P.lib.pump_split_offsets(g["ot"].p, T, n_mc, split.p, None))
    nw = (n_p + 63) // 64
    src = P.DeviceArray(shape=(n_par, nw), dtype=np.uint64, fill=0)
    P.check("pump_setup_valid_bits", P.lib.pump_setup_valid_bits(src.p, n_p, n_par, None))
    # room for every (path, waypoint) pair in the binned form
    ws_bytes = int(P.lib.pump_propagate_workspace_pairs(n_edges, T, n_edges * W))
    ws = P.DeviceArray(shape=(ws_bytes,), dtype=np.uint8, fill=0)
    stream = P.Stream()
    timer = P.Timer(stream)
    for f, k in ((0.01, 1), (0.03, 1), (0.1, 1), (0.1, 4), (0.3, 4), (1.0, 4), (1.0, 8)):
        n_wp = n_edges * W
        a = rng.normal(0, 0.05, (n_wp, max_hs, 3)).astype(np.float32)
        b = (0.5 * (a.astype(np.float64) ** 2).sum(-1)).astype(np.float32)
        b[:, k:] = 1111.0
        b[rng.uniform(size=n_wp) >= f] = 1111.0
        ga, gb = P.DeviceArray(a.reshape(-1, 3)), P.DeviceArray(b.reshape(-1))
        marked = int((b < 1000).any(axis=1).sum())
        tests = int((b < 1000).sum()) * n_p
        mask = np.zeros(n_edges, np.uint64)
        has = (b < 1000).any(axis=1).reshape(n_edges, W)
        for kk in range(W):
            mask |= np.where(has[:, kk], np.uint64(1) << np.uint64(kk), np.uint64(0))
        gm = P.DeviceArray(mask)
        row = {"f": f, "k": k, "marked_waypoints": marked, "halfspace_tests": tests}
        cps = {}
        for name, prec in (("fp32", 0), ("split_bf16", 1)):
            dst = P.DeviceArray(shape=(n_edges, nw), dtype=np.uint64, fill=0)
            cp = P.DeviceArray(shape=(n_edges,), dtype=np.float32, fill=0)
            band = P.DeviceArray(shape=(1,), dtype=np.uint64, fill=0)

            def run():
                P.check("pump_propagate_cp_bits", P.lib.pump_propagate_cp_bits(
                    n_edges, 0, g["off"].p, g["edge"].p, g["prev"].p, ga.p, gb.p, gm.p,
                    g["times"].p, g["topts"].p, prm.dt, max_hs, g["ot"].p, T, prm.offset_mult,
                    n_mc, src.p, dst.p, 0, n_p, cp.p, 0, prec, split.p if prec else None,
                    band.p if prec else None, None, ws.p, ws_bytes, stream.h))

            # each form at its best particles per wave
            times_ms = {}
            for ppw in (64, 128, 256, 512):
                P.check("pump_hsmc_set_particles_per_wave",
                        P.lib.pump_hsmc_set_particles_per_wave(1, prec, n_p, max_hs, ppw))
                # the first timed configuration also warms the clocks up
                for _ in range(1 if cps or times_ms or f != 0.01 else 20):
                    run()
                stream.sync()
                c_now = cp.numpy()
                if name in cps:
                    assert np.array_equal(c_now, cps[name]), (name, ppw)
                cps[name] = c_now
                timer.start()
                for _ in range(reps):
                    run()
                times_ms[ppw] = timer.stop() / reps
            P.check("pump_hsmc_set_particles_per_wave",
                    P.lib.pump_hsmc_set_particles_per_wave(1, prec, n_p, max_hs, 0))
            best = min(times_ms, key=times_ms.get)
            ms = times_ms[best]
            if prec:
                band.upload(np.zeros(1, np.uint64))
                run()
                stream.sync()
                row["band_retests"] = int(band.numpy()[0])
            row[name + "_ms"] = ms
            row[name + "_particles_per_wave"] = best
            row[name + "_tests_per_s"] = tests / (ms / 1e3)
            dst.free()
            cp.free()
            band.free()
        row["split_vs_fp32"] = row["fp32_ms"] / row["split_bf16_ms"]
        row["cps_equal"] = bool(np.array_equal(cps["fp32"], cps["split_bf16"]))
        log(f"density f={f} k={k}: fp32 {row['fp32_ms']:.3f} ms, split-BF16 "
            f"{row['split_bf16_ms']:.3f} ms ({row['split_vs_fp32']:.2f}x), "
            f"CPs equal {row['cps_equal']}")
        print(json.dumps(row), flush=True)
        ga.free()
        gb.free()
        gm.free()


def search_pass(P, rm, prm):
    """The whole PUMP search (PUMP.cu:3-405) on the roadmap, with its
    bookkeeping on the device and on the host; both must return the same
    path. The roadmap keeps the search's device buffers between searches:
    the first search grows them (hipMalloc/hipFree inside the call), the
    ones after it reuse them. Both are reported; the second is the figure
    for a planner that queries the same roadmap again."""
    import ctypes
    out = {}
    for name, dev in (("device_first", 1), ("device", 1), ("host", 0)):
        q = P.Params()
        ctypes.memmove(ctypes.byref(q), ctypes.byref(prm), ctypes.sizeof(P.Params))
        q.validity_bits = 1
        q.search_device = dev
        r = P.search_result_dict(rm.search(q))
        out[name] = r
    # smoothing and controller discretisation of the device search's result
    sm = None
    q = P.Params()
    ctypes.memmove(ctypes.byref(q), ctypes.byref(prm), ctypes.sizeof(P.Params))
    q.validity_bits, q.search_device = 1, 1
    r = P.SearchResult()
    P.check("pump_search_run", P.lib.pump_search_run(rm._h, ctypes.byref(q), ctypes.byref(r)))
    if r.status == 0:
        so, path, _, disc = rm.smooth(r, q)
        sm = {"status": so.status, "itrs": so.itrs, "mccp_launches": so.mccp_launches,
              "ms": so.ms_total, "cost_before": so.cost_before, "cost": so.cost, "cp": so.cp,
              "path_rows": so.path_rows, "discretized_rows": so.discretized_rows}
        log(f"smoothing: {so.itrs} steps in {so.mccp_launches} full-MC launches, "
            f"{so.ms_total:.1f} ms, cost {so.cost_before:.2f} -> {so.cost:.2f}, cp {so.cp:.4f}")
    d, h = out["device"], out["host"]
    keys = ("status", "itrs", "num_paths", "num_goal_paths", "soln_nodes", "soln_cost",
            "soln_cp", "particle_steps")
    same = all(d[k] == h[k] == out["device_first"][k] for k in keys)
    log(f"search: {d['ms_total']:.1f} ms with device bookkeeping (first search "
        f"{out['device_first']['ms_total']:.1f} ms), {h['ms_total']:.1f} ms on the "
        f"host, {d['itrs']} iterations, {d['num_paths']} paths, same path: {same}")
    return {"ms_device_bookkeeping": d["ms_total"],
            "ms_device_first_search": out["device_first"]["ms_total"],
            "ms_host_bookkeeping": h["ms_total"],
            "ms_propagate": d["ms_propagate"],
            "particle_steps": d["particle_steps"],
            "propagate_particle_steps_per_s": (d["particle_steps"] / (d["ms_propagate"] / 1e3)
                                               if d["ms_propagate"] > 0 else None),
            "itrs": d["itrs"], "num_paths": d["num_paths"],
            "num_considered": d["num_considered"], "status": d["status"],
            "soln_cost": d["soln_cost"], "soln_cp": d["soln_cp"], "same_result": same,
            "smoothing": sm}


def mccp_pass(P, rm, prm, samples, num_paths, reps, n=64):
    """Full Monte Carlo CP (collisionProbability.cu:4-92) in its batched form
    (pump_collision_probability_many) over the n deepest paths of the last
    search's path table, each traced back to the init and sampled by
    findOptimalPath.  A segment-check is one particle's segment between
    consecutive path rows tested against every box (the reference checks rows
    t -> t+1 for t < T-2; a particle stops at its first hit, rare here)."""
    tab = rm.paths(num_paths)
    prev, node = tab["path_prev"], tab["path_node"]
    depth = np.zeros(prev.size, np.int64)
    for k in range(1, prev.size):
        depth[k] = depth[prev[k]] + 1
    pick = np.lexsort((np.arange(prev.size), -depth))[:n]
    cap = int(prm.path_cap)
    s = np.ascontiguousarray(samples, np.float32)
    d = rm.device()
    info = rm.info
    paths = np.zeros((pick.size, cap, 6), np.float32)
    T = np.zeros(pick.size, np.int32)
    for q, k in enumerate(pick.tolist()):
        chain = []
        while k != -2:
            chain.append(int(node[k]))
            k = int(prev[k])
        xs = s[np.asarray(chain[::-1])]
        rows, _ = P.find_optimal_path(xs, dt=prm.dt, cap=cap)
        paths[q, :rows.shape[0]] = rows
        T[q] = rows.shape[0]
    g_p, g_T = P.DeviceArray(paths), P.DeviceArray(T)
    g_h = P.DeviceArray(shape=(pick.size,), dtype=np.int32, fill=0)
    stream = P.Stream()
    timer = P.Timer(stream)

    def run():
        P.check("pump_collision_probability_many", P.lib.pump_collision_probability_many(
            d.obstacles, int(info.n_obs), d.xcomb, int(info.tmax_steps), int(info.n_mc),
            np.float32(prm.offset_mult), g_p.p, cap, g_T.p, int(pick.size), g_h.p, 1, stream.h))

    run()
    stream.sync()
    timer.start()
    for _ in range(reps):
        run()
    ms = timer.stop() / reps
    hits = g_h.numpy()
    nseg = np.minimum(np.maximum(T.astype(np.int64) - 2, 0), min(int(info.tmax_steps), cap) - 1)
    checks = int(info.n_mc) * int(nseg.sum())
    out = {"paths": int(pick.size), "rows_min": int(T.min()), "rows_max": int(T.max()),
           "n_mc": int(info.n_mc), "ms": ms, "segment_checks": checks,
           "segment_checks_per_s": checks / (ms / 1e3), "mean_cp": float(hits.mean() / info.n_mc)}
    log(f"full MC (MCCP) of {pick.size} paths ({T.min()}-{T.max()} rows, {info.n_mc} particles): "
        f"{ms:.3f} ms, {out['segment_checks_per_s']:.3e} segment-checks/s")
    for a in (g_p, g_T, g_h):
        a.free()
    return out, (paths, T, hits)


def cpu_mccp(prm, obstacles, xcomb, paths, T, gpu_hits, seconds):
    """The CPU oracle's MCCP (collisionProbability.cu:41-92, one thread) over
    the same paths as the GPU's full-MC batch, path after path until about
    `seconds`; hits compared with the GPU's."""
    import oracle
    O = oracle.load()
    obs = np.ascontiguousarray(obstacles, np.float32)
    xc = np.ascontiguousarray(xcomb, np.float32)
    n_mc, tsteps = xc.shape[0], xc.shape[1]
    t0 = time.perf_counter()
    done, checks, agree = 0, 0, True
    for q in range(paths.shape[0]):
        h = O.po_mccp_hits(oracle.fp(obs), obs.shape[0], oracle.fp(xc), tsteps, n_mc,
                           np.float32(prm.offset_mult), oracle.fp(paths[q]), int(T[q]), None)
        agree &= int(h) == int(gpu_hits[q])
        done += 1
        checks += n_mc * max(int(T[q]) - 2, 0)
        if time.perf_counter() - t0 > seconds:
            break
    dt = time.perf_counter() - t0
    log(f"cpu full MC: {done} paths ({checks} segment-checks) in {dt:.2f} s on one thread, "
        f"hits agree: {agree}")
    return {"value": checks / dt, "unit": "segment-checks/s", "cores": 1, "kind": "port",
            "sample": f"first {done} of the GPU's {paths.shape[0]} full-MC paths", "seconds": dt,
            "hits_match_gpu": bool(agree)}


def batch_pass(P, n_inst, reps):
    """The first n_inst problems of the uneven batch workload built together
    (pump_batch_build), then the graph replays of its pair pass and of its
    edge pass + half-space cache, against the same device passes launched
    instance by instance on one stream (the single-instance entry points on
    the batch's own buffers)."""
    from pump_amd import configs
    cfgs = configs.batch_config()[:n_inst]
    probs = []
    for c in cfgs:
        c["params"].keep_pairs = 1
        s, o = P.gen_problem(c["spec"])
        probs.append((s, o, s[-1], c["params"]))
    ot, xc = P.make_noise(cfgs[0]["noise"])
    t0 = time.perf_counter()
    b = P.Batch(probs, ot, xc, autotune=True)
    build_s = time.perf_counter() - t0
    info = b.info
    log(f"batch of {info.n_inst}: built in {build_s:.1f} s, pairs {info.n_pairs}, "
        f"E={info.n_edges}, waypoints={info.n_waypoints}, launch shapes "
        f"{info.pairs_blocks_per_cu} WG/CU (pairs), {info.edges_per_block} edges/WG")
    b.replay(0)
    b.replay(1)
    ms_pairs = b.replay(0, reps)
    ms_edges = b.replay(1, reps)
    rms = [b.roadmap(k) for k in range(info.n_inst)]
    devs = [rm.device() for rm in rms]
    stream = P.Stream()
    timer = P.Timer(stream)

    def single_pairs():
        for rm, d, p in zip(rms, devs, probs):
            P.check("pump_fill_copts_topts", P.lib.pump_fill_copts_topts(
                d.samples, rm.info.num, p[3].tmax, d.copts, d.topts, stream.h))

    def single_edges():
        for rm, d in zip(rms, devs):
            I = rm.info
            P.check("pump_edges_discretize_check", P.lib.pump_edges_discretize_check(
                d.samples, d.edge_src, d.edge_dst, d.topts_edge, I.n_edges, d.wp_offsets, 0,
                d.obstacles, I.n_obs, d.disc, d.is_free_edges, stream.h))
            P.check("pump_cache_halfspaces", P.lib.pump_cache_halfspaces(
                I.n_edges, d.disc, d.is_free_edges, I.n_obs, I.max_hs, d.obstacles, 0,
                d.wp_offsets, d.as_, d.bs, d.count_stored, I.maxd2, 1, d.edge_hs_mask,
                stream.h))

    def timed(fn):
        fn()
        stream.sync()
        timer.start()
        for _ in range(reps):
            fn()
        return timer.stop() / reps

    ms_pairs_single = timed(single_pairs)
    ms_edges_single = timed(single_edges)
    log(f"batch pair pass: {ms_pairs:.2f} ms graph vs {ms_pairs_single:.2f} ms per instance; "
        f"edge pass + half-spaces: {ms_edges:.3f} ms graph vs {ms_edges_single:.3f} ms")
    # the HSMC step of every instance (each instance's free edges expanded
    # once from 4096 all-valid parents), captured as the third graph
    b.mc_build(n_parents=4096, max_paths=1 << 20, seed=1)
    b.replay(2)
    ms_mc = b.replay(2, reps)
    ms_mc_eager = b.mc_eager(reps)
    mc_steps = 0
    for k, rm in enumerate(rms):
        e = b.mc_get(k)["edge"]
        wo = rm.flags()["wp_offsets"]
        mc_steps += int(probs[k][3].n_p) * int((wo[e + 1] - wo[e]).sum())
    one_launch = int(b.info.mc_one_launch)
    log(f"batch HSMC step: {ms_mc:.2f} ms graph vs {ms_mc_eager:.2f} ms eager, "
        f"{mc_steps / (ms_mc / 1e3):.3e} particle-steps/s, one launch per stage {one_launch} "
        f"({b.info.mc_particles_per_wave} particles per wave)")
    out = {"n_inst": int(info.n_inst), "n_pairs": int(info.n_pairs),
           "n_edges": int(info.n_edges), "n_waypoints": int(info.n_waypoints),
           "pairs_blocks_per_cu": int(info.pairs_blocks_per_cu),
           "edges_per_block": int(info.edges_per_block), "build_s": build_s,
           "ms_autotune": float(info.ms_autotune),
           "pairs": {"ms_batched": ms_pairs, "ms_per_instance": ms_pairs_single,
                     "speedup": ms_pairs_single / ms_pairs,
                     "pair_solves_per_s": info.n_pairs / (ms_pairs / 1e3)},
           "edges_and_halfspaces": {"ms_batched": ms_edges, "ms_per_instance": ms_edges_single,
                                    "speedup": ms_edges_single / ms_edges,
                                    "edge_checks_per_s": info.n_edges / (ms_edges / 1e3)},
           "edge_obs_tile": int(info.edge_obs_tile),
           "hsmc": {"ms_graph": ms_mc, "ms_eager": ms_mc_eager, "particle_steps": mc_steps,
                    "particle_steps_per_s": mc_steps / (ms_mc / 1e3),
                    "one_launch_per_stage": one_launch,
                    "particles_per_wave": int(b.info.mc_particles_per_wave)}}
    b.close()
    return out


def _cpu_edges(rm, samples, obstacles, R, cores):
    """The CPU oracle's edge pass over the first R candidate edges; returns
    the seconds and whether its flags equal the GPU's."""
    import oracle
    from oracle import bp, fp, ip, lp
    O = oracle.load()
    e = rm.edges()
    fl = rm.flags()
    obs = np.ascontiguousarray(obstacles, np.float32)
    num = samples.shape[0]
    rs = e["row_start"]
    last_row = int(np.searchsorted(rs, R, side="left"))
    rs_cut = np.minimum(rs[: last_row + 1], R).astype(np.int64)
    rs_cut = np.concatenate([rs_cut, np.full(num + 1 - rs_cut.shape[0], R, np.int64)])
    wo = np.ascontiguousarray(fl["wp_offsets"][: R + 1])
    disc = np.zeros((int(wo[-1]) + 1, 6), np.float32)
    flags = np.zeros(R + 1, np.uint8)
    dst = np.ascontiguousarray(e["edge_dst"][:R])
    t0 = time.perf_counter()
    O.po_discretize_edges(fp(samples), lp(rs_cut), ip(dst), num, lp(wo), fp(disc), cores)
    O.po_free_edges(fp(obs), obs.shape[0], fp(disc), lp(wo), R, bp(flags), cores)
    dt = time.perf_counter() - t0
    return dt, bool(np.array_equal(flags[:R], fl["is_free_edges"][:R]))


def _sized(run, n0, n_max, seconds):
    """run(n) -> (seconds, ...): a first run at n0, then one sized to take
    about `seconds`."""
    n = min(n_max, n0)
    res = run(n)
    if res[0] < seconds * 0.5 and n < n_max:
        n = int(min(n_max, n * max(1.0, seconds / max(res[0], 1e-3))))
        res = run(n)
    return n, res


def cpu_baseline(rm, samples, obstacles, prm, seconds):
    """The CPU oracle's edge pass (findDiscretizedPath + freeEdges) over the
    first R candidate edges, OpenMP over all host cores; R is sized so that
    the run takes about `seconds`."""
    cores = len(os.sched_getaffinity(0))
    E = int(rm.info.n_edges)
    R, (dt, agree) = _sized(lambda r: _cpu_edges(rm, samples, obstacles, r, cores), 20000, E,
                            seconds)
    log(f"cpu baseline: {R} edges in {dt:.2f} s on {cores} threads, flags agree: {agree}")
    return {"value": R / dt, "unit": "edge-checks/s", "cores": cores, "kind": "port",
            "sample": f"first {R} of {E} candidate edges (CPU oracle findDiscretizedPath + "
                      f"freeEdges, OpenMP)", "seconds": dt, "flags_match_gpu": agree}


def cpu_edges_one_thread(rm, samples, obstacles, seconds):
    """The same CPU edge pass on one thread."""
    E = int(rm.info.n_edges)
    R, (dt, agree) = _sized(lambda r: _cpu_edges(rm, samples, obstacles, r, 1), 2000, E,
                            seconds)
    log(f"cpu baseline, one thread: {R} edges in {dt:.2f} s, flags agree: {agree}")
    return {"value": R / dt, "unit": "edge-checks/s", "cores": 1, "kind": "port",
            "sample": f"first {R} of {E} candidate edges, one thread", "seconds": dt,
            "flags_match_gpu": agree}


def cpu_pairs(rm, samples, prm, g_c, g_t, seconds):
    """The CPU oracle's all-pairs 2pBVP (fillCoptsTopts, 2pBVP.cu:22-50) over
    the first r source rows (r x (NUM-1) pair solves), all host cores; its
    rows are compared with the device's tables g_c, g_t (cost, arrival time):
    the share of bit-equal entries and the largest relative difference (the
    oracle's pow() misrounds a few tau^3, tau^4 in the last place, so a few
    entries may differ by an ulp)."""
    import oracle
    import pump_amd as P
    O = oracle.load()
    cores = len(os.sched_getaffinity(0))
    num = samples.shape[0]
    s = np.ascontiguousarray(samples, np.float32)

    def run(r):
        rows = np.arange(r, dtype=np.int32)
        c = np.zeros((r, num - 1), np.float32)
        t = np.zeros((r, num - 1), np.float32)
        t0 = time.perf_counter()
        O.po_fill_rows(oracle.fp(s), num, np.float32(prm.tmax), oracle.ip(rows), r,
                       oracle.fp(c), oracle.fp(t), cores)
        return time.perf_counter() - t0, c, t

    r, (dt, c, t) = _sized(run, 8, num, seconds)
    n = r * (num - 1)
    out = {"value": n / dt, "unit": "pair-solves/s", "cores": cores, "kind": "port",
           "sample": f"first {r} of {num} source rows ({n} pair solves)", "seconds": dt}
    if g_c is not None:
        gh = {}
        for name, g in (("c", g_c), ("t", g_t)):
            a = np.empty(n, np.float32)
            P.check("pump_dev_download", P.lib.pump_dev_download(a.ctypes.data, g.ptr, n * 4))
            gh[name] = a
        ref = {"c": c.reshape(-1), "t": t.reshape(-1)}
        same = [float(np.mean(gh[k].view(np.uint32) == ref[k].view(np.uint32))) for k in "ct"]
        rel = [float(np.max(np.abs(gh[k].astype(np.float64) - ref[k]) /
                            np.maximum(np.abs(ref[k].astype(np.float64)), 1e-30))) for k in "ct"]
        out["rows_bit_equal_fraction"] = min(same)
        out["rows_max_rel_diff"] = max(rel)
    log(f"cpu pair solves: {n} in {dt:.2f} s on {cores} threads, bit-equal share "
        f"{out.get('rows_bit_equal_fraction')}, largest relative difference "
        f"{out.get('rows_max_rel_diff')}")
    return out


def cpu_particles(rm, prm, offsets_time, gpu_cp, seconds):
    """The CPU oracle's HSMC step (propagateValidParticles + calculateCP_HSMC,
    hsmc.cu:101-199) over the first R children of the synthetic wavefront,
    all host cores; its CPs are compared with the device drop-in form's."""
    import oracle
    from oracle import bp, fp, ip, lp
    O = oracle.load()
    cores = len(os.sched_getaffinity(0))
    info = rm.info
    wf = synthetic_wavefront(rm)
    hs = rm.halfspaces()
    topts = rm.edges()["topts_edge"]
    wo = wf["wp_offsets"]
    n_p, n_par = int(prm.n_p), wf["n_par"]
    ot = np.ascontiguousarray(offsets_time, np.float32)

    def run(R):
        edge = np.ascontiguousarray(wf["free"][:R])
        prev = np.ascontiguousarray(wf["prev"][:R])
        valid = np.zeros((n_par + R, n_p), np.uint8)
        valid[:n_par] = 1
        cp = np.zeros(n_par + R, np.float32)
        t0 = time.perf_counter()
        O.po_propagate_valid_particles(R, lp(wo), ip(edge), ip(prev), n_par, fp(hs["as_"]),
                                       fp(hs["bs"]), fp(wf["times"]), fp(topts), prm.dt,
                                       info.max_hs, fp(ot), info.tmax_steps, prm.offset_mult,
                                       info.n_mc, bp(valid), n_p, cores)
        O.po_calculate_cp_hsmc(R, n_par, fp(cp), bp(valid), n_p)
        dt = time.perf_counter() - t0
        steps = n_p * int((wo[edge + 1] - wo[edge]).sum())
        return dt, steps, bool(np.array_equal(cp[n_par:], gpu_cp[:R]))

    n_new = int(wf["free"].shape[0])
    R, (dt, steps, agree) = _sized(run, 2000, n_new, seconds)
    log(f"cpu particle steps: {steps} ({R} children) in {dt:.2f} s on {cores} threads, "
        f"CPs agree: {agree}")
    return {"value": steps / dt, "unit": "particle-steps/s", "cores": cores, "kind": "port",
            "sample": f"first {R} of {n_new} children of the synthetic wavefront", "seconds": dt,
            "cp_match_gpu": agree}


def edge_traffic(form):
    """HBM bytes per launch of the edge pass at NUM=16000 from the committed
    counter runs (profiles/edge_pass_num16000_traffic.json): the exact fused
    pass, or the FP32-waypoint form with form="fp32_form"; None if absent."""
    path = os.path.join(HERE, "profiles", "edge_pass_num16000_traffic.json")
    try:
        j = json.load(open(path))
    except (OSError, ValueError):
        return None
    if form:
        j = j.get(form) or {}
    return j.get("bytes_per_launch")


def mc_counters():
    """The MC pass's VALU and MFMA activity from the committed counter run
    (profiles/hsmc_num16000_counters.json), or None."""
    path = os.path.join(HERE, "profiles", "hsmc_num16000_counters.json")
    try:
        return json.load(open(path))
    except (OSError, ValueError):
        return None


if __name__ == "__main__":
    main()
