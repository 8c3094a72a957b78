"""Batched roadmap build (pump_batch_build) against single-instance builds.

The batch runs the pair pass, the fused edge pass and the half-space cache of
every instance in one launch each (captured in hipGraphs); every table it
produces must be bit-identical to the single-instance build of the same
problem, which the other parity tests pin to the CPU oracle.  Instances are
taken from the uneven 256-problem batch workload (NUM 1000-8000, 5-50 boxes,
per-instance noise multiplier and particle count), plus an instance with no
boxes and a fixed waypoint count and a 40-sample instance.
"""
import ctypes

import numpy as np
import pytest

import pump_amd as P
from pump_amd import configs

pytestmark = pytest.mark.gpu


def download(ptr, n, dtype):
    out = np.empty(n, dtype)
    if n:
        P.check("pump_dev_download",
                P.lib.pump_dev_download(out.ctypes.data_as(ctypes.c_void_p), ptr, out.nbytes))
    return out


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint8)


def tables(rm):
    d = {}
    c, t = rm.pairs()
    d["copts"], d["topts"] = c, t
    d.update(rm.edges())
    d.update(rm.flags())
    d["disc"] = rm.waypoints()
    d.update(rm.halfspaces())
    dev = rm.device()
    d["edge_hs_mask"] = download(dev.edge_hs_mask, rm.info.n_edges, np.uint64)
    for k in ("rn", "n_edges", "n_waypoints", "max_nn", "max_hs", "n_free_edges", "rn_idx"):
        d["info_" + k] = getattr(rm.info, k)
    return d


def assert_same(got, ref, label):
    assert got.keys() == ref.keys()
    for k in ref:
        if isinstance(ref[k], np.ndarray):
            assert got[k].shape == ref[k].shape, (label, k)
            assert np.array_equal(bits(got[k]), bits(ref[k])), (label, k)
        else:
            assert got[k] == ref[k], (label, k)


def problems():
    batch = configs.batch_config()
    order = np.argsort([b["spec"].num for b in batch])
    out = []
    for k in list(order[:6]) + [int(np.argmax([b["spec"].n_obs for b in batch]))]:
        cfg = batch[int(k)]
        prm = cfg["params"]
        prm.keep_pairs = 1
        s, o = P.gen_problem(cfg["spec"])
        out.append((cfg["name"], s, o, s[-1], prm))
    # no boxes and a fixed waypoint count; a 40-sample instance
    cfg = configs.small(num=300, n_obs=0, w_fixed=20)
    s, o = P.gen_problem(cfg["spec"])
    cfg["params"].keep_pairs = 1
    out.append(("small_num300_noboxes", s, o, s[-1], cfg["params"]))
    cfg = configs.small(num=40, n_obs=3, w_fixed=5, k=8)
    s, o = P.gen_problem(cfg["spec"])
    cfg["params"].keep_pairs = 1
    out.append(("small_num40", s, o, s[-1], cfg["params"]))
    return out


@pytest.fixture(scope="module")
def noise():
    return P.make_noise(configs.batch_config(1)[0]["noise"])


def test_batch_matches_single_builds(noise):
    ot, xc = noise
    probs = problems()
    print("instances", [p[0] for p in probs], flush=True)
    b = P.Batch([(s, o, g, prm) for _, s, o, g, prm in probs], ot, xc, autotune=True)
    info = b.info
    assert info.n_inst == len(probs)
    assert info.pairs_blocks_per_cu in (4, 8, 16)
    assert info.edges_per_block in (64, 128, 256)
    max_obs = max(int(o.shape[0]) for _, _, o, _, _ in probs)
    assert info.edge_obs_tile == 32 or info.edge_obs_tile == min(64, max_obs)
    assert info.autotune_cached == 0
    print("batch", info.pairs_blocks_per_cu, info.edges_per_block, info.edge_obs_tile,
          info.ms_pairs, info.ms_edges, flush=True)
    # the same configuration again: the tuned shapes are reused, not re-timed
    b2 = P.Batch([(s, o, g, prm) for _, s, o, g, prm in probs], ot, xc, autotune=True)
    i2 = b2.info
    assert i2.autotune_cached == 3 and i2.ms_autotune == 0
    assert (i2.pairs_blocks_per_cu, i2.edges_per_block, i2.edge_obs_tile) == (
        info.pairs_blocks_per_cu, info.edges_per_block, info.edge_obs_tile)
    b2.close()
    singles = []
    n_pairs = n_edges = n_wp = 0
    for k, (name, s, o, g, prm) in enumerate(probs):
        rm = P.Roadmap(s, o, g, ot, xc, prm)
        ref = tables(rm)
        singles.append((rm, ref))
        got = tables(b.roadmap(k))
        assert_same(got, ref, name)
        n_pairs += rm.info.n_pairs
        n_edges += rm.info.n_edges
        n_wp += rm.info.n_waypoints
    assert (info.n_pairs, info.n_edges, info.n_waypoints) == (n_pairs, n_edges, n_wp)

    # graph replays rewrite the same tables
    for stage in (1, 0, 1):
        assert b.replay(stage, reps=2) > 0
    for k in (0, len(probs) - 2, len(probs) - 1):
        assert_same(tables(b.roadmap(k)), singles[k][1], probs[k][0] + " after replay")

    # the search over a batch roadmap returns the single-instance result
    for k in (0, len(probs) - 2):
        got = P.search_result_dict(b.roadmap(k).search())
        ref = P.search_result_dict(singles[k][0].search())
        for key in ("status", "itrs", "num_paths", "num_goal_paths", "soln_nodes", "soln_cost",
                    "soln_cp", "soln_hsmc_cp", "particle_steps"):
            assert got[key] == ref[key], (probs[k][0], key)
    for rm, _ in singles:
        rm.close()
    b.close()


def test_batch_without_kept_pairs(noise):
    """Pairs freed after the build: the pair graph is dropped (replay refused),
    the edge graph still replays, and pump_roadmap_free leaves batch roadmaps
    alone."""
    ot, xc = noise
    probs = problems()[-2:]
    for p in probs:
        p[4].keep_pairs = 0
    b = P.Batch([(s, o, g, prm) for _, s, o, g, prm in probs], ot, xc, autotune=False)
    assert b.info.pairs_blocks_per_cu == 8 and b.info.edges_per_block == 128
    with pytest.raises(P.PumpError) as ex:
        b.replay(0)
    assert ex.value.code == -1
    assert b.replay(1) > 0
    rm = b.roadmap(1)
    P.lib.pump_roadmap_free(rm._h)  # ignored for a batch roadmap
    single = P.Roadmap(probs[1][1], probs[1][2], probs[1][3], ot, xc, probs[1][4])
    assert np.array_equal(bits(b.roadmap(1).waypoints()), bits(single.waypoints()))
    hs_b, hs_s = b.roadmap(1).halfspaces(), single.halfspaces()
    for k in hs_s:
        assert np.array_equal(bits(hs_b[k]), bits(hs_s[k])), k
    single.close()
    b.close()


def test_batch_all_256_instances(noise):
    """The whole uneven 256-problem workload in one batch; a sample of its
    instances (the smallest, the largest, the most boxes, the first, the last
    and six drawn at random) against single-instance builds, and every
    instance's counts against the batch totals."""
    ot, xc = noise
    batch = configs.batch_config()
    assert len(batch) == 256
    nums = np.array([c["spec"].num for c in batch])
    n_obs = np.array([c["spec"].n_obs for c in batch])
    rng = np.random.default_rng(256)
    sample = sorted({0, 255, int(np.argmin(nums)), int(np.argmax(nums)), int(np.argmax(n_obs)),
                     *(int(k) for k in rng.choice(256, 6, replace=False))})
    probs = []
    for k, cfg in enumerate(batch):
        prm = cfg["params"]
        prm.keep_pairs = 1 if k in sample else 0
        s, o = P.gen_problem(cfg["spec"])
        probs.append((cfg["name"], s, o, s[-1], prm))
    print("sampled instances", [probs[k][0] for k in sample], flush=True)
    b = P.Batch([(s, o, g, prm) for _, s, o, g, prm in probs], ot, xc, autotune=True)
    info = b.info
    assert info.n_inst == 256
    print("batch of 256:", info.n_pairs, "pairs", info.n_edges, "edges", info.n_waypoints,
          "waypoints", info.ms_pairs, "ms pairs", info.ms_edges, "ms edges", flush=True)
    tot = np.zeros(3, np.int64)
    for k in range(256):
        i = b.roadmap(k).info
        assert i.num == nums[k] and i.n_edges > 0 and i.n_waypoints >= 2 * i.n_edges
        tot += (i.n_pairs, i.n_edges, i.n_waypoints)
    assert tuple(tot) == (info.n_pairs, info.n_edges, info.n_waypoints)
    for k in sample:
        name, s, o, g, prm = probs[k]
        rm = P.Roadmap(s, o, g, ot, xc, prm)
        assert_same(tables(b.roadmap(k)), tables(rm), name)
        rm.close()
    # the edge graph replays over all 256 and rewrites the same tables
    assert b.replay(1, reps=2) > 0
    k = sample[len(sample) // 2]
    name, s, o, g, prm = probs[k]
    rm = P.Roadmap(s, o, g, ot, xc, prm)
    assert_same(tables(b.roadmap(k)), tables(rm), name + " after replay")
    rm.close()
    b.close()


def check_mc_step(probs, noise, one_launch):
    """Runs the batch's MC stage on probs and checks every instance's
    children's CPs against the CPU oracle, after graph replays and an eager
    pass too; one_launch: the form the build must pick."""
    import oracle
    from oracle import bp, fp, ip, lp
    O = oracle.load()
    ot, xc = noise
    b = P.Batch([(s, o, g, prm) for _, s, o, g, prm in probs], ot, xc, autotune=False)
    n_par = 64
    ms = b.mc_build(n_parents=n_par, max_paths=20000, seed=7)
    print("batch MC step, first launch", ms, "ms, one launch per stage", b.info.mc_one_launch,
          flush=True)
    assert b.info.mc_one_launch == one_launch

    def check_all(label):
        for k in range(len(probs)):
            name, s, o, g, prm = probs[k]
            rm = b.roadmap(k)
            got = b.mc_get(k)
            n_new = got["edge"].size
            fl, hs, e = rm.flags(), rm.halfspaces(), rm.edges()
            free = np.nonzero(fl["is_free_edges"])[0][:20000]
            assert np.array_equal(got["edge"], free), (name, label)
            assert np.array_equal(got["prev"], np.arange(n_new) % n_par)
            n_p = int(prm.n_p)
            valid = np.zeros((n_par + n_new, n_p), np.uint8)
            valid[:n_par] = 1
            O.po_propagate_valid_particles(
                n_new, lp(fl["wp_offsets"]), ip(got["edge"]), ip(got["prev"]), n_par,
                fp(hs["as_"]), fp(hs["bs"]), fp(got["times"]), fp(e["topts_edge"]), prm.dt,
                rm.info.max_hs, fp(ot), rm.info.tmax_steps, prm.offset_mult, rm.info.n_mc,
                bp(valid), n_p, 8)
            cp = np.zeros(n_par + n_new, np.float32)
            O.po_calculate_cp_hsmc(n_new, n_par, fp(cp), bp(valid), n_p)
            assert np.array_equal(got["cp"], cp[n_par:]), (name, label)
            # children that meet no half-space keep their parent's row
            own = got["row_src"] >= 0
            assert np.array_equal(got["row_src"][~own], -1 - got["prev"][~own])

    check_all("first launch")
    assert b.replay(2, reps=2) > 0
    check_all("after replays")
    print("eager", b.mc_eager(reps=2), "ms; graph", b.replay(2, reps=3), "ms", flush=True)
    check_all("after the eager pass")
    b.close()


def test_batch_mc_step(noise):
    """The HSMC step of every instance captured as the batch's third graph
    (pump_batch_mc_build, one launch per stage for all instances): each
    instance's children's CPs equal the CPU oracle's propagateValidParticles
    + calculateCP_HSMC on the same wavefront and tables, after graph replays
    and an eager pass too."""
    check_mc_step(problems(), noise, 1)


def test_batch_mc_step_side_streams(noise):
    """An instance whose bit rows are 8 bytes (64 particles) cannot take the
    one-launch form: the build runs every instance's step on the side
    streams instead, with the same results."""
    probs = problems()
    name, s, o, g, prm = probs[-1]
    prm.n_p = 64
    probs[-1] = (name + "_p64", s, o, g, prm)
    check_mc_step(probs, noise, 0)
