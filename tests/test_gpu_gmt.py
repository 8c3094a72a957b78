"""GPU parity of the Euclidean neighbour precompute (pump_precomp_nn_count /
pump_precomp_nn, precomp.cu:3-76) and the Group Marching Tree wavefront
(pump_gmt_wavefront, GMT.cu:3-261) against the oracle, through the C ABI.

Bar: bit-exact -- counts, neighbour lists, distances (sqrtf of the same
float32 sum), costs (float32 sums of the same terms in the same order),
parent edges, goal cost and iteration count.  GMT runs on random geometric
graphs with boxes and on roadmaps built by the pipeline, with the connecting
edge found both through the dense nnIdxs table and through the CSR rows, and
with per-edge waypoint offsets as well as the fixed numDisc layout."""
import numpy as np
import pytest

import oracle
import oracle.pipeline as OP
import pump_amd as P
from test_gpu_parity import CASES, build_pair
from test_oracle_gmt import brute_nn, geometric_graph, graph_tables, oracle_nn, straight_disc

pytestmark = pytest.mark.gpu

O = oracle.load()
fp, ip, lp, bp = oracle.fp, oracle.ip, oracle.lp, oracle.bp


@pytest.mark.parametrize("num,r,cut", [(1, 0.5, None), (700, 0.55, None), (2049, 0.5, None),
                                       (1500, 0.6, 4)])
def test_precomp_nn_parity(num, r, cut):
    rng = np.random.default_rng(num)
    s = rng.random((num, 6), dtype=np.float32)
    print("precomp", num, r, cut, flush=True)
    r2 = np.float32(r) * np.float32(r)
    counts, dist, edges = P.precomp_nn(s, r, nn_size=cut)
    m = int(counts.max()) if cut is None else cut
    sizes, d_ref, e_ref = oracle_nn(s, r2, m)
    assert np.array_equal(counts, sizes)
    assert np.array_equal(edges, e_ref)
    assert np.array_equal(dist.view(np.uint32), d_ref.view(np.uint32))
    if cut is None and num > 1:
        assert np.array_equal(counts, brute_nn(s, r2)[0].sum(1))


def oracle_gmt(obs, go, come, dcome, disc, wp, row_start, dst, free, goal):
    num = len(go)
    obs = np.ascontiguousarray(obs, np.float32).reshape(-1, 12)
    costs = np.zeros(num, np.float32)
    edges = np.zeros(num, np.int32)
    cg, it = np.zeros(1, np.float32), np.zeros(1, np.int32)
    rc = O.po_gmt_wavefront(fp(obs), len(obs), fp(np.ascontiguousarray(dcome, np.float32)),
                            ip(np.ascontiguousarray(go)), ip(np.ascontiguousarray(come)),
                            go.shape[1], fp(np.ascontiguousarray(disc, np.float32)),
                            lp(np.ascontiguousarray(wp, np.int64)),
                            lp(np.ascontiguousarray(row_start, np.int64)),
                            ip(np.ascontiguousarray(dst, np.int32)),
                            bp(np.ascontiguousarray(free, np.uint8)), num, fp(costs), ip(edges),
                            goal, fp(cg), ip(it))
    assert rc == 0
    return costs, edges, float(cg[0]), int(it[0])


def assert_same(got, ref):
    assert np.array_equal(got[1], ref[1])
    assert np.array_equal(got[0].view(np.uint32), ref[0].view(np.uint32))
    assert got[2:] == ref[2:]


def dense_idx(num, src, dst):
    d = np.full((num, num), -3, np.int32)
    d[dst, src] = np.arange(len(dst), dtype=np.int32)
    return d


@pytest.mark.parametrize("seed,n_obs", [(0, 0), (1, 12), (2, 30)])
def test_gmt_geometric_parity(seed, n_obs):
    num, nd = 600, 5
    s, adj, w = geometric_graph(num, 0.85, seed)
    go, come, dcome, row_start, dst, src = graph_tables(num, adj, w)
    disc = straight_disc(s, src, dst, nd)
    wp = np.arange(len(dst) + 1, dtype=np.int64) * (nd + 1)
    rng = np.random.default_rng(seed + 50)
    lo = rng.random((n_obs, 3), dtype=np.float32) * np.float32(1.7)
    z = np.zeros((n_obs, 3), np.float32)
    obs = np.concatenate([lo, z, lo + np.float32(0.25), z], 1).astype(np.float32)
    free = np.ones(num, np.uint8)
    free[rng.integers(1, num - 1, 10)] = 0
    print("gmt geometric", seed, n_obs, len(dst), flush=True)
    ref = oracle_gmt(obs, go, come, dcome, disc, wp, row_start, dst, free, num - 1)
    dense = dense_idx(num, src, dst)
    for kw in (dict(wp_offsets=wp, row_start=row_start, edge_dst=dst),
               dict(wp_offsets=wp, nn_idxs=dense),
               dict(num_disc=nd, nn_idxs=dense),
               dict(num_disc=nd, row_start=row_start, edge_dst=dst)):
        got = P.gmt_wavefront(obs, go, come, dcome, disc, free, num - 1, **kw)
        assert_same(got, ref)
    assert ref[3] >= 1


@pytest.mark.parametrize("case", [CASES[0], CASES[1], CASES[3]])
def test_gmt_roadmap_parity(case):
    """GMT over a pipeline-built roadmap: its padded go/come tables, edge
    costs as distances, waypoints and free flags."""
    cfg, s, o, ot, xc, rm, orc = build_pair(case)
    num = s.shape[0]
    nn = rm.nn_tables(dense=True)
    f = rm.flags()
    disc = rm.waypoints()
    m = rm.info.max_nn
    go = np.empty((num, m), np.int32)
    come = np.empty((num, m), np.int32)
    dist = np.empty((num, m), np.float32)
    O.po_nn_tables(oracle.lp(orc["row_start"]), oracle.ip(orc["edge_dst"]),
                   oracle.fp(orc["copts_edge"]), num, m, ip(go), ip(come), fp(dist))
    ref = oracle_gmt(o, go, come, dist, orc["disc"], orc["wp_offsets"], orc["row_start"],
                     orc["edge_dst"], orc["is_free_samples"], num - 1)
    print("gmt roadmap", case, "itrs", ref[3], "cost", ref[2], flush=True)
    variants = [dict(wp_offsets=f["wp_offsets"], nn_idxs=nn["nn_idxs"]),
                dict(wp_offsets=f["wp_offsets"], row_start=orc["row_start"],
                     edge_dst=orc["edge_dst"])]
    w_fixed = case[2]
    if w_fixed:
        variants.append(dict(num_disc=w_fixed - 1, nn_idxs=nn["nn_idxs"]))
    for kw in variants:
        got = P.gmt_wavefront(o, nn["nn_go_edges"], nn["nn_come_edges"], nn["distances_come"],
                              disc, f["is_free_samples"], num - 1, **kw)
        assert_same(got, ref)
    rm.close()
