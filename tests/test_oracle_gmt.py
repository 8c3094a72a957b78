"""Pins of the oracle's restatement of the Euclidean neighbour precompute
(precomp.cu:3-76) and the Group Marching Tree wavefront (GMT.cu:3-261).

precomp is checked against a numpy float32 brute force in the same summation
order.  GMT is checked against closed forms (a chain, a complete graph, a
blocked direct edge, a blocked sample) and, on random geometric graphs,
against breadth-first hop layers and Dijkstra from scipy plus the invariants
that every kept connection satisfies."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as csg

import oracle

O = oracle.load()
fp, ip, lp, bp = oracle.fp, oracle.ip, oracle.lp, oracle.bp


def brute_nn(samples, r2):
    x = samples.astype(np.float32)
    d2 = np.zeros((len(x), len(x)), np.float32)
    for d in range(6):
        diff = (x[:, None, d] - x[None, :, d]).astype(np.float32)
        d2 = (d2 + diff * diff).astype(np.float32)
    np.fill_diagonal(d2, np.inf)
    return d2 < np.float32(r2), d2


def oracle_nn(samples, r2, nn_size):
    num = len(samples)
    sizes = np.zeros(num, np.int32)
    dist = np.zeros((num, nn_size), np.float32)
    edges = np.zeros((num, nn_size), np.int32)
    O.po_precomp_nn(fp(np.ascontiguousarray(samples, np.float32)), num, np.float32(r2), nn_size,
                    ip(sizes), fp(dist), ip(edges))
    return sizes, dist, edges


@pytest.mark.parametrize("num,r", [(1, 0.5), (50, 0.3), (400, 0.45), (300, 2.0)])
def test_precomp_matches_brute_force(num, r):
    rng = np.random.default_rng(num)
    s = rng.random((num, 6), dtype=np.float32)
    r2 = np.float32(r) * np.float32(r)
    within, d2 = brute_nn(s, r2)
    counts = within.sum(1)
    m = int(counts.max()) if num else 0
    sizes, dist, edges = oracle_nn(s, r2, m)
    assert np.array_equal(sizes, counts)
    for n in range(num):
        nbr = np.flatnonzero(within[n])
        assert np.array_equal(edges[n, :len(nbr)], nbr)
        assert np.array_equal(dist[n, :len(nbr)], np.sqrt(d2[n, nbr]).astype(np.float32))
        assert np.all(edges[n, len(nbr):] == -1) and np.all(dist[n, len(nbr):] == 0)


def test_precomp_truncates_rows_at_nn_size():
    rng = np.random.default_rng(7)
    s = rng.random((200, 6), dtype=np.float32)
    r2 = np.float32(0.5)
    within, _ = brute_nn(s, r2)
    full = oracle_nn(s, r2, int(within.sum(1).max()))
    cut = oracle_nn(s, r2, 3)
    assert np.array_equal(cut[0], full[0])          # counts are never truncated
    assert np.array_equal(cut[2], full[2][:, :3])   # the first three in sample order
    assert np.array_equal(cut[1], full[1][:, :3])


# ------------------------------------------------------------------ GMT ----
def straight_disc(samples, src, dst, num_disc):
    """num_disc+1 waypoints per edge on the segment between its end samples."""
    t = (np.arange(num_disc + 1, dtype=np.float32) / np.float32(num_disc))[None, :, None]
    a, b = samples[src][:, None, :], samples[dst][:, None, :]
    return (a + t * (b - a)).astype(np.float32).reshape(-1, 6)


def graph_tables(num, adj, w):
    """Padded go/come tables (go = out-neighbours in node order, come =
    in-neighbours in node order), CSR rows of the out-edges and edge weights."""
    m = max(int(adj.sum(1).max()), int(adj.sum(0).max()), 1)
    go = np.full((num, m), -1, np.int32)
    come = np.full((num, m), -1, np.int32)
    dcome = np.zeros((num, m), np.float32)
    for n in range(num):
        out = np.flatnonzero(adj[n])
        go[n, :len(out)] = out
        inn = np.flatnonzero(adj[:, n])
        come[n, :len(inn)] = inn
        dcome[n, :len(inn)] = w[inn, n]
    src, dst = np.nonzero(adj)
    row_start = np.zeros(num + 1, np.int64)
    np.add.at(row_start, src + 1, 1)
    row_start = np.cumsum(row_start).astype(np.int64)
    return go, come, dcome, row_start, dst.astype(np.int32), src.astype(np.int32)


def run_gmt(samples, adj, w, obstacles, goal, num_disc=4, free=None):
    num = len(samples)
    go, come, dcome, row_start, dst, src = graph_tables(num, adj, w)
    disc = straight_disc(samples, src, dst, num_disc)
    wp = (np.arange(len(dst) + 1, dtype=np.int64) * (num_disc + 1))
    free = np.ones(num, np.uint8) if free is None else free.astype(np.uint8)
    obs = np.ascontiguousarray(obstacles, np.float32).reshape(-1, 12)
    costs = np.zeros(num, np.float32)
    edges = np.zeros(num, np.int32)
    cg, it = np.zeros(1, np.float32), np.zeros(1, np.int32)
    rc = O.po_gmt_wavefront(fp(obs), len(obs), fp(dcome), ip(go), ip(come), go.shape[1], fp(disc),
                            lp(wp), lp(row_start), ip(dst), bp(free), num, fp(costs), ip(edges),
                            goal, fp(cg), ip(it))
    assert rc == 0
    return costs, edges, float(cg[0]), int(it[0]), dict(src=src, dst=dst, disc=disc, wp=wp)


def line_samples(xs):
    s = np.zeros((len(xs), 6), np.float32)
    s[:, 0] = xs
    return s


def box(lo, hi):
    return np.array(list(lo) + [0, 0, 0] + list(hi) + [0, 0, 0], np.float32)


def test_gmt_chain_is_prefix_sums():
    num = 9
    s = line_samples(np.cumsum(np.linspace(0.1, 0.5, num)).astype(np.float32))
    adj = np.zeros((num, num), bool)
    adj[np.arange(num - 1), np.arange(1, num)] = True
    w = np.abs(s[:, None, 0] - s[None, :, 0]).astype(np.float32)
    costs, edges, cg, it, _ = run_gmt(s, adj, w, np.zeros((0, 12)), num - 1)
    want = np.zeros(num, np.float32)
    for i in range(1, num):
        want[i] = np.float32(want[i - 1] + w[i - 1, i])
    assert np.array_equal(costs, want)
    assert np.array_equal(edges, np.r_[-1, np.arange(num - 1)])
    assert it == num - 1 and cg == want[-1]


def test_gmt_complete_graph_one_step():
    rng = np.random.default_rng(3)
    num = 30
    s = rng.random((num, 6), dtype=np.float32)
    adj = ~np.eye(num, dtype=bool)
    w = rng.random((num, num), dtype=np.float32) + np.float32(0.1)
    costs, edges, cg, it, _ = run_gmt(s, adj, w, np.zeros((0, 12)), num - 1)
    assert it == 1
    assert np.array_equal(edges[1:], np.zeros(num - 1, np.int32))
    assert np.array_equal(costs[1:], w[0, 1:])
    assert costs[0] == 0 and edges[0] == -1 and cg == w[0, num - 1]


def test_gmt_blocked_direct_edge_takes_the_detour():
    # 0 -> 2 crosses a box; 0 -> 1 -> 2 passes above it
    s = line_samples([0.0, 0.5, 1.0])
    s[1, 1] = 1.0
    adj = np.zeros((3, 3), bool)
    adj[0, 1] = adj[0, 2] = adj[1, 2] = True
    w = np.full((3, 3), 0.0, np.float32)
    w[0, 1], w[0, 2], w[1, 2] = 1.25, 1.0, 0.75
    wall = box((0.45, -0.2, -1), (0.55, 0.2, 1))
    c0, e0, cg0, it0, _ = run_gmt(s, adj, w, np.zeros((0, 12)), 2)
    assert (it0, cg0, e0[2]) == (1, np.float32(1.0), 0)
    c1, e1, cg1, it1, _ = run_gmt(s, adj, w, wall[None], 2)
    assert (it1, e1[2], e1[1]) == (2, 1, 0)
    assert cg1 == np.float32(np.float32(1.25) + np.float32(0.75))


def test_gmt_blocked_sample_keeps_its_marks():
    rng = np.random.default_rng(4)
    num = 12
    s = rng.random((num, 6), dtype=np.float32)
    adj = ~np.eye(num, dtype=bool)
    w = rng.random((num, num), dtype=np.float32) + np.float32(0.1)
    free = np.ones(num, np.uint8)
    free[[2, 5]] = 0
    costs, edges, cg, it, _ = run_gmt(s, adj, w, np.zeros((0, 12)), num - 1, free=free)
    assert np.all(costs[[2, 5]] == -11) and np.all(edges[[2, 5]] == -2)
    keep = np.flatnonzero(free)[1:]
    assert np.array_equal(costs[keep], w[0, keep]) and it == 1


def geometric_graph(num, r, seed):
    rng = np.random.default_rng(seed)
    s = (rng.random((num, 6), dtype=np.float32) * np.float32(2)).astype(np.float32)
    r2 = np.float32(r) * np.float32(r)
    m = int(brute_nn(s, r2)[0].sum(1).max())
    sizes, dist, nn = oracle_nn(s, r2, m)
    adj = np.zeros((num, num), bool)
    w = np.zeros((num, num), np.float32)
    for n in range(num):
        k = sizes[n]
        adj[n, nn[n, :k]] = True
        w[n, nn[n, :k]] = dist[n, :k]
    return s, adj, w


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_gmt_obstacle_free_follows_hop_layers(seed):
    num = 300
    s, adj, w = geometric_graph(num, 0.9, seed)
    hop = csg.shortest_path(sp.csr_matrix(adj.astype(np.float64)), unweighted=True, indices=0)
    goal = int(np.argmax(np.where(np.isfinite(hop), hop, -1)))
    costs, edges, cg, it, _ = run_gmt(s, adj, w, np.zeros((0, 12)), goal)
    H = int(hop[goal])
    assert it == H
    reached = np.flatnonzero(edges >= 0)
    assert set(reached) == set(np.flatnonzero((hop >= 1) & (hop <= H)))
    for n in reached:
        p = edges[n]
        assert adj[p, n] and hop[p] == hop[n] - 1
        assert costs[n] == np.float32(costs[p] + w[p, n])
        prev = np.flatnonzero(adj[:, n] & (hop == hop[n] - 1))
        assert np.all(costs[n] <= (costs[prev] + w[prev, n]).astype(np.float32))
    dij = csg.dijkstra(sp.csr_matrix(np.where(adj, w, 0).astype(np.float64)), indices=0)
    assert cg == costs[goal] and cg >= dij[goal] * (1 - 1e-6)


@pytest.mark.parametrize("seed", [0, 5])
def test_gmt_kept_connections_are_free(seed):
    num = 300
    s, adj, w = geometric_graph(num, 0.9, seed)
    rng = np.random.default_rng(seed + 100)
    lo = rng.random((12, 3), dtype=np.float32) * np.float32(1.6)
    obs = np.concatenate([lo, np.zeros((12, 3), np.float32), lo + np.float32(0.3),
                          np.zeros((12, 3), np.float32)], 1).astype(np.float32)
    costs, edges, cg, it, g = run_gmt(s, adj, w, obs, num - 1)
    free_e = np.zeros(len(g["dst"]), np.uint8)
    O.po_free_edges(fp(obs), len(obs), fp(g["disc"]), lp(g["wp"]), len(g["dst"]), bp(free_e), 1)
    eid = {(int(a), int(b)): k for k, (a, b) in enumerate(zip(g["src"], g["dst"]))}
    reached = np.flatnonzero(edges >= 0)
    assert len(reached) > 0
    n_blocked = 0
    for n in reached:
        p = int(edges[n])
        assert free_e[eid[(p, int(n))]] == 1
        assert costs[n] == np.float32(costs[p] + w[p, n])
    for n in range(1, num):   # a free node left without a connection keeps (0, -1)
        if edges[n] == -1:
            assert costs[n] == 0
            n_blocked += 1
    c0, e0, _, _, _ = run_gmt(s, adj, w, np.zeros((0, 12)), num - 1)
    assert n_blocked > 0 and not np.array_equal(e0, edges)   # the boxes changed the tree
