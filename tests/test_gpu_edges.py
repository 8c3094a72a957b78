"""GPU parity of the fused edge pass (pump_edges_discretize_check: waypoints
of findDiscretizedPath, 2pBVP.cu:53-69, and the freeEdges flags,
collisionCheck.cu:5-37) on synthetic edge sets built to reach the kernel's
less common paths: more boxes than a workgroup stages (the remainder is
tested from global memory), workgroups whose waypoints overflow the
waypoint-to-edge map (binary search instead), edges longer than a tile, the
minimum of two waypoints per edge, and the fixed-W layout.

Bar: flags bit-exact; waypoints bit-exact except the rare last-place
double rounding of pow() (the same tolerance as the pipeline parity tests).
The oracle recomputes each edge's arrival time by its own bisection, as
findDiscretizedPath does; the edges' arrival times are taken from it.

The FP32-waypoint form (pump_edges_discretize_check_fp32) must give the same
flags bit for bit, and waypoints within its stated bound of the oracle's:
12 u S per position component, S = |p0| + |v0| tau + |c2| tau^2 + |c3| tau^3
of the axis's cubic Hermite segment, u = 2^-24."""
import numpy as np
import pytest

import oracle
import pump_amd as P
from test_gpu_parity import assert_close_f32, dev, host, ptr

pytestmark = pytest.mark.gpu

O = oracle.load()
fp, ip, lp, bp = oracle.fp, oracle.ip, oracle.lp, oracle.bp


def edge_set(seed, num, per_node, n_obs, w_lo, w_hi, side=(0.02, 0.08)):
    rng = np.random.default_rng(seed)
    s = np.concatenate([rng.uniform(0, 1, (num, 3)), rng.uniform(-0.5, 0.5, (num, 3))], 1)
    s = s.astype(np.float32)
    src = np.repeat(np.arange(num), per_node).astype(np.int32)
    dst = np.concatenate([np.sort(rng.choice(np.delete(np.arange(num), i), per_node,
                                             replace=False)) for i in range(num)]).astype(np.int32)
    row_start = np.arange(num + 1, dtype=np.int64) * per_node
    tau = np.array([O.po_topt_bisection(fp(s[a]), fp(s[b]), 5.0) for a, b in zip(src, dst)],
                   np.float32)
    w = rng.integers(w_lo, w_hi + 1, len(src)).astype(np.int64)
    off = np.zeros(len(src) + 1, np.int64)
    np.cumsum(w, out=off[1:])
    lo = rng.uniform(0.05, 0.95, (n_obs, 3))
    size = rng.uniform(side[0], side[1], (n_obs, 3))
    obs = np.concatenate([lo, np.full((n_obs, 3), -3.1), lo + size, np.full((n_obs, 3), 3.1)], 1)
    return dict(s=s, src=src, dst=dst, row_start=row_start, tau=tau, off=off,
                obs=obs.astype(np.float32).reshape(-1, 12))


def oracle_pass(d):
    n_wp = int(d["off"][-1])
    disc = np.zeros((n_wp, 6), np.float32)
    O.po_discretize_edges(fp(d["s"]), lp(d["row_start"]), ip(d["dst"]), len(d["s"]),
                          lp(d["off"]), fp(disc), 8)
    free = np.zeros(len(d["src"]), np.uint8)
    O.po_free_edges(fp(d["obs"]), len(d["obs"]), fp(disc), lp(d["off"]), len(d["src"]), bp(free),
                    8)
    return disc, free


def gpu_pass(d, fixed_w=0):
    E, n_wp = len(d["src"]), int(d["off"][-1])
    g = {k: dev(d[k]) for k in ("s", "src", "dst", "tau", "obs")}
    goff = dev(d["off"]) if not fixed_w else None
    disc = dev(np.zeros((n_wp, 6), np.float32))
    flags = dev(np.full(E, 9, np.uint8))
    P.check("edges", P.lib.pump_edges_discretize_check(
        ptr(g["s"]), ptr(g["src"]), ptr(g["dst"]), ptr(g["tau"]), E, ptr(goff), fixed_w,
        ptr(g["obs"]), len(d["obs"]), ptr(disc), ptr(flags), None))
    # the flags alone (no waypoint output)
    flags2 = dev(np.full(E, 9, np.uint8))
    P.check("edges", P.lib.pump_edges_discretize_check(
        ptr(g["s"]), ptr(g["src"]), ptr(g["dst"]), ptr(g["tau"]), E, ptr(goff), fixed_w,
        ptr(g["obs"]), len(d["obs"]), None, ptr(flags2), None))
    return host(disc), host(flags), host(flags2)


def gpu_pass32(d, fixed_w=0, write=True):
    E, n_wp = len(d["src"]), int(d["off"][-1])
    g = {k: dev(d[k]) for k in ("s", "src", "dst", "tau", "obs")}
    goff = dev(d["off"]) if not fixed_w else None
    disc = dev(np.zeros((n_wp, 6), np.float32))
    flags = dev(np.full(E, 9, np.uint8))
    cnt = dev(np.zeros(1, np.uint64))
    P.check("edges32", P.lib.pump_edges_discretize_check_fp32(
        ptr(g["s"]), ptr(g["src"]), ptr(g["dst"]), ptr(g["tau"]), E, ptr(goff), fixed_w,
        ptr(g["obs"]), len(d["obs"]), ptr(disc) if write else None, ptr(flags), ptr(cnt), None))
    return host(disc), host(flags), int(host(cnt)[0])


def fp32_bound(s, src, dst, tau, off):
    """Per waypoint row, per position axis: the stated bound 12 u S."""
    t = tau.astype(np.float64)
    a, b = s[src].astype(np.float64), s[dst].astype(np.float64)
    p0, v0, p1, v1 = a[:, :3], a[:, 3:], b[:, :3], b[:, 3:]
    c2 = 3 * (p1 - p0) / (t * t)[:, None] - (2 * v0 + v1) / t[:, None]
    c3 = 2 * (p0 - p1) / (t * t * t)[:, None] + (v0 + v1) / (t * t)[:, None]
    S = (np.abs(p0) + np.abs(v0) * t[:, None] + np.abs(c2) * (t * t)[:, None]
         + np.abs(c3) * (t * t * t)[:, None])
    eps = 12 * 2.0 ** -24 * S
    return np.repeat(eps, np.diff(off), axis=0)


def check_fp32_waypoints(disc32, disc_ref, eps):
    err = np.abs(disc32[:, :3].astype(np.float64) - disc_ref[:, :3])
    assert np.all(err <= eps), float((err - eps).max())
    vel = np.abs(disc32[:, 3:].astype(np.float64) - disc_ref[:, 3:])
    assert np.all(vel <= 1e-4 * (1 + np.abs(disc_ref[:, 3:]))), float(vel.max())
    return float((err / np.maximum(eps, 1e-30)).max())


@pytest.mark.parametrize("name,seed,per_node,n_obs,w_lo,w_hi", [
    ("300 boxes, past the staged 256", 1, 12, 300, 2, 40),
    ("long edges overflow the waypoint map", 2, 10, 40, 2, 400),
    ("two waypoints per edge", 3, 16, 60, 2, 2),
])
def test_fused_edge_pass_paths(name, seed, per_node, n_obs, w_lo, w_hi):
    d = edge_set(seed, 160, per_node, n_obs, w_lo, w_hi)
    print(name, "edges", len(d["src"]), "waypoints", int(d["off"][-1]), flush=True)
    disc_ref, free_ref = oracle_pass(d)
    assert 0 < free_ref.sum() < len(free_ref)   # both outcomes occur
    disc, flags, flags2 = gpu_pass(d)
    assert np.array_equal(flags, free_ref)
    assert np.array_equal(flags2, free_ref)
    assert_close_f32(disc, disc_ref)


def test_fused_edge_pass_fixed_w():
    d = edge_set(4, 120, 12, 50, 9, 9)
    disc_ref, free_ref = oracle_pass(d)
    disc, flags, _ = gpu_pass(d, fixed_w=9)
    assert np.array_equal(flags, free_ref)
    assert_close_f32(disc, disc_ref)


@pytest.mark.parametrize("name,seed,per_node,n_obs,w_lo,w_hi,side", [
    ("300 boxes, past the staged 256", 1, 12, 300, 2, 40, (0.02, 0.08)),
    ("long edges overflow the waypoint map", 2, 10, 40, 2, 400, (0.02, 0.08)),
    ("two waypoints per edge", 3, 16, 60, 2, 2, (0.02, 0.08)),
    ("dense large boxes", 5, 12, 120, 2, 30, (0.05, 0.25)),
])
def test_fp32_edge_pass(name, seed, per_node, n_obs, w_lo, w_hi, side):
    d = edge_set(seed, 160, per_node, n_obs, w_lo, w_hi, side=side)
    disc_ref, free_ref = oracle_pass(d)
    assert 0 < free_ref.sum() < len(free_ref)
    disc32, flags32, retests = gpu_pass32(d)
    _, flags_only, _ = gpu_pass32(d, write=False)
    print(name, "edges", len(d["src"]), "blocked", int((free_ref == 0).sum()), "exact re-tests",
          retests, flush=True)
    assert np.array_equal(flags32, free_ref)
    assert np.array_equal(flags_only, free_ref)
    worst = check_fp32_waypoints(disc32, disc_ref,
                                 fp32_bound(d["s"], d["src"], d["dst"], d["tau"], d["off"]))
    print("largest error / bound", worst, flush=True)


def test_fp32_edge_pass_faces_on_waypoints():
    """Boxes whose faces pass through the reference's waypoints (exactly, or
    one float step off): the segments there sit on the boundary of the test,
    where only the exact re-test may decide."""
    d = edge_set(6, 160, 12, 1, 2, 30)
    disc_ref, _ = oracle_pass(d)
    rng = np.random.default_rng(6)
    n_obs = 150
    rows = rng.choice(len(disc_ref), n_obs, replace=False)
    size = rng.uniform(0.02, 0.1, (n_obs, 3)).astype(np.float32)
    lo = (disc_ref[rows, :3] - size * rng.uniform(0, 1, (n_obs, 3))).astype(np.float32)
    hi = (lo + size).astype(np.float32)
    # one face per box on the waypoint: lo or hi, exactly or one step off
    ax = rng.integers(0, 3, n_obs)
    side = rng.integers(0, 2, n_obs)
    step = rng.integers(-1, 2, n_obs)
    for i in range(n_obs):
        c = disc_ref[rows[i], ax[i]]
        c = np.nextafter(c, np.float32(np.inf * step[i])) if step[i] else c
        if side[i]:
            hi[i, ax[i]] = c
            lo[i, ax[i]] = min(lo[i, ax[i]], np.nextafter(c, np.float32(-np.inf)))
        else:
            lo[i, ax[i]] = c
            hi[i, ax[i]] = max(hi[i, ax[i]], np.nextafter(c, np.float32(np.inf)))
    obs = np.concatenate([lo, np.full((n_obs, 3), -3.1), hi, np.full((n_obs, 3), 3.1)], 1)
    d["obs"] = obs.astype(np.float32)
    disc_ref, free_ref = oracle_pass(d)
    assert 0 < free_ref.sum() < len(free_ref)
    _, flags32, retests = gpu_pass32(d)
    _, flags, _ = gpu_pass(d)
    print("edges", len(d["src"]), "blocked", int((free_ref == 0).sum()), "exact re-tests",
          retests, flush=True)
    assert np.array_equal(flags, free_ref)
    assert np.array_equal(flags32, free_ref)
    assert retests > 0


def test_fp32_edge_pass_fixed_w():
    d = edge_set(4, 120, 12, 50, 9, 9)
    disc_ref, free_ref = oracle_pass(d)
    disc32, flags32, _ = gpu_pass32(d, fixed_w=9)
    assert np.array_equal(flags32, free_ref)
    check_fp32_waypoints(disc32, disc_ref,
                         fp32_bound(d["s"], d["src"], d["dst"], d["tau"], d["off"]))
