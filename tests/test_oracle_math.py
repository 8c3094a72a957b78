"""Pins of the CPU oracle by what the mathematics fixes: closed forms,
brute-force geometry, invariants and independent FP64 evaluations.  Each
check is chosen so that a plausible slip in the restatement (dropped term,
wrong sign, swapped lo/hi, transposed index) fails it."""
import ctypes

import numpy as np
import pytest

import oracle
from oracle import bp, dp, fp, ip, lp

O = oracle.load()
G = np.float32(-9.81)


def rand_state(rng, pos=(0, 1), vel=(-0.5, 0.5)):
    return np.concatenate([rng.uniform(*pos, 3), rng.uniform(*vel, 3)]).astype(np.float32)


# ---------------------------------------------------------------- 2pBVP ----
def closed_form_cost(tau, a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    dp_ = b[:3] - a[:3]
    D = (dp_ ** 2).sum()
    S = (dp_ * (a[3:] + b[3:])).sum()
    R = (a[3:] ** 2 + a[3:] * b[3:] + b[3:] ** 2).sum()
    g = float(G)
    return 12 * D / tau ** 3 - 12 * S / tau ** 2 + 4 * R / tau + 2 * g * (a[5] - b[5]) + (
        1 + g * g) * tau


def test_cost_matches_closed_form():
    rng = np.random.default_rng(1)
    for _ in range(4000):
        a, b = rand_state(rng), rand_state(rng)
        tau = float(rng.uniform(0.05, 5))
        c = O.po_cost(np.float32(tau), fp(a), fp(b))
        ref = closed_form_cost(float(np.float32(tau)), a, b)
        # float-rounded terms of magnitude ~ the individual summands
        scale = abs(ref) + 12 * 3 / tau ** 3 + 100 * tau
        assert abs(c - ref) <= 2e-6 * scale, (c, ref)
        assert abs(O.po64_cost(float(np.float32(tau)), fp(a), fp(b)) - ref) <= 1e-12 * scale


def float_term_mass(tau, a, b):
    """Sum of |float-rounded terms| of the cost numerator (2pBVP.cu:100-105):
    their rounding, scaled by 1/tau^3, bounds the cost's deviation from the
    closed form beyond its final rounding."""
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    m = 24 * abs(a[1] * b[1])
    m += 12 * tau * (abs(a[4] * a[1]) + abs(a[1] * b[4]) + abs(a[3] * b[0]) + abs(b[3] * b[0]) +
                     abs(a[5] * b[2]) + abs(b[5] * b[2]) + abs(a[4] * b[1]) + abs(b[4] * b[1]))
    m += 4 * tau ** 2 * (abs(a[4] * b[4]) + abs(a[3] * b[3]) + abs(a[5] * b[5]))
    m += 2 * 9.81 * tau ** 3 * (abs(a[5]) + abs(b[5]))
    m += 12 * abs(a[0]) * (2 * abs(b[0]) + abs(a[3] + b[3]) * tau)
    m += 12 * abs(a[2]) * (2 * abs(b[2]) + abs(a[5] + b[5]) * tau)
    return m / tau ** 3


def test_dcost_is_derivative():
    rng = np.random.default_rng(2)
    h = np.float32(0.000001)
    for _ in range(500):
        a, b = rand_state(rng), rand_state(rng)
        tau = np.float32(rng.uniform(0.3, 4))
        d = O.po_dcost(tau, fp(a), fp(b))
        # the evaluation points are the float sums tau +- h/2
        tp = float(np.float32(tau + h / np.float32(2)))
        tm = float(np.float32(tau - h / np.float32(2)))
        ref = (closed_form_cost(tp, a, b) - closed_form_cost(tm, a, b)) / float(h)
        c = abs(closed_form_cost(float(tau), a, b))
        err_cost = float(np.spacing(np.float32(c))) + 3 * 2.0 ** -24 * float_term_mass(
            float(tau), a, b)
        assert abs(d - ref) <= 2 * err_cost / float(h) + 1e-4 * abs(ref)


def exact_dcost(tau, a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    dp_ = b[:3] - a[:3]
    D = (dp_ ** 2).sum()
    S = (dp_ * (a[3:] + b[3:])).sum()
    R = (a[3:] ** 2 + a[3:] * b[3:] + b[3:] ** 2).sum()
    g = float(G)
    return -36 * D / tau ** 4 + 24 * S / tau ** 3 - 4 * R / tau ** 2 + 1 + g * g


def test_topt_is_stationary_point():
    """toptBisection returns tmax when c'(tmax) < 0, else a point where the
    exact derivative is within the float finite-difference noise of zero, and
    its cost is within that noise of the FP64 minimum."""
    rng = np.random.default_rng(3)
    h = 1e-6
    for _ in range(2000):
        a, b = rand_state(rng), rand_state(rng)
        t = float(O.po_topt_bisection(fp(a), fp(b), 5.0))
        t64 = O.po64_topt(fp(a), fp(b), 5.0)
        if t == 5.0:
            assert exact_dcost(5.0, a, b) < 1.0 and t64 == 5.0
            continue
        c = abs(closed_form_cost(t, a, b))
        noise = 2 * (float(np.spacing(np.float32(c))) + 3 * 2.0 ** -24 *
                     float_term_mass(t, a, b)) / h
        # half-ulp steps of tau +- h/2 change the difference quotient too
        noise += abs(exact_dcost(t, a, b)) * 0.5
        assert abs(exact_dcost(t, a, b)) <= 2 * noise + 1e-3, (t, t64)
        assert closed_form_cost(t, a, b) >= closed_form_cost(t64, a, b) - 1e-9 * c


def test_topt_tmax_branch():
    # a pair far apart with opposing velocities needs more than tmax: the
    # bisection returns tmax exactly when c'(tmax) < 0
    a = np.array([0, 0, 0, -0.5, -0.5, -0.5], np.float32)
    b = np.array([1, 1, 1, -0.5, -0.5, -0.5], np.float32)
    t = O.po_topt_bisection(fp(a), fp(b), np.float32(0.5))
    assert t == np.float32(0.5)
    assert O.po_dcost(np.float32(0.5), fp(a), fp(b)) < 0


def test_path_point_endpoints_and_velocity():
    rng = np.random.default_rng(4)
    out0, out1, outm, outp = (np.zeros(6, np.float32) for _ in range(4))
    for _ in range(500):
        a, b = rand_state(rng), rand_state(rng)
        tau = np.float32(rng.uniform(0.2, 4))
        O.po_path_point(np.float32(0), tau, fp(a), fp(b), fp(out0))
        O.po_path_point(tau, tau, fp(a), fp(b), fp(out1))
        assert np.allclose(out0, a, atol=2e-6)
        assert np.allclose(out1, b, atol=5e-5)
        t = np.float32(rng.uniform(0.1, 0.9) * tau)
        h = np.float32(1e-3)
        O.po_path_point(t - h, tau, fp(a), fp(b), fp(outm))
        O.po_path_point(t + h, tau, fp(a), fp(b), fp(outp))
        vel_fd = (outp[:3].astype(np.float64) - outm[:3]) / (2 * float(h))
        O.po_path_point(t, tau, fp(a), fp(b), fp(out0))
        assert np.allclose(vel_fd, out0[3:], atol=5e-3)


def hermite64(t, tau, a, b):
    """The double integrator's minimum-energy trajectory between states a and
    b over [0, tau] in FP64: the cubic Hermite segment through (p0, v0) and
    (p1, v1), position and velocity."""
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    p0, v0, p1, v1 = a[:3], a[3:], b[:3], b[3:]
    c2 = 3 * (p1 - p0) / tau ** 2 - (2 * v0 + v1) / tau
    c3 = 2 * (p0 - p1) / tau ** 3 + (v0 + v1) / tau ** 2
    pos = p0 + v0 * t + c2 * t ** 2 + c3 * t ** 3
    vel = v0 + 2 * c2 * t + 3 * c3 * t ** 2
    return np.concatenate([pos, vel])


def test_path_point_is_cubic_hermite():
    """pathPoint (2pBVP.cu:117-131) evaluates the cubic Hermite segment: every
    component against the FP64 closed form (a dropped term, a wrong sign or a
    swapped coefficient of any of the six expressions fails)."""
    rng = np.random.default_rng(41)
    out = np.zeros(6, np.float32)
    worst = 0.0
    for _ in range(2000):
        a, b = rand_state(rng), rand_state(rng)
        tau = np.float32(rng.uniform(0.2, 4.5))
        t = np.float32(rng.uniform(0, 1) * tau)
        O.po_path_point(t, tau, fp(a), fp(b), fp(out))
        ref = hermite64(float(t), float(tau), a, b)
        worst = max(worst, float(np.max(np.abs(out - ref) / (1 + np.abs(ref)))))
    assert worst < 2e-6, worst


@pytest.mark.parametrize("num_disc", [1, 2, 8, 19])
def test_find_discretized_path_spacing(num_disc):
    """findDiscretizedPath (2pBVP.cu:53-69): row 0 and row numDisc are exact
    copies of the end states, row i is the trajectory at i * (tau/numDisc)
    with tau the bisected arrival time, which it returns."""
    rng = np.random.default_rng(42 + num_disc)
    wp = np.zeros((num_disc + 1, 6), np.float32)
    for _ in range(300):
        a, b = rand_state(rng), rand_state(rng)
        tau = O.po_find_discretized_path(fp(wp), fp(a), fp(b), num_disc)
        assert np.float32(tau) == np.float32(O.po_topt_bisection(fp(a), fp(b), np.float32(5)))
        assert np.array_equal(wp[0].view(np.uint32), a.view(np.uint32))
        assert np.array_equal(wp[num_disc].view(np.uint32), b.view(np.uint32))
        step = float(np.float32(np.float32(tau) / np.float32(num_disc)))
        for i in range(1, num_disc):
            ref = hermite64(float(np.float32(step * i)), float(tau), a, b)
            assert np.allclose(wp[i], ref, rtol=0, atol=5e-6 * (1 + np.abs(ref)).max()), i


def test_find_optimal_path_samples_the_global_time_grid():
    """findOptimalPath (2pBVP.cu:5-36): hop i gets T = (int)((tau - toff)/dt +
    1) rows at local times toff + k*dt and carries toff = (t_last - tau) + dt,
    so that row r of the whole path is the trajectory at global time r*dt and
    there are floor(sum tau / dt) + 1 rows.  Checked against the FP64
    trajectory of the chained hops; one row more or less in a hop, or a carry
    without its + dt, shifts every later row by dt and fails."""
    rng = np.random.default_rng(43)
    cap = 4096
    split = np.zeros((cap, 6), np.float32)
    ln = np.zeros(1, np.int32)
    checked = 0
    for trial in range(60):
        n = int(rng.integers(2, 7))
        xs = np.stack([rand_state(rng) for _ in range(n)]).astype(np.float32)
        dt = np.float32([0.05, 0.1, 0.0625][trial % 3])
        taus = [float(O.po_topt_bisection(fp(xs[k]), fp(xs[k + 1]), np.float32(5)))
                for k in range(n - 1)]
        if min(taus) <= 2 * float(dt):
            continue  # a hop shorter than its carry gets no rows (a quirk of the loop)
        total = O.po_find_optimal_path(dt, fp(split), fp(xs), n, ip(ln), cap)
        assert np.float32(total) == np.float32(sum(np.float32(t) for t in taus)) or \
            abs(total - sum(taus)) < 1e-5
        cum = np.concatenate([[0.0], np.cumsum(taus)])
        frac = cum[-1] / float(dt) - np.floor(cum[-1] / float(dt))
        if 0.01 < frac < 0.99:
            assert ln[0] == int(np.floor(cum[-1] / float(dt))) + 1, (ln[0], cum[-1])
        for r in range(ln[0]):
            g = r * float(dt)
            k = min(int(np.searchsorted(cum, g, side="right")) - 1, n - 2)
            ref = hermite64(g - cum[k], taus[k], xs[k], xs[k + 1])
            assert np.allclose(split[r], ref, rtol=0, atol=2e-5 * (1 + np.abs(ref)).max()), \
                (trial, r, k)
        checked += 1
    assert checked >= 40


# ----------------------------------------------------- collision checks ----
def box(lo, hi):
    o = np.zeros(12, np.float32)
    o[:3], o[6:9] = lo, hi
    o[3:6], o[9:12] = -3.1, 3.1
    return o


def test_sample_free_strict_interior():
    obs = np.stack([box([0.2, 0.2, 0.2], [0.4, 0.4, 0.4])])
    pts = np.array([[0.3, 0.3, 0.3, 9, 9, 9],      # inside (velocity ignored)
                    [0.2, 0.3, 0.3, 0, 0, 0],      # on lo face -> free (strict)
                    [0.3, 0.4, 0.3, 0, 0, 0],      # on hi face -> free
                    [0.1, 0.3, 0.3, 0, 0, 0],      # outside in x
                    [0.3, 0.3, 0.41, 0, 0, 0]],    # outside in z
                   np.float32)
    out = np.zeros(5, np.uint8)
    O.po_sample_free(fp(obs), 1, fp(pts), 5, bp(out))
    assert out.tolist() == [0, 1, 1, 1, 1]


def seg_box_fp64(v, w, lo, hi):
    """Slab intersection of segment v->w with the closed box; returns
    (hits, robust) where robust means the answer is unchanged under a small
    inflation/deflation of the box."""
    def clip(eps):
        t0, t1 = 0.0, 1.0
        for d in range(3):
            u = w[d] - v[d]
            l, h = lo[d] - eps, hi[d] + eps
            if abs(u) < 1e-15:
                if v[d] < l or v[d] > h:
                    return False
                continue
            a, b = (l - v[d]) / u, (h - v[d]) / u
            if a > b:
                a, b = b, a
            t0, t1 = max(t0, a), min(t1, b)
            if t0 > t1:
                return False
        return True
    hit = clip(0.0)
    return hit, clip(1e-4) == clip(-1e-4) == hit


def test_segment_box_brute_force():
    rng = np.random.default_rng(5)
    lo = np.array([0.4, 0.4, 0.4])
    hi = np.array([0.6, 0.7, 0.5])
    obs = box(lo, hi)
    n_checked = n_hits = 0
    for _ in range(20000):
        v = rng.uniform(0.2, 0.9, 3).astype(np.float32)
        w = (v + rng.normal(0, 0.15, 3)).astype(np.float32)
        v64, w64 = v.astype(np.float64), w.astype(np.float64)
        inside_v = np.all(v64 >= lo - 1e-4) and np.all(v64 <= hi + 1e-4)
        hit, robust = seg_box_fp64(v64, w64, lo, hi)
        if inside_v or not robust:
            continue
        bmin, bmax = np.minimum(v, w), np.maximum(v, w)
        valid = O.po_is_motion_valid(fp(v), fp(w), fp(bmin), fp(bmax), 1, fp(obs))
        n_checked += 1
        n_hits += hit
        # from outside, the line-face test reports exactly the geometric
        # intersections of the segment with the box
        assert valid == (0 if hit else 1), (v, w)
    assert n_checked > 10000 and n_hits > 500


def test_free_edges_is_and_over_segments():
    rng = np.random.default_rng(6)
    obs = np.stack([box([0.4, 0.4, 0.4], [0.6, 0.6, 0.6])])
    W = 5
    n = 300
    disc = rng.uniform(0.0, 1.0, (n * W, 6)).astype(np.float32)
    off = np.arange(0, (n + 1) * W, W, dtype=np.int64)
    flags = np.zeros(n, np.uint8)
    O.po_free_edges(fp(obs), 1, fp(disc), lp(off), n, bp(flags), 1)
    for e in range(n):
        ok = 1
        for k in range(W - 1):
            v, w = disc[e * W + k], disc[e * W + k + 1]
            bmin, bmax = np.minimum(v, w), np.maximum(v, w)
            ok &= O.po_is_motion_valid(fp(v.copy()), fp(w.copy()), fp(bmin), fp(bmax), 1, fp(obs))
        assert flags[e] == ok


# --------------------------------------------------------- half-spaces ----
def test_cache_halfspaces_geometry():
    rng = np.random.default_rng(7)
    n_obs = 12
    obs = np.stack([box(c - s / 2, c + s / 2) for c, s in
                    zip(rng.uniform(0.2, 0.8, (n_obs, 3)), rng.uniform(0.05, 0.2, (n_obs, 3)))])
    n_wp = 4000
    disc = np.concatenate([rng.uniform(0, 1, (n_wp, 3)), rng.uniform(-0.5, 0.5, (n_wp, 3))],
                          1).astype(np.float32)
    disc[:50, 3:] = 0  # stationary waypoints (vdotv == 0 branch)
    off = np.array([0, n_wp], np.int64)
    free = np.ones(1, np.uint8)
    max_hs = 5
    maxd2 = O.po_maxd2(np.float32(0.5))
    as_ = np.zeros((n_wp * max_hs, 3), np.float32)
    bs = np.zeros(n_wp * max_hs, np.float32)
    cnt = np.zeros(n_wp, np.int32)
    O.po_setup_as_bs(n_wp * max_hs, fp(bs), fp(as_))
    O.po_cache_halfspaces(1, fp(disc), lp(off), bp(free), n_obs, max_hs, fp(obs), fp(as_), fp(bs),
                          ip(cnt), maxd2, 1)
    lo, hi = obs[:, :3].astype(np.float64), obs[:, 6:9].astype(np.float64)
    n_stored = 0
    for g in range(n_wp):
        pos, vel = disc[g, :3].astype(np.float64), disc[g, 3:].astype(np.float64)
        x = np.clip(pos, lo, hi) - pos
        d2 = (x ** 2).sum(1)
        near = np.where(d2 <= maxd2)[0]
        ambiguous = np.any(np.abs(d2 - maxd2) < 1e-6)
        if not ambiguous:
            assert cnt[g] == min(len(near), max_hs)
            for h, o in enumerate(near[:max_hs]):
                a = as_[g * max_hs + h].astype(np.float64)
                vv = vel if (vel ** 2).sum() > 0 else None
                xo = x[o]
                if vv is not None:
                    expect = xo - (xo @ vv) / (vv @ vv) * vv
                    assert abs(a @ vv) <= 1e-5 * (np.linalg.norm(xo) * np.linalg.norm(vv) + 1e-12)
                else:
                    expect = xo
                assert np.allclose(a, expect, atol=1e-6)
                assert abs(bs[g * max_hs + h] - a @ a) <= 1e-6 * (a @ a) + 1e-12
        for h in range(cnt[g], max_hs):
            assert bs[g * max_hs + h] == 1111 and np.all(as_[g * max_hs + h] == -1)
        n_stored += cnt[g]
    assert n_stored > 100


def test_waypoint_inside_box_blocks_every_particle():
    # x = 0 => a = 0, b = 0 => LHS*m < 0 is false for every particle
    obs = np.stack([box([0.4, 0.4, 0.4], [0.6, 0.6, 0.6])])
    disc = np.array([[0.5, 0.5, 0.5, 0.1, 0, 0], [0.52, 0.5, 0.5, 0.1, 0, 0]], np.float32)
    off = np.array([0, 2], np.int64)
    as_ = np.zeros((10, 3), np.float32)
    bs = np.zeros(10, np.float32)
    cnt = np.zeros(2, np.int32)
    O.po_setup_as_bs(10, fp(bs), fp(as_))
    O.po_cache_halfspaces(1, fp(disc), lp(off), bp(np.ones(1, np.uint8)), 1, 5, fp(obs), fp(as_),
                          fp(bs), ip(cnt), np.float32(0.1), 1)
    assert cnt.tolist() == [1, 1] and bs[0] == 0 and bs[5] == 0
    n_p, n_mc, T = 64, 64, 4
    rng = np.random.default_rng(8)
    offs = rng.normal(0, 0.05, (T, n_mc, 3)).astype(np.float32)
    valid = np.ones((2, n_p), np.uint8)
    times = np.zeros(2, np.float32)
    topts = np.array([0.1], np.float32)
    O.po_propagate_valid_particles(1, lp(off), ip(np.zeros(1, np.int32)),
                                   ip(np.zeros(1, np.int32)), 1, fp(as_), fp(bs), fp(times),
                                   fp(topts), np.float32(0.05), 5, fp(offs), T, np.float32(1.0),
                                   n_mc, bp(valid), n_p, 1)
    assert valid[1].sum() == 0 and valid[0].sum() == n_p


# ---------------------------------------------------------------- HSMC ----
def test_propagate_against_fp64_and_monotone():
    rng = np.random.default_rng(9)
    n_edges, W, max_hs, n_p, n_mc, T = 40, 6, 3, 200, 256, 50
    off = np.arange(0, (n_edges + 1) * W, W, dtype=np.int64)
    hs = n_edges * W * max_hs
    as_ = rng.normal(0, 0.05, (hs, 3)).astype(np.float32)
    bs = (np.abs(rng.normal(0, 0.05, hs)) ** 2 * 0.5).astype(np.float32)
    bs[rng.uniform(size=hs) < 0.5] = 1111
    offs = rng.normal(0, 0.05, (T, n_mc, 3)).astype(np.float32)
    topts = rng.uniform(0.3, 1.5, n_edges).astype(np.float32)
    n_parent = 5
    valid = np.zeros((n_parent + 30, n_p), np.uint8)
    valid[:n_parent] = rng.uniform(size=(n_parent, n_p)) < 0.9
    times = np.zeros(n_parent + 30, np.float32)
    times[:n_parent] = rng.uniform(0, 0.5, n_parent)
    wf_edge = rng.integers(0, n_edges, 30).astype(np.int32)
    wf_prev = rng.integers(0, n_parent, 30).astype(np.int32)
    om = np.float32(0.7)
    dt = np.float32(0.05)
    O.po_propagate_valid_particles(30, lp(off), ip(wf_edge), ip(wf_prev), n_parent, fp(as_),
                                   fp(bs), fp(times), fp(topts), dt, max_hs, fp(offs), T, om,
                                   n_mc, bp(valid), n_p, 1)
    amb = 0
    for w in range(30):
        e, pr = wf_edge[w], wf_prev[w]
        expect = valid[pr].astype(bool).copy()
        near = np.zeros(n_p, bool)
        dtl = np.float32(topts[e] / np.float32(W - 1))
        for i in range(W):
            ti = int(np.float32(times[pr] + np.float32(i) * dtl) / dt)
            for h in range(max_hs):
                k = (e * W + i) * max_hs + h
                if bs[k] >= 1000:
                    continue
                lhs = offs[ti, :n_p].astype(np.float64) @ as_[k].astype(np.float64) * float(om)
                expect &= lhs < bs[k]
                near |= np.abs(lhs - bs[k]) < 1e-6
        got = valid[n_parent + w].astype(bool)
        assert np.all(got[~near] == expect[~near])
        assert not np.any(got & ~valid[pr].astype(bool))  # monotone
        amb += near.sum()
    assert amb < 30 * n_p * 0.01


def test_calculate_cp_counts():
    rng = np.random.default_rng(10)
    valid = (rng.uniform(size=(7, 333)) < 0.8).astype(np.uint8)
    cp = np.zeros(7, np.float32)
    O.po_calculate_cp_hsmc(5, 2, fp(cp), bp(valid), 333)
    for w in range(5):
        assert cp[2 + w] == np.float32(np.float32((valid[2 + w] == 0).sum()) / np.float32(333))


# ------------------------------------------------------------- full MC ----
def test_mccp_zero_noise_and_fp64():
    rng = np.random.default_rng(11)
    obs = np.stack([box([0.45, 0.45, 0.0], [0.55, 0.55, 1.0])])
    T, n_mc, tsteps = 40, 128, 60
    path = np.zeros((T, 6), np.float32)
    path[:, 0] = np.linspace(0.1, 0.9, T)
    path[:, 1] = 0.5
    path[:, 2] = 0.5
    xc0 = np.zeros((n_mc, tsteps, 3), np.float32)
    hits = O.po_mccp_hits(fp(obs), 1, fp(xc0), tsteps, n_mc, np.float32(1.0), fp(path), T, None)
    assert hits == n_mc  # nominal path crosses the wall
    path[:, 1] = 0.2
    hits = O.po_mccp_hits(fp(obs), 1, fp(xc0), tsteps, n_mc, np.float32(1.0), fp(path), T, None)
    assert hits == 0
    # the final segment is skipped (collisionProbability.cu:50)
    path2 = path.copy()
    path2[-1, :3] = (0.5, 0.5, 0.5)
    hits = O.po_mccp_hits(fp(obs), 1, fp(xc0), tsteps, n_mc, np.float32(1.0), fp(path2), T, None)
    assert hits == 0
    # noisy: float and FP64 twins agree up to boundary particles
    path[:, 1] = 0.38
    xc = rng.normal(0, 0.05, (n_mc, tsteps, 3)).astype(np.float32)
    h32 = O.po_mccp_hits(fp(obs), 1, fp(xc), tsteps, n_mc, np.float32(1.0), fp(path), T, None)
    h64 = O.po64_mccp_hits(fp(obs), 1, fp(xc), tsteps, n_mc, 1.0,
                           dp(path.astype(np.float64)), T)
    assert 0 < h32 < n_mc and abs(h32 - h64) <= 2


# -------------------------------------------------------- rn and edges ----
def test_rn_and_edges_against_numpy():
    rng = np.random.default_rng(12)
    num = 120
    s = np.concatenate([rng.uniform(0, 1, (num, 3)), rng.uniform(-0.5, 0.5, (num, 3))],
                       1).astype(np.float32)
    n_pairs = num * (num - 1)
    c = np.empty(n_pairs, np.float32)
    t = np.empty(n_pairs, np.float32)
    O.po_fill_copts_topts(fp(s), num, 5.0, fp(c), fp(t), 2)
    rn_perc = np.float32(20 / (num - 1))
    k = O.po_rn_index(num, rn_perc)
    assert k == int(np.float32(np.float32(n_pairs) * rn_perc))
    rn = O.po_select_rn(fp(c), n_pairs, k)
    assert rn == np.sort(c)[k]
    rs = np.empty(num + 1, np.int64)
    E = O.po_count_edges(fp(c), num, rn, lp(rs))
    assert E == (c < rn).sum()
    dst = np.empty(E, np.int32)
    ce = np.empty(E, np.float32)
    te = np.empty(E, np.float32)
    O.po_fill_edges(fp(c), fp(t), num, rn, lp(rs), ip(dst), fp(ce), fp(te))
    # brute force in (i, j) order over the dense matrix
    e = 0
    for i in range(num):
        for j in range(num):
            if i == j:
                continue
            idx = i * (num - 1) + j - (1 if i < j else 0)
            if c[idx] < rn:
                assert dst[e] == j and ce[e] == c[idx] and te[e] == t[idx]
                e += 1
    assert e == E
    m = O.po_max_nn_size(lp(rs), ip(dst), E, num)
    indeg = np.bincount(dst, minlength=num)
    assert m == max(np.diff(rs).max(), indeg.max())
    go = np.empty((num, m), np.int32)
    come = np.empty((num, m), np.int32)
    dist = np.empty((num, m), np.float32)
    O.po_nn_tables(lp(rs), ip(dst), fp(ce), num, m, ip(go), ip(come), fp(dist))
    src = np.repeat(np.arange(num), np.diff(rs))
    for j in range(num):
        ins = src[dst == j]
        assert np.array_equal(come[j, :len(ins)], ins) and np.all(come[j, len(ins):] == -1)
        assert np.array_equal(dist[j, :len(ins)], ce[dst == j])


# --------------------------------------------------------------- search ----
@pytest.mark.parametrize("om", [0.5, 1.0])
def test_search_invariants(om):
    import pump_amd as P
    from pump_amd import configs
    import oracle.pipeline as OP
    cfg = configs.small(num=300, n_obs=12, n_p=256, n_mc=512, offset_mult=om, k=30)
    s, o = P.gen_problem(cfg["spec"])
    ot, xc = P.make_noise(cfg["noise"])
    rm = OP.roadmap(s, o, s[-1], cfg["params"], nthreads=4)
    out = OP.search(rm, s, s[-1], o, ot, xc, cfg["params"], nthreads=4)
    assert out["status"] in (0, 1)
    if out["soln_path_idx"] < 0:
        pytest.skip("no goal path in this instance")
    nodes = out["soln_nodes"]
    assert nodes[0] == 0 and nodes[-1] == s.shape[0] - 1
    rs, dst = rm["row_start"], rm["edge_dst"]
    cost = np.float32(0)
    for a, b in zip(nodes[:-1], nodes[1:]):
        row = dst[rs[a]:rs[a + 1]]
        hit = np.where(row == b)[0]
        assert len(hit) == 1
        e = rs[a] + hit[0]
        assert rm["is_free_edges"][e] and rm["is_free_samples"][b]
        cost = np.float32(cost + rm["copts_edge"][e])
    assert cost == np.float32(out["soln_cost"])
    if out["status"] == 0:
        assert out["soln_cp"] < cfg["params"].cp_target


def test_fill_rows_matches_whole_table():
    """po_fill_rows (sampled source rows, used by the full-size GPU checks)
    equals the corresponding rows of po_fill_copts_topts."""
    rng = np.random.default_rng(12)
    num = 150
    s = np.concatenate([rng.uniform(0, 1, (num, 3)), rng.uniform(-0.5, 0.5, (num, 3))],
                       1).astype(np.float32)
    c = np.zeros(num * (num - 1), np.float32)
    t = np.zeros(num * (num - 1), np.float32)
    O.po_fill_copts_topts(oracle.fp(s), num, 5.0, oracle.fp(c), oracle.fp(t), 4)
    rows = np.array([0, 7, 80, 149], np.int32)
    rc = np.zeros((rows.size, num - 1), np.float32)
    rt = np.zeros((rows.size, num - 1), np.float32)
    O.po_fill_rows(oracle.fp(s), num, 5.0, oracle.ip(rows), rows.size, oracle.fp(rc),
                   oracle.fp(rt), 4)
    for k, r in enumerate(rows):
        assert np.array_equal(rc[k], c[r * (num - 1):(r + 1) * (num - 1)])
        assert np.array_equal(rt[k], t[r * (num - 1):(r + 1) * (num - 1)])
