"""CPU-side checks of the product library: it loads, exports every symbol the
headers declare, and its host logic (Philox, problem and noise generation,
the host 2pBVP arithmetic used by findOptimalPath) is right.  No GPU calls."""
import os
import re

import numpy as np
import pytest

import pump_amd as P
from pump_amd import configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in ("pump_abi.h", "pump_host.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(pump_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    names = declared_symbols()
    assert len(names) > 40
    missing = [n for n in sorted(names) if not hasattr(P.lib, n)]
    assert not missing, missing


def test_status_strings_and_version():
    assert P.lib.pump_abi_version() == 1
    assert P.lib.pump_status_string(0) == b"ok"
    assert P.lib.pump_status_string(-5).startswith(b"argument")


def test_invalid_arguments_rejected_without_gpu():
    # argument validation happens before any device call
    assert P.lib.pump_fill_copts_topts(None, 100, 5.0, None, None, None) == -1
    assert P.lib.pump_select_rn(None, 10, 3, None, None, 0, None) == -1
    assert P.lib.pump_propagate_cp(-1, 9, None, None, None, None, None, None, None, None, 0.05, 5,
                                   None, 10, 1.0, 64, None, None, 0, 64, None, 0, None, None, 0,
                                   None) == -1
    cg, it = np.zeros(1, np.float32), np.zeros(1, np.int32)
    assert P.lib.pump_gmt_wavefront(None, 0, None, None, None, 4, None, None, 0, None, None, None,
                                    None, 100, None, None, 99, None, 0, P.fp(cg), P.ip(it),
                                    None) == -1
    assert P.lib.pump_precomp_nn(None, 10, 0.25, 4, None, None, None) == -1
    assert P.lib.pump_gmt_workspace(0) == 0


def test_noalloc_rejects_oversize_wavefront_without_gpu():
    # a wavefront whose batched workspace exceeds the fixed 64 MiB device array
    # is rejected before any device call
    assert P.lib.pump_propagate_workspace(566564, 2001) <= 64 << 20
    assert P.lib.pump_propagate_workspace(566565, 2001) > 64 << 20
    assert P.lib.pump_propagate_cp_bits_noalloc(
        566565, 9, None, None, None, None, None, None, None, None, 0.05, 5, None, 2001, 1.0, 64,
        None, None, 0, 64, None, 0, 0, None, None, None, None) == -3


def test_philox_known_answers():
    # Random123 known-answer vectors for philox4x32-10
    assert P.philox4x32([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert P.philox4x32([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E,
                                                              0xA20BC7C6, 0x6D5451FD]
    assert P.philox4x32([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                        [0xA4093822, 0x299F31D0]) == [0xD16CFE09, 0x94FDCCEB, 0x5001E420,
                                                      0x24126EA1]


def test_problem_generator_deterministic_and_bounded():
    cfg = configs.config(0)
    s1, o1 = P.gen_problem(cfg["spec"])
    s2, o2 = P.gen_problem(cfg["spec"])
    assert np.array_equal(s1, s2) and np.array_equal(o1, o2)
    assert s1.shape == (4000, 6) and o1.shape == (10, 12)
    assert np.array_equal(s1[0], np.float32([0.1, 0.1, 0.1, 0, 0, 0]))
    assert np.array_equal(s1[-1], np.float32([0.9, 0.9, 0.9, 0, 0, 0]))
    assert np.all((s1[1:-1, :3] >= 0) & (s1[1:-1, :3] < 1))
    assert np.all((s1[1:-1, 3:] >= -0.5) & (s1[1:-1, 3:] < 0.5))
    side = o1[:, 6:9] - o1[:, :3]
    centre = (o1[:, 6:9] + o1[:, :3]) / 2
    assert np.all((side > 0.0499) & (side < 0.2001))
    assert np.all((centre > 0.1999) & (centre < 0.8001))
    assert np.all(o1[:, 3:6] == np.float32(-3.1)) and np.all(o1[:, 9:12] == np.float32(3.1))
    # uniformity of the position draws
    m = s1[1:-1, :3].mean(0)
    assert np.all(np.abs(m - 0.5) < 0.02)


def test_walls_generator():
    cfg = configs.config(3, num=1000)
    s, o = P.gen_problem(cfg["spec"])
    side = np.sort(o[:, 6:9] - o[:, :3], axis=1)
    assert np.allclose(side[:, 0], 0.02, atol=1e-6)
    assert np.all((side[:, 1] >= 0.05 - 1e-6) & (side[:, 2] <= 0.8 + 1e-6))


def test_noise_tables_layout_and_covariance():
    ns = P.noise_spec(n_mc=4096, tmax_steps=120)
    ot, xc, sig = P.make_noise(ns, sigma=True)
    assert np.array_equal(ot.transpose(1, 0, 2), xc)
    assert np.all(ot[0] == 0)  # particles start on the nominal path
    # sample covariance of the offsets matches the FP32-propagated Sigma_t
    for t in (5, 20, 60, 119):
        emp = np.cov(ot[t].astype(np.float64).T)
        ref = sig[t].astype(np.float64)
        se = np.sqrt((ref ** 2 + np.outer(np.diag(ref), np.diag(ref))) / ns.n_mc)
        assert np.all(np.abs(emp - ref) <= 5 * se + 1e-9), t
    sd = np.sqrt(sig[119][0, 0])
    assert 0.03 < sd < 0.08
    # independent particles: lag-0 correlation across particles ~ 0
    c = np.corrcoef(ot[60, :-1, 0], ot[60, 1:, 0])[0, 1]
    assert abs(c) < 0.06


def test_host_bvp_arithmetic_matches_oracle():
    """The library's 2pBVP arithmetic (used on the host by findOptimalPath and,
    identically, by the kernels) agrees with the oracle's glibc-pow
    restatement: bit-exact except for rare last-place double differences."""
    import oracle
    O = oracle.load()
    rng = np.random.default_rng(5)
    n, diff_t, diff_c = 20000, 0, 0
    for _ in range(n):
        a = np.concatenate([rng.uniform(0, 1, 3), rng.uniform(-0.5, 0.5, 3)]).astype(np.float32)
        b = np.concatenate([rng.uniform(0, 1, 3), rng.uniform(-0.5, 0.5, 3)]).astype(np.float32)
        t1 = P.lib.pump_host_topt(P.fp(a), P.fp(b), 5.0)
        t2 = O.po_topt_bisection(oracle.fp(a), oracle.fp(b), 5.0)
        diff_t += np.float32(t1).tobytes() != np.float32(t2).tobytes()
        c1 = P.lib.pump_host_cost(np.float32(t2), P.fp(a), P.fp(b))
        c2 = O.po_cost(np.float32(t2), oracle.fp(a), oracle.fp(b))
        diff_c += np.float32(c1).tobytes() != np.float32(c2).tobytes()
    assert diff_t <= n * 1e-3 and diff_c <= n * 1e-3, (diff_t, diff_c)


def test_host_find_optimal_path_matches_golden(golden_dir):
    g = np.load(os.path.join(golden_dir, "bvp_optimal_path.npz"))
    exact = 0
    for k in range(len(g["totals"])):
        nwp, plen = g["lens"][k]
        path, tot = P.find_optimal_path(g["chains"][k][:nwp], dt=0.05)
        assert path.shape[0] == plen
        exact += path.tobytes() == g["paths"][k][:plen].tobytes()
        assert np.allclose(path, g["paths"][k][:plen], rtol=1e-5, atol=1e-6)
    assert exact >= len(g["totals"]) - 2


def test_configs_map_reference_parameters():
    c0 = configs.config(0)
    assert c0["params"].w_fixed == 20 and c0["params"].n_p == 1000
    assert abs(c0["params"].offset_mult - 0.02) < 1e-9
    rn_idx = P.lib.pump_rn_index(4000, c0["params"].rn_perc)
    assert abs(rn_idx / 4000 - 50) < 0.01
    c1 = configs.config(1)
    assert c1["params"].w_fixed == 0 and c1["params"].w_cap == 64 and c1["params"].n_p == 2000
    assert P.lib.pump_maxd2(np.float32(0.5)) == np.float32(0.1 * np.float32(0.5) ** 2)
    b = configs.batch_config(16)
    assert all(1000 <= x["spec"].num <= 8000 for x in b)


def _path_point_ieee(t, tau, x0, x1):
    """The kernels' pathPoint expression tree (2pBVP.cu:117-131 with exact
    powers) evaluated by numpy with IEEE division; vectorised over rows."""
    f32, f64 = np.float32, np.float64
    t, tau = t.astype(f32), tau.astype(f32)
    td, taud = t.astype(f64), tau.astype(f64)
    s2 = td * td
    s3 = s2 * td
    t2 = taud * taud
    t3 = t2 * taud
    out = np.zeros((t.shape[0], 6), f32)
    for d in (0, 1):
        p0, v0, p1, v1 = x0[:, d], x0[:, d + 3], x1[:, d], x1[:, d + 3]
        lin = t * v0 + p0
        c3 = f32(2) * p0 - f32(2) * p1 + (v0 + v1) * tau
        c2 = f32(3) * p0 - f32(3) * p1 + (f32(2) * v0 + v1) * tau
        out[:, d] = (lin.astype(f64) + (s3 * c3.astype(f64)) / t3
                     - (s2 * c2.astype(f64)) / t2).astype(f32)
        e = f32(3) * p0 - f32(3) * p1 + v1 * tau
        f = -((f32(2) * t) * e)
        c = f32(2) * p0 - f32(2) * p1 + v1 * tau
        q = 3.0 * s2 - ((f32(4) * t) * tau).astype(f64) + t2
        out[:, d + 3] = (f.astype(f64) / t2 + ((3.0 * s2) * c.astype(f64)) / t3
                         + (v0.astype(f64) * q) / t2).astype(f32)
    z0, vz0, z1, vz1 = (x.astype(f64) for x in (x0[:, 2], x0[:, 5], x1[:, 2], x1[:, 5]))
    a = (t * x0[:, 5]).astype(f64) * t3
    b = z0 * t3
    c = (s2 * taud) * ((((-3.0 * z0) + 3.0 * z1) - (2.0 * vz0) * taud) - (1.0 * vz1) * taud)
    dd = s3 * ((2.0 * z0 - 2.0 * z1) + ((x0[:, 5] + x1[:, 5]) * tau).astype(f64))
    out[:, 2] = ((((a + b) + c) + dd) / t3).astype(f32)
    a = vz0 * t3
    b = (t * tau).astype(f64) * ((((-6.0 * z0) + 6.0 * z1) - (4.0 * vz0) * taud)
                                 - (2.0 * vz1) * taud)
    c = s2 * ((((6.0 * z0) - 6.0 * z1) + (3.0 * vz0) * taud) + (3.0 * vz1) * taud)
    out[:, 5] = (((a + b) + c) / t3).astype(f32)
    return out


def test_path_point_division_is_correctly_rounded():
    """pathPoint's FP64 divisions run as a reciprocal plus two exact FMA
    residual steps; every waypoint must equal the one computed with IEEE
    division, bit for bit, over taus spanning the bisection's range."""
    rng = np.random.default_rng(11)
    n = 200000
    x0 = np.concatenate([rng.uniform(0, 1, (n, 3)), rng.uniform(-0.5, 0.5, (n, 3))], 1)
    x1 = np.concatenate([rng.uniform(0, 1, (n, 3)), rng.uniform(-0.5, 0.5, (n, 3))], 1)
    x0, x1 = x0.astype(np.float32), x1.astype(np.float32)
    tau = np.exp(rng.uniform(np.log(0.01), np.log(5.0), n)).astype(np.float32)
    k = rng.integers(1, 63, n)
    t = ((tau / np.float32(63)) * k.astype(np.float32)).astype(np.float32)
    ref = _path_point_ieee(t, tau, x0, x1)
    got = np.zeros((n, 6), np.float32)
    o = np.zeros(6, np.float32)
    for i in range(n):
        P.lib.pump_host_path_point(t[i], tau[i], P.fp(x0[i]), P.fp(x1[i]), P.fp(o))
        got[i] = o
    bad = np.nonzero(np.any(got.view(np.uint32) != ref.view(np.uint32), axis=1))[0]
    assert bad.size == 0, (bad[:5], got[bad[:2]], ref[bad[:2]])


@pytest.mark.parametrize("scale", [1.0, 30.0])
def test_edge_bbox_contains_every_waypoint(scale):
    """The edge pass drops a box for a whole edge when it is separated from
    this bounding box, so every waypoint position of the edge (interior
    points at k * RN(tau / (W-1)), endpoints copied) must lie inside it:
    random edges at the workload's magnitudes and at 30x larger states."""
    rng = np.random.default_rng(int(scale))
    n = 40000
    lo_s = np.float32([0, 0, 0, -0.5, -0.5, -0.5]) * scale
    hi_s = np.float32([1, 1, 1, 0.5, 0.5, 0.5]) * scale
    x0 = rng.uniform(lo_s, hi_s, (n, 6)).astype(np.float32)
    x1 = rng.uniform(lo_s, hi_s, (n, 6)).astype(np.float32)
    tau = np.exp(rng.uniform(np.log(0.01), np.log(5.0), n)).astype(np.float32)
    W = rng.integers(2, 65, n)
    lo, hi = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32)
    lo_i, hi_i = np.zeros(3, np.float32), np.zeros(3, np.float32)
    for i in range(n):
        P.lib.pump_host_edge_bbox(P.fp(x0[i]), P.fp(x1[i]), tau[i], P.fp(lo_i), P.fp(hi_i))
        lo[i], hi[i] = lo_i, hi_i
    assert np.all(lo <= x0[:, :3]) and np.all(x0[:, :3] <= hi)
    assert np.all(lo <= x1[:, :3]) and np.all(x1[:, :3] <= hi)
    rows = np.repeat(np.arange(n), W - 2)
    k = np.concatenate([np.arange(1, w - 1) for w in W]).astype(np.float32)
    dt = (tau / (W - 1).astype(np.float32)).astype(np.float32)
    t = (dt[rows] * k).astype(np.float32)
    pts = _path_point_ieee(t, tau[rows], x0[rows], x1[rows])[:, :3]
    assert np.all(lo[rows] <= pts) and np.all(pts <= hi[rows])
    # the box is the control-point hull plus a 1e-5 relative margin, not more
    S = (np.abs(x0[:, :3]) + np.abs(x1[:, :3]) + (np.abs(x0[:, 3:]) + np.abs(x1[:, 3:]))
         * tau[:, None])
    assert np.all(hi - lo <= np.abs(x1[:, :3] - x0[:, :3]) + S + 3e-5 * (1 + S))


def test_fp64_division_matches_ieee():
    """The kernels' FP64 division (correctly rounded reciprocal, then two
    residual steps with exact FMA) equals IEEE a/b: random operands across
    the magnitudes pathPoint produces, divisors with all-ones significands,
    near-exact quotients and quotients close to a rounding midpoint."""
    rng = np.random.default_rng(5)
    n = 400000
    sig = rng.uniform(1, 2, n)
    b = sig * 2.0 ** rng.integers(-30, 10, n)
    a = rng.uniform(-1, 1, n) * 2.0 ** rng.integers(-40, 12, n)
    # divisors whose significand is all ones
    b[: n // 8] = np.ldexp(2.0 - 2.0 ** -52, rng.integers(-20, 8, n // 8))
    # exact and near-midpoint quotients: a = q*b (rounded), q = random double
    q = rng.uniform(1, 2, n) * 2.0 ** rng.integers(-20, 20, n)
    sl = slice(n // 8, n // 2)
    a[sl] = q[sl] * b[sl]
    half = slice(n // 2, 3 * n // 4)
    a[half] = (q[half] + np.spacing(q[half]) / 2) * b[half]
    a[3 * n // 4: 3 * n // 4 + 2] = [0.0, -0.0]
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    out = np.zeros(n)
    P.lib.pump_host_div_rn(a.ctypes.data, b.ctypes.data, out.ctypes.data, n)
    ref = a / b
    assert np.array_equal(out.view(np.uint64), ref.view(np.uint64))
    # the test set separates: the plain product a*RN(1/b) is not always exact
    r = 1.0 / b
    q0 = a * r
    assert not np.array_equal(q0, ref)


def test_fp64_division_by_a_float_square_matches_ieee():
    """The one-step division used for the tau^2 divisors of pathPoint equals
    IEEE a/b for b = tau*tau with tau a float: random quotients, quotients
    one bit above and below a rounding midpoint, exact quotients, taus with
    all-ones significands and the magnitudes of pathPoint's operands."""
    rng = np.random.default_rng(11)
    n = 2_000_000
    tau = (rng.uniform(1, 2, n) * 2.0 ** rng.integers(-8, 5, n)).astype(np.float32)
    tau[: n // 10] = np.ldexp(np.float32(2 - 2 ** -23), rng.integers(-8, 5, n // 10)).astype(
        np.float32)
    b = tau.astype(np.float64) ** 2
    q = rng.uniform(1, 2, n) * 2.0 ** rng.integers(-30, 12, n) * rng.choice([-1.0, 1.0], n)
    a = q * b
    sl = slice(n // 4, n // 2)          # a / b just off a midpoint of q's binade
    mid = q[sl] + np.spacing(q[sl]) / 2
    a[sl] = mid * b[sl]
    a[n // 2: 5 * n // 8] = np.nextafter(a[n // 2: 5 * n // 8], np.inf)
    ex = slice(5 * n // 8, 3 * n // 4)  # exact quotients with short significands, signed zeros
    a[ex] = np.round(q[ex] * 2.0 ** 20) * 2.0 ** -20 * b[ex]
    assert np.any(a[ex] == 0) and np.any(np.signbit(a[ex]) & (a[ex] == 0))
    a = np.ascontiguousarray(a)
    out = np.zeros(n)
    P.lib.pump_host_div_sq(a.ctypes.data, np.ascontiguousarray(tau).ctypes.data,
                           out.ctypes.data, n)
    ref = a / b
    assert np.array_equal(out.view(np.uint64), ref.view(np.uint64))
    assert not np.array_equal(a * (1.0 / b), ref)


def test_pair_row_division_by_magic_number():
    """The pair pass finds a pair's row i = idx / (num - 1) by a multiply-high
    with the divisor's magic number; it must equal integer division for
    every divisor of the workloads (1 .. 65535: num up to 65,536 keeps the
    pair count below 2^32) over random and boundary dividends: 0, 2^32 - 1,
    and the multiples of d and the values just below them."""
    rng = np.random.default_rng(5)
    ds = np.concatenate([np.arange(1, 1025), rng.integers(1025, 65536, 600),
                         [3999, 15999, 31999, 63999, 65535, 32768, 65534]]).astype(np.uint32)
    for d in ds.tolist():
        n = np.concatenate([rng.integers(0, 1 << 32, 4000, dtype=np.uint64),
                            [0, 1, d - 1, d, d + 1, (1 << 32) - 1],
                            (rng.integers(1, (1 << 32) // d, 200, dtype=np.uint64) * d),
                            (rng.integers(1, (1 << 32) // d, 200, dtype=np.uint64) * d - 1)])
        n = np.ascontiguousarray(n.astype(np.uint32))
        q = np.zeros_like(n)
        P.lib.pump_host_div_magic(n.ctypes.data, n.size, d, q.ctypes.data)
        assert np.array_equal(q, n // np.uint32(d)), d


def test_dcost_magnitude_threshold():
    """The pair pass tests |dcost| > 0.0001 (dcost = RN((ca - cb) / 1e-6f),
    compared as a double) as |ca - cb| >= kDcostBig without dividing.  The
    float quotient is monotone in the numerator, so the two tests agree for
    every float exactly when kDcostBig satisfies the test and its predecessor
    does not; checked there, over a window of 2^20 floats each side, over
    random floats of both signs, and at 0, infinities and NaN."""
    text = open(os.path.join(ROOT, "pump_amd", "csrc", "pump_device.hpp")).read()
    m = re.search(r"constexpr float kDcostBig = (0x[0-9a-fp.+-]+)f;", text)
    assert m
    big = np.float32(float.fromhex(m.group(1)))
    h = np.float32(0.000001)

    def ref(d):
        with np.errstate(invalid="ignore", over="ignore"):
            return np.abs((d / h).astype(np.float32)).astype(np.float64) > 0.0001

    bits = big.view(np.uint32)
    window = (np.arange(-(1 << 20), 1 << 20, dtype=np.int64) + int(bits)).astype(np.uint32)
    rng = np.random.default_rng(11)
    rand = rng.integers(0, 0x7F800000, 1 << 20, dtype=np.uint32)
    for u in (window, rand):
        d = u.view(np.float32)
        for x in (d, -d):
            assert np.array_equal(ref(x), np.abs(x) >= big)
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan], dtype=np.float32)
    assert np.array_equal(ref(special), np.abs(special) >= big)
    assert ref(np.array([big]))[0] and not ref(np.array([np.nextafter(big, np.float32(0))]))[0]


def test_workspace_sizing_for_marked_pairs():
    """pump_propagate_workspace_pairs: no smaller than the default workspace,
    equal to it up to the default capacity (4 per path + 65536 pairs), and
    16 bytes plus a chunk record's share per further pair (capped at 2^30)."""
    for n, T in ((4096, 2001), (1000000, 2001), (6000, 120)):
        base = P.lib.pump_propagate_workspace(n, T)
        cap = 4 * n + 65536
        assert P.lib.pump_propagate_workspace_pairs(n, T, 0) == base
        assert P.lib.pump_propagate_workspace_pairs(n, T, cap) == base
        more = P.lib.pump_propagate_workspace_pairs(n, T, cap + 160000)
        assert 16 * 160000 <= more - base <= 17 * 160000 + 1024
        huge = P.lib.pump_propagate_workspace_pairs(n, T, 1 << 40)
        assert huge == P.lib.pump_propagate_workspace_pairs(n, T, 1 << 30)
