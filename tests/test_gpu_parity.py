"""GPU parity: every entry point of include/pump_abi.h, called through the C
ABI on an MI355X, against the CPU oracle on the same seeded inputs.

Bar: bit-exact for flags, indices, neighbour sets, counts and the returned
path; arrival times / costs / waypoints bit-exact in practice, with a
tolerance of 1e-5 relative written into each assertion for the rare
last-place double rounding of pow() (see pump_device.hpp)."""
import ctypes

import numpy as np
import pytest

import oracle
import oracle.pipeline as OP
import pump_amd as P
from pump_amd import configs

pytestmark = pytest.mark.gpu

O = oracle.load()
REL = 1e-5


def dev(a):
    return P.DeviceArray(np.ascontiguousarray(a))


def ptr(t):
    return t.p if t is not None else None


def host(t):
    P.sync()
    return t.numpy()


def assert_close_f32(got, ref, rel=REL, max_mismatch_frac=1e-5):
    got = np.asarray(got, np.float32)
    ref = np.asarray(ref, np.float32)
    assert got.shape == ref.shape
    same = got.view(np.uint32) == ref.view(np.uint32)
    frac = 1 - same.mean() if same.size else 0.0
    assert frac <= max_mismatch_frac, f"{(~same).sum()} of {same.size} differ"
    if not same.all():
        d = np.abs(got[~same].astype(np.float64) - ref[~same])
        assert np.all(d <= rel * np.abs(ref[~same]) + 1e-30)


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    assert P.lib.pump_dev_count() >= 1, "no GPU visible to libpump_amd"
    print("gpu:", P.device_name(), flush=True)


# ------------------------------------------------------------ pipeline ----
CASES = [
    # (num, n_obs, w_fixed, offset_mult, n_p, k, seed)
    (192, 10, 20, 0.5, 256, 30, 0),
    (256, 12, 0, 1.0, 300, 25, 1),     # variable waypoints, ragged n_p
    (300, 0, 9, 0.5, 128, 30, 2),      # no obstacles
    (200, 40, 0, 0.7, 200, 40, 3),     # many boxes
]


def build_pair(case, **over):
    num, n_obs, w_fixed, om, n_p, k, seed = case
    cfg = configs.small(num=num, n_obs=n_obs, seed=seed, w_fixed=w_fixed, n_p=n_p,
                        offset_mult=om, k=k, n_mc=512, tmax_steps=1200, keep_pairs=1,
                        **over)
    s, o = P.gen_problem(cfg["spec"])
    ot, xc = P.make_noise(cfg["noise"])
    print("case", case, flush=True)
    rm = P.Roadmap(s, o, s[-1], ot, xc, cfg["params"])
    orc = OP.roadmap(s, o, s[-1], cfg["params"], nthreads=8)
    return cfg, s, o, ot, xc, rm, orc


@pytest.mark.parametrize("case", CASES)
def test_roadmap_parity(case):
    cfg, s, o, ot, xc, rm, orc = build_pair(case)
    c, t = rm.pairs()
    assert_close_f32(t, orc["topts"])
    assert_close_f32(c, orc["copts"])
    assert rm.info.rn_idx == orc["rn_idx"]
    assert np.float32(rm.info.rn) == orc["rn"]
    e = rm.edges()
    assert rm.info.n_edges == orc["n_edges"]
    assert np.array_equal(e["row_start"], orc["row_start"])
    assert np.array_equal(e["edge_src"], orc["edge_src"])
    assert np.array_equal(e["edge_dst"], orc["edge_dst"])
    assert_close_f32(e["copts_edge"], orc["copts_edge"])
    assert_close_f32(e["topts_edge"], orc["topts_edge"])
    assert_close_f32(e["h"], orc["h"])
    assert rm.info.max_nn == orc["max_nn"]
    f = rm.flags()
    assert np.array_equal(f["is_free_samples"], orc["is_free_samples"])
    assert np.array_equal(f["wp_offsets"], orc["wp_offsets"])
    assert np.array_equal(f["is_free_edges"], orc["is_free_edges"])
    assert_close_f32(rm.waypoints(), orc["disc"])
    hs = rm.halfspaces()
    assert np.array_equal(hs["count_stored"], orc["count_stored"])
    assert_close_f32(hs["bs"], orc["bs"])
    assert_close_f32(hs["as_"], orc["as_"])
    rm.close()


def test_nn_tables_parity():
    cfg, s, o, ot, xc, rm, orc = build_pair(CASES[0])
    nn = rm.nn_tables(dense=True)
    num, m = s.shape[0], rm.info.max_nn
    go = np.empty((num, m), np.int32)
    come = np.empty((num, m), np.int32)
    dist = np.empty((num, m), np.float32)
    O.po_nn_tables(oracle.lp(orc["row_start"]), oracle.ip(orc["edge_dst"]),
                   oracle.fp(orc["copts_edge"]), num, m, oracle.ip(go), oracle.ip(come),
                   oracle.fp(dist))
    assert np.array_equal(nn["nn_go_edges"], go)
    assert np.array_equal(nn["nn_come_edges"], come)
    assert np.array_equal(nn["distances_come"], dist)
    dense = np.full((num, num), -3, np.int32)
    dense[orc["edge_dst"], orc["edge_src"]] = np.arange(orc["n_edges"], dtype=np.int32)
    assert np.array_equal(nn["nn_idxs"], dense)
    rm.close()


@pytest.mark.parametrize("case", CASES[:2] + CASES[3:])
@pytest.mark.parametrize("bits", [0, 1])
@pytest.mark.parametrize("on_device", [1, 0])
def test_search_parity(case, bits, on_device):
    """The search with its bookkeeping on the device (default) and on the
    host, against the oracle's restatement of PUMP.cu:3-405."""
    cfg, s, o, ot, xc, rm, orc = build_pair(case, validity_bits=bits)
    prm = cfg["params"]
    prm.search_device = on_device
    got = P.search_result_dict(rm.search(prm))
    ref = OP.search(orc, s, s[-1], o, ot, xc, prm, nthreads=8, path_cap=prm.path_cap)
    print("gpu", got, "\noracle", ref, flush=True)
    for k in ("status", "itrs", "num_paths", "num_considered", "num_goal_paths", "soln_path_idx",
              "soln_nodes", "particle_steps", "clamped_time_idx"):
        assert got[k] == ref[k], k
    for k in ("soln_cost", "soln_cp", "soln_time", "soln_hsmc_cp"):
        assert np.float32(got[k]) == np.float32(ref[k]), k
    rm.close()


def test_search_repeated_on_one_roadmap():
    """The roadmap keeps the search's device buffers from one search to the
    next. Searches in a row on one roadmap (validity bits, then bytes with
    their wider rows, the host bookkeeping, a smaller particle count, and
    smoothing between them) each return the oracle's result for their own
    parameters, as a search on a fresh roadmap does."""
    import ctypes
    cfg, s, o, ot, xc, rm, orc = build_pair(CASES[1])
    base = cfg["params"]

    def params(**kw):
        q = P.Params()
        ctypes.memmove(ctypes.byref(q), ctypes.byref(base), ctypes.sizeof(P.Params))
        for k, v in kw.items():
            setattr(q, k, v)
        return q

    refs = {}
    for step, kw in enumerate([dict(validity_bits=1, search_device=1),
                               dict(validity_bits=0, search_device=1),
                               dict(validity_bits=1, search_device=0),
                               dict(validity_bits=1, search_device=1, n_p=97),
                               dict(validity_bits=0, search_device=1, n_p=97),
                               dict(validity_bits=1, search_device=1)]):
        q = params(**kw)
        r = rm.search(q)
        got = P.search_result_dict(r)
        if q.n_p not in refs:
            refs[q.n_p] = OP.search(orc, s, s[-1], o, ot, xc, q, nthreads=8,
                                    path_cap=q.path_cap)
        ref = refs[q.n_p]
        print("step", step, kw, got["status"], got["num_paths"], flush=True)
        for k in ("status", "itrs", "num_paths", "num_considered", "num_goal_paths",
                  "soln_path_idx", "soln_nodes", "particle_steps", "clamped_time_idx"):
            assert got[k] == ref[k], (step, k)
        for k in ("soln_cost", "soln_cp", "soln_time", "soln_hsmc_cp"):
            assert np.float32(got[k]) == np.float32(ref[k]), (step, k)
        if r.status == 0 and step in (0, 3):
            rm.smooth(r, q)
    assert len(refs) == 2
    rm.close()


# ------------------------------------------- entry points called directly ----
def random_hsmc_inputs(seed, n_edges=60, W=7, max_hs=4, n_p=333, n_mc=512, T=80, n_new=50,
                       n_parent=9, variable=False, w_range=(2, 40)):
    rng = np.random.default_rng(seed)
    if variable:
        w = rng.integers(w_range[0], w_range[1], n_edges)
    else:
        w = np.full(n_edges, W)
    off = np.zeros(n_edges + 1, np.int64)
    np.cumsum(w, out=off[1:])
    n_wp = int(off[-1])
    hs = n_wp * max_hs
    as_ = rng.normal(0, 0.05, (hs, 3)).astype(np.float32)
    bs = (rng.normal(0, 0.05, hs) ** 2 * 0.5).astype(np.float32)
    bs[rng.uniform(size=hs) < 0.6] = 1111
    bs[rng.uniform(size=hs) < 0.02] = 0.0
    offs = rng.normal(0, 0.05, (T, n_mc, 3)).astype(np.float32)
    topts = rng.uniform(0.2, 2.0, n_edges).astype(np.float32)
    valid = np.zeros((n_parent + n_new, n_p), np.uint8)
    valid[:n_parent] = rng.uniform(size=(n_parent, n_p)) < 0.85
    times = np.zeros(n_parent + n_new, np.float32)
    times[:n_parent] = rng.uniform(0, 2.0, n_parent)
    times[n_parent - 1] = 3.9  # pushes time indices past the table end
    wf_edge = rng.integers(0, n_edges, n_new).astype(np.int32)
    wf_prev = rng.integers(0, n_parent, n_new).astype(np.int32)
    return dict(off=off, w=w, as_=as_, bs=bs, offs=offs, topts=topts, valid=valid, times=times,
                wf_edge=wf_edge, wf_prev=wf_prev, max_hs=max_hs, n_p=n_p, n_mc=n_mc, T=T,
                n_new=n_new, n_parent=n_parent, om=np.float32(0.8), dt=np.float32(0.05))


def edge_mask_host(off, bs, max_hs):
    """Per-edge waypoint summary: bit min(k, 63) set iff waypoint k of the
    edge holds a slot whose b is not >= 1000 (the sentinel test of the
    propagate step)."""
    has = ~(bs.reshape(-1, max_hs) >= 1000).all(axis=1)
    n_edges = off.shape[0] - 1
    edge_of = np.repeat(np.arange(n_edges), np.diff(off))
    k = np.arange(off[-1]) - off[edge_of]
    bits = np.where(has, np.left_shift(np.uint64(1), np.minimum(k, 63).astype(np.uint64)),
                    np.uint64(0)).astype(np.uint64)
    out = np.zeros(n_edges, np.uint64)
    np.bitwise_or.at(out, edge_of, bits)
    return out


def oracle_propagate(d):
    v = d["valid"].copy()
    cl = O.po_propagate_valid_particles(
        d["n_new"], oracle.lp(d["off"]), oracle.ip(d["wf_edge"]), oracle.ip(d["wf_prev"]),
        d["n_parent"], oracle.fp(d["as_"]), oracle.fp(d["bs"]), oracle.fp(d["times"]),
        oracle.fp(d["topts"]), d["dt"], d["max_hs"], oracle.fp(d["offs"]), d["T"], d["om"],
        d["n_mc"], oracle.bp(v), d["n_p"], 1)
    cp = np.ones(d["n_parent"] + d["n_new"], np.float32)
    O.po_calculate_cp_hsmc(d["n_new"], d["n_parent"], oracle.fp(cp), oracle.bp(v), d["n_p"])
    return v, cp, cl


@pytest.mark.parametrize("variable,n_new", [(False, 50), (True, 50), (True, 6000)])
def test_propagate_entry_points(variable, n_new):
    # 6000 new paths: the drop-in runs the batched form in place
    d = random_hsmc_inputs(7 + variable, variable=variable, n_new=n_new)
    ref_v, ref_cp, ref_cl = oracle_propagate(d)
    assert ref_cl > 0
    g = {k: dev(d[k]) for k in ("off", "as_", "bs", "offs", "topts", "times", "wf_edge",
                                 "wf_prev")}
    W = int(d["w"][0])
    off_ptr = ptr(g["off"]) if variable else None
    # reference-shaped pair: in-place propagate, then calculateCP_HSMC
    valid = dev(d["valid"])
    cp = dev(np.ones(d["n_parent"] + d["n_new"], np.float32))
    clamp = dev(np.zeros(2, np.uint64))
    P.check("propagate", P.lib.pump_propagate_valid_particles(
        d["n_new"], W, off_ptr, ptr(g["wf_edge"]), ptr(g["wf_prev"]), d["n_parent"],
        ptr(g["as_"]), ptr(g["bs"]), None, ptr(g["times"]), ptr(g["topts"]), d["dt"],
        d["max_hs"], ptr(g["offs"]), d["T"], d["om"], d["n_mc"], ptr(valid), d["n_p"],
        ptr(clamp), None))
    P.check("cp", P.lib.pump_calculate_cp_hsmc(d["n_new"], d["n_parent"], ptr(cp), ptr(valid),
                                               d["n_p"], None))
    assert np.array_equal(host(valid), ref_v)
    assert np.array_equal(host(cp)[d["n_parent"]:], ref_cp[d["n_parent"]:])
    assert int(host(clamp)[0]) == ref_cl
    # fused propagate + CP with separate parent / child arrays
    src = dev(d["valid"][: d["n_parent"]])
    dst = dev(np.zeros((d["n_new"], d["n_p"]), np.uint8))
    cp2 = dev(np.zeros(d["n_new"], np.float32))
    P.check("propagate_cp", P.lib.pump_propagate_cp(
        d["n_new"], W, off_ptr, ptr(g["wf_edge"]), ptr(g["wf_prev"]), ptr(g["as_"]),
        ptr(g["bs"]), None, ptr(g["times"]), ptr(g["topts"]), d["dt"], d["max_hs"],
        ptr(g["offs"]), d["T"], d["om"], d["n_mc"], ptr(src), ptr(dst), 0, d["n_p"], ptr(cp2), 0,
        None, None, 0, None))
    assert np.array_equal(host(dst), ref_v[d["n_parent"]:])
    assert np.array_equal(host(cp2), ref_cp[d["n_parent"]:])
    # bit masks
    nw = (d["n_p"] + 63) // 64
    bits_src = np.zeros((d["n_parent"], nw), np.uint64)
    for r in range(d["n_parent"]):
        packed = np.packbits(d["valid"][r].astype(np.uint8), bitorder="little")
        packed = np.pad(packed, (0, nw * 8 - packed.size))
        bits_src[r] = packed.view(np.uint64)
    bsrc = dev(bits_src)
    bdst = dev(np.zeros((d["n_new"], nw), np.uint64))
    cp3 = dev(np.zeros(d["n_new"], np.float32))
    P.check("propagate_cp_bits", P.lib.pump_propagate_cp_bits(
        d["n_new"], W, off_ptr, ptr(g["wf_edge"]), ptr(g["wf_prev"]), ptr(g["as_"]),
        ptr(g["bs"]), None, ptr(g["times"]), ptr(g["topts"]), d["dt"], d["max_hs"],
        ptr(g["offs"]), d["T"], d["om"], d["n_mc"], ptr(bsrc), ptr(bdst), 0, d["n_p"], ptr(cp3),
        0, 0, None, None, None, None, 0, None))
    out = host(bdst)
    for r in range(d["n_new"]):
        unpacked = np.unpackbits(out[r].view(np.uint8), bitorder="little")[: d["n_p"]]
        assert np.array_equal(unpacked, ref_v[d["n_parent"] + r])
    assert np.array_equal(host(cp3), ref_cp[d["n_parent"]:])


def split_table(g_offs, T, n_mc):
    """The split-BF16 particle-vector table of a device offsets table."""
    nbytes = P.lib.pump_split_offsets_bytes(T, n_mc)
    t = P.DeviceArray(shape=(nbytes,), dtype=np.uint8, fill=0)
    P.check("split_offsets", P.lib.pump_split_offsets(ptr(g_offs), T, n_mc, t.p, None))
    return t


def pack_bits(rows, n_p):
    nw = (n_p + 63) // 64
    out = np.zeros((rows.shape[0], nw), np.uint64)
    for r in range(rows.shape[0]):
        packed = np.packbits(rows[r].astype(np.uint8), bitorder="little")
        packed = np.pad(packed, (0, nw * 8 - packed.size))
        out[r] = packed.view(np.uint64)
    return out


def unpack_bits(words, n_p):
    return np.stack([np.unpackbits(w.view(np.uint8), bitorder="little")[:n_p] for w in words])


@pytest.mark.parametrize("n_p,w_range,T", [(333, (2, 40), 120), (2000, (2, 40), 120),
                                            (512, (2, 40), 3000), (640, (40, 300), 120)])
def test_propagate_batched_form(n_p, w_range, T):
    """Large wavefronts (>= 4096 paths) with a workspace take the batched
    form: classify, copy, then the half-space tests binned by time index, or,
    when the binned pairs overflow the workspace (the 40-300 waypoint case),
    one wave per (path, particle group).  Results must equal the oracle, with
    and without the per-edge waypoint summary (edges past 64 waypoints
    exercise the summary's inexact case and waypoint groups past 256)."""
    d = random_hsmc_inputs(21 + n_p, n_edges=3000, W=9, max_hs=5, n_p=n_p, n_mc=2048, T=T,
                           n_new=6000, n_parent=64, variable=True, w_range=w_range)
    # most edges hold no half-space, as in the real workloads
    keep = np.random.default_rng(1).uniform(size=d["off"].shape[0] - 1) < 0.15
    for e in np.nonzero(~keep)[0]:
        d["bs"][d["off"][e] * d["max_hs"]: d["off"][e + 1] * d["max_hs"]] = 1111
    ref_v, ref_cp, ref_cl = oracle_propagate(d)
    g = {k: dev(d[k]) for k in ("off", "as_", "bs", "offs", "topts", "times", "wf_edge",
                                 "wf_prev")}
    mask = edge_mask_host(d["off"], d["bs"], d["max_hs"])
    assert (mask >> np.uint64(63)).any() == (w_range[1] > 64)
    g_mask = dev(mask)
    g_split = split_table(g["offs"], d["T"], d["n_mc"])
    ws_bytes = P.lib.pump_propagate_workspace(d["n_new"], d["T"])
    ws = P.DeviceArray(shape=(ws_bytes,), dtype=np.uint8, fill=0)
    npar, nn = d["n_parent"], d["n_new"]
    for m in (None, ptr(g_mask)):
        src = dev(d["valid"][:npar])
        dst = dev(np.zeros((nn, n_p), np.uint8))
        cp = dev(np.zeros(nn, np.float32))
        clamp = dev(np.zeros(2, np.uint64))
        P.check("propagate_cp", P.lib.pump_propagate_cp(
            nn, 0, ptr(g["off"]), ptr(g["wf_edge"]), ptr(g["wf_prev"]), ptr(g["as_"]),
            ptr(g["bs"]), m, ptr(g["times"]), ptr(g["topts"]), d["dt"], d["max_hs"],
            ptr(g["offs"]), d["T"], d["om"], d["n_mc"], ptr(src), ptr(dst), 0, n_p, ptr(cp), 0,
            ptr(clamp), ws.p, ws_bytes, None))
        assert np.array_equal(host(dst), ref_v[npar:])
        assert np.array_equal(host(cp), ref_cp[npar:])
        assert int(host(clamp)[0]) == ref_cl
        # control words: listed paths, binned pairs, overflow flag
        ctl = ws.numpy()[:16].view(np.int32)
        assert ctl[0] > 0 and ctl[2] == (1 if w_range[1] > 64 else 0), ctl[:4]
        for precision, split in ((0, None), (1, g_split)):
            bsrc = dev(pack_bits(d["valid"][:npar], n_p))
            bdst = dev(np.zeros((nn, (n_p + 63) // 64), np.uint64))
            cpb = dev(np.zeros(nn, np.float32))
            clamp_b = dev(np.zeros(2, np.uint64))
            P.check("propagate_cp_bits", P.lib.pump_propagate_cp_bits(
                nn, 0, ptr(g["off"]), ptr(g["wf_edge"]), ptr(g["wf_prev"]), ptr(g["as_"]),
                ptr(g["bs"]), m, ptr(g["times"]), ptr(g["topts"]), d["dt"], d["max_hs"],
                ptr(g["offs"]), d["T"], d["om"], d["n_mc"], ptr(bsrc), ptr(bdst), 0, n_p,
                ptr(cpb), 0, precision, split.p if split else None, clamp_b.ptr + 8,
                ptr(clamp_b), ws.p, ws_bytes, None))
            assert np.array_equal(unpack_bits(host(bdst), n_p), ref_v[npar:]), precision
            assert np.array_equal(host(cpb), ref_cp[npar:]), precision
            assert int(host(clamp_b)[0]) == ref_cl
        # the single-launch form takes the summary too
        dst1 = dev(np.zeros((nn, n_p), np.uint8))
        cp1 = dev(np.zeros(nn, np.float32))
        P.check("propagate_cp", P.lib.pump_propagate_cp(
            nn, 0, ptr(g["off"]), ptr(g["wf_edge"]), ptr(g["wf_prev"]), ptr(g["as_"]),
            ptr(g["bs"]), m, ptr(g["times"]), ptr(g["topts"]), d["dt"], d["max_hs"],
            ptr(g["offs"]), d["T"], d["om"], d["n_mc"], ptr(src), ptr(dst1), 0, n_p, ptr(cp1), 0,
            None, None, 0, None))
        assert np.array_equal(host(dst1), ref_v[npar:])
        assert np.array_equal(host(cp1), ref_cp[npar:])
    if w_range[1] > 64:
        # a workspace sized for every (path, waypoint) pair of the wavefront
        # keeps the binned form where the default one overflowed
        n_pairs = int((d["off"][d["wf_edge"] + 1] - d["off"][d["wf_edge"]]).sum())
        big = int(P.lib.pump_propagate_workspace_pairs(nn, d["T"], n_pairs))
        assert big > ws_bytes
        wsb = P.DeviceArray(shape=(big,), dtype=np.uint8, fill=0)
        for precision, split in ((0, None), (1, g_split)):
            bsrc = dev(pack_bits(d["valid"][:npar], n_p))
            bdst = dev(np.zeros((nn, (n_p + 63) // 64), np.uint64))
            cpb = dev(np.zeros(nn, np.float32))
            P.check("propagate_cp_bits", P.lib.pump_propagate_cp_bits(
                nn, 0, ptr(g["off"]), ptr(g["wf_edge"]), ptr(g["wf_prev"]), ptr(g["as_"]),
                ptr(g["bs"]), ptr(g_mask), ptr(g["times"]), ptr(g["topts"]), d["dt"],
                d["max_hs"], ptr(g["offs"]), d["T"], d["om"], d["n_mc"], ptr(bsrc), ptr(bdst),
                0, n_p, ptr(cpb), 0, precision, split.p if split else None, None, None, wsb.p,
                big, None))
            assert np.array_equal(unpack_bits(host(bdst), n_p), ref_v[npar:]), precision
            assert np.array_equal(host(cpb), ref_cp[npar:]), precision
            ctl = wsb.numpy()[:16].view(np.int32)
            assert ctl[2] == 0 and ctl[1] > 4 * nn + 65536, ctl[:4]


@pytest.mark.parametrize("n_new", [6000, 500])
def test_propagate_copy_on_write(n_new):
    """Copy-on-write step (pump_propagate_cp_cow): children that meet no
    stored half-space point at their parent's row and take its CP; the rows
    gathered for every child (pump_gather_rows_cow) equal the oracle's, in byte
    and bit validity, FP32 and split-BF16; a small wavefront (per-path form)
    writes every row."""
    n_p = 700
    d = random_hsmc_inputs(91, n_edges=3000, W=9, max_hs=5, n_p=n_p, n_mc=1024, T=300,
                           n_new=n_new, n_parent=64)
    keep = np.random.default_rng(3).uniform(size=d["off"].shape[0] - 1) < 0.15
    for e in np.nonzero(~keep)[0]:
        d["bs"][d["off"][e] * d["max_hs"]: d["off"][e + 1] * d["max_hs"]] = 1111
    ref_v, ref_cp, ref_cl = oracle_propagate(d)
    npar = d["n_parent"]
    pcp = (np.count_nonzero(d["valid"][:npar] == 0, axis=1).astype(np.float32) /
           np.float32(n_p)).astype(np.float32)
    g = {k: dev(d[k]) for k in ("off", "as_", "bs", "offs", "topts", "times", "wf_edge",
                                 "wf_prev")}
    g_mask = dev(edge_mask_host(d["off"], d["bs"], d["max_hs"]))
    g_split = split_table(g["offs"], d["T"], d["n_mc"])
    g_pcp = dev(pcp)
    ws_bytes = P.lib.pump_propagate_workspace(n_new, d["T"])
    ws = P.DeviceArray(shape=(ws_bytes,), dtype=np.uint8, fill=0)
    has_hs = (edge_mask_host(d["off"], d["bs"], d["max_hs"]) != 0)[d["wf_edge"]]
    for bits, precision in ((0, 0), (1, 0), (1, 1)):
        width = (n_p + 63) // 64 * 8 if bits else n_p
        src = dev(pack_bits(d["valid"][:npar], n_p) if bits else d["valid"][:npar])
        stage = dev(np.zeros((n_new, width), np.uint8))
        row_src = dev(np.full(n_new, 12345, np.int32))
        cp = dev(np.zeros(n_new, np.float32))
        cnt = dev(np.zeros(2, np.uint64))
        P.check("propagate_cp_cow", P.lib.pump_propagate_cp_cow(
            bits, n_new, 0, ptr(g["off"]), ptr(g["wf_edge"]), ptr(g["wf_prev"]), ptr(g["as_"]),
            ptr(g["bs"]), ptr(g_mask), ptr(g["times"]), ptr(g["topts"]), d["dt"], d["max_hs"],
            ptr(g["offs"]), d["T"], d["om"], d["n_mc"], ptr(src), ptr(g_pcp), ptr(stage), 0,
            ptr(row_src), n_p, ptr(cp), 0, precision, g_split.p if precision else None,
            cnt.ptr + 8, ptr(cnt), ws.p, ws_bytes, None))
        rs = host(row_src)
        if n_new >= 4096:
            assert np.array_equal(rs[~has_hs], -1 - d["wf_prev"][~has_hs])
            assert np.array_equal(rs[has_hs], np.arange(n_new)[has_hs])
        else:
            assert np.array_equal(rs, np.arange(n_new))
        assert np.array_equal(host(cp), ref_cp[npar:]), (bits, precision)
        assert int(host(cnt)[0]) == ref_cl
        out = dev(np.zeros((n_new + 3, width), np.uint8))
        P.check("gather_rows_cow", P.lib.pump_gather_rows_cow(
            ptr(stage), ptr(src), ptr(row_src), None, n_new, ptr(out), 3, width, bits, None))
        rows = host(out)[3:]
        got = unpack_bits(rows.view(np.uint64), n_p) if bits else rows
        assert np.array_equal(got, ref_v[npar:]), (bits, precision)
        # a selection: every third child, in reverse
        sel = np.arange(n_new)[::-3].astype(np.int32).copy()
        out2 = dev(np.zeros((sel.size, width), np.uint8))
        g_sel = dev(sel)
        P.check("gather_rows_cow", P.lib.pump_gather_rows_cow(
            ptr(stage), ptr(src), ptr(row_src), ptr(g_sel), sel.size, ptr(out2), 0, width, bits,
            None))
        rows2 = host(out2)
        got2 = unpack_bits(rows2.view(np.uint64), n_p) if bits else rows2
        assert np.array_equal(got2, ref_v[npar:][sel])
        if n_new < 4096:
            continue
        # every particles-per-wave shape of the binned test gives the same
        # rows and CPs; then the autotuner picks one of them
        for ppw in (64, 128, 512, 0):
            P.check("set_ppw", P.lib.pump_hsmc_set_particles_per_wave(
                bits, precision, n_p, d["max_hs"], ppw))
            assert P.lib.pump_hsmc_particles_per_wave(bits, precision, n_p, d["max_hs"]) == (
                ppw or 256)
            stage2 = dev(np.zeros((n_new, width), np.uint8))
            cp2 = dev(np.zeros(n_new, np.float32))
            P.check("propagate_cp_cow", P.lib.pump_propagate_cp_cow(
                bits, n_new, 0, ptr(g["off"]), ptr(g["wf_edge"]), ptr(g["wf_prev"]),
                ptr(g["as_"]), ptr(g["bs"]), ptr(g_mask), ptr(g["times"]), ptr(g["topts"]),
                d["dt"], d["max_hs"], ptr(g["offs"]), d["T"], d["om"], d["n_mc"], ptr(src),
                ptr(g_pcp), ptr(stage2), 0, ptr(row_src), n_p, ptr(cp2), 0, precision,
                g_split.p if precision else None, None, None, ws.p, ws_bytes, None))
            assert np.array_equal(host(cp2), ref_cp[npar:]), (bits, precision, ppw)
            assert np.array_equal(host(stage2)[has_hs], host(stage)[has_hs]), (bits, ppw)
        stage3 = dev(np.zeros((n_new, width), np.uint8))
        cp3 = dev(np.zeros(n_new, np.float32))
        cnt3 = dev(np.zeros(2, np.uint64))
        chosen = ctypes.c_int(0)
        P.check("hsmc_autotune_cow", P.lib.pump_hsmc_autotune_cow(
            bits, n_new, 0, ptr(g["off"]), ptr(g["wf_edge"]), ptr(g["wf_prev"]), ptr(g["as_"]),
            ptr(g["bs"]), ptr(g_mask), ptr(g["times"]), ptr(g["topts"]), d["dt"], d["max_hs"],
            ptr(g["offs"]), d["T"], d["om"], d["n_mc"], ptr(src), ptr(g_pcp), ptr(stage3), 0,
            ptr(row_src), n_p, ptr(cp3), 0, precision, g_split.p if precision else None,
            cnt3.ptr + 8, ptr(cnt3), ws.p, ws_bytes, None, ctypes.byref(chosen)))
        assert chosen.value in (64, 128, 256, 512)
        assert P.lib.pump_hsmc_particles_per_wave(bits, precision, n_p, d["max_hs"]) == chosen.value
        assert np.array_equal(host(cp3), ref_cp[npar:])
        assert np.array_equal(host(stage3)[has_hs], host(stage)[has_hs])
        # the counters come from the final call only
        assert int(host(cnt3)[0]) == ref_cl and host(cnt3)[1] == host(cnt)[1]
        P.check("set_ppw", P.lib.pump_hsmc_set_particles_per_wave(
            bits, precision, n_p, d["max_hs"], 0))


def test_split_bf16_band_decisions():
    """Split-BF16 MFMA tests with half-spaces placed exactly on a particle's
    decision boundary (b = RN(LHS*offsetMult) in the reference's FP32 order,
    and the next float above): those tests fall in the margin band, are redone
    in FP32 and counted, and every row and CP equals the oracle's."""
    d = random_hsmc_inputs(77, n_edges=3000, W=9, max_hs=5, n_p=700, n_mc=1024, T=40,
                           n_new=6000, n_parent=64)
    keep = np.random.default_rng(2).uniform(size=d["off"].shape[0] - 1) < 0.2
    hs_rows = d["bs"].reshape(-1, d["max_hs"])
    for e in np.nonzero(~keep)[0]:
        hs_rows[d["off"][e]:d["off"][e + 1]] = 1111
    # every waypoint of every path at time index 0
    d["times"][:] = 0
    d["topts"][:] = np.float32(1e-6)
    f32 = np.float32
    om = f32(d["om"])
    rng = np.random.default_rng(3)
    live = np.nonzero(d["bs"] < 1000)[0]
    pick = live[rng.uniform(size=live.size) < 0.5]
    part = rng.integers(0, d["n_p"], pick.size)
    o = d["offs"][0, part]
    a = d["as_"][pick]
    lhs = f32(0) + o[:, 0] * a[:, 0]
    lhs = lhs + o[:, 1] * a[:, 1]
    lhs = lhs + o[:, 2] * a[:, 2]
    edge = (lhs * om).astype(np.float32)
    up = rng.uniform(size=pick.size) < 0.5
    d["bs"][pick] = np.where(up, np.nextafter(edge, np.float32(np.inf)), edge)
    ref_v, ref_cp, _ = oracle_propagate(d)
    g = {k: dev(d[k]) for k in ("off", "as_", "bs", "offs", "topts", "times", "wf_edge",
                                 "wf_prev")}
    W = int(d["w"][0])
    mask = edge_mask_host(d["off"], d["bs"], d["max_hs"])
    g_mask = dev(mask)
    g_split = split_table(g["offs"], d["T"], d["n_mc"])
    ws_bytes = P.lib.pump_propagate_workspace(d["n_new"], d["T"])
    ws = P.DeviceArray(shape=(ws_bytes,), dtype=np.uint8, fill=0)
    npar, nn, n_p = d["n_parent"], d["n_new"], d["n_p"]
    bsrc = dev(pack_bits(d["valid"][:npar], n_p))
    bdst = dev(np.zeros((nn, (n_p + 63) // 64), np.uint64))
    cpb = dev(np.zeros(nn, np.float32))
    cnt = dev(np.zeros(2, np.uint64))
    P.check("propagate_cp_bits", P.lib.pump_propagate_cp_bits(
        nn, W, None, ptr(g["wf_edge"]), ptr(g["wf_prev"]), ptr(g["as_"]), ptr(g["bs"]),
        ptr(g_mask), ptr(g["times"]), ptr(g["topts"]), d["dt"], d["max_hs"], ptr(g["offs"]),
        d["T"], d["om"], d["n_mc"], ptr(bsrc), ptr(bdst), 0, n_p, ptr(cpb), 0, 1, g_split.p,
        cnt.ptr + 8, ptr(cnt), ws.p, ws_bytes, None))
    assert np.array_equal(unpack_bits(host(bdst), n_p), ref_v[npar:])
    assert np.array_equal(host(cpb), ref_cp[npar:])
    ctl = ws.numpy()[:16].view(np.int32)
    assert ctl[2] == 0  # the binned (MFMA) pass ran
    assert int(host(cnt)[1]) > 0


def test_gather_rows():
    rng = np.random.default_rng(3)
    src = rng.integers(0, 255, (40, 100), dtype=np.uint8)
    rows = np.array([3, 7, 7, 0, 39], np.int32)
    dst = np.zeros((12, 100), np.uint8)
    gs, gr, gd = dev(src), dev(rows), dev(dst)
    P.check("gather", P.lib.pump_gather_rows(ptr(gs), ptr(gr), 5, ptr(gd), 4, 100, None))
    out = host(gd)
    assert np.array_equal(out[4:9], src[rows]) and not out[:4].any() and not out[9:].any()


def test_collision_probability_entry_points():
    rng = np.random.default_rng(11)
    n_obs = 15
    obs = np.zeros((n_obs, 12), np.float32)
    c = rng.uniform(0.2, 0.8, (n_obs, 3))
    sd = rng.uniform(0.05, 0.2, (n_obs, 3))
    obs[:, :3], obs[:, 6:9] = c - sd / 2, c + sd / 2
    obs[:, 3:6], obs[:, 9:12] = -3.1, 3.1
    n_mc, tsteps = 700, 300
    xc = rng.normal(0, 0.05, (n_mc, tsteps, 3)).astype(np.float32)
    n_paths, cap = 6, 260
    paths = np.zeros((n_paths, cap, 6), np.float32)
    Ts = rng.integers(3, cap, n_paths).astype(np.int32)
    Ts[0] = 2  # no segment at all
    Ts[1] = 130
    for k in range(n_paths):
        a, b = rng.uniform(0, 1, 3), rng.uniform(0, 1, 3)
        lam = np.linspace(0, 1, cap)[:, None]
        paths[k, :, :3] = a + lam * (b - a)
    om = np.float32(1.0)
    ref = [O.po_mccp_hits(oracle.fp(obs), n_obs, oracle.fp(xc), tsteps, n_mc, om,
                          oracle.fp(np.ascontiguousarray(paths[k])), int(Ts[k]), None)
           for k in range(n_paths)]
    assert max(ref) > 0 and min(ref) == 0
    gobs, gxc, gp, gT = dev(obs), dev(xc), dev(paths), dev(Ts)
    hits = dev(np.full(n_paths, 77, np.int32))
    P.check("mccp_many", P.lib.pump_collision_probability_many(
        ptr(gobs), n_obs, ptr(gxc), tsteps, n_mc, om, ptr(gp), cap, ptr(gT), n_paths, ptr(hits), 1,
        None))
    assert host(hits).tolist() == ref
    cp = ctypes.c_float(0)
    gp1 = dev(paths[1])
    P.check("mccp", P.lib.pump_collision_probability(ptr(gobs), n_obs, ptr(gxc), tsteps, n_mc, om,
                                                     ptr(gp1), int(Ts[1]), ctypes.byref(cp), None))
    assert np.float32(cp.value) == np.float32(np.float32(ref[1]) / np.float32(n_mc))


def fixed_rows(samples, orc, W):
    """W waypoints per edge (numDisc = W - 1), from the oracle's
    findDiscretizedPath (2pBVP.cu:53-69) of every edge."""
    E = orc["n_edges"]
    off = np.arange(0, (E + 1) * W, W, dtype=np.int64)
    rows = np.zeros((E * W, 6), np.float32)
    O.po_discretize_edges(oracle.fp(samples), oracle.lp(orc["row_start"]),
                          oracle.ip(orc["edge_dst"]), samples.shape[0], oracle.lp(off),
                          oracle.fp(rows), 4)
    return rows


def test_free_edges_and_discretize_entry_points():
    cfg, s, o, ot, xc, rm, orc = build_pair(CASES[1])
    E, n_wp = orc["n_edges"], orc["n_wp"]
    gs, go = dev(s), dev(o)
    gsrc, gdst = dev(orc["edge_src"]), dev(orc["edge_dst"])
    gte, goff = dev(orc["topts_edge"]), dev(orc["wp_offsets"])
    disc = dev(np.zeros((n_wp, 6), np.float32))
    P.check("discretize", P.lib.pump_discretize_edges(ptr(gs), ptr(gsrc), ptr(gdst), ptr(gte), E,
                                                      ptr(goff), 0, ptr(disc), None))
    assert_close_f32(host(disc), orc["disc"])
    gdisc = dev(orc["disc"])
    flags = dev(np.full(E, 9, np.uint8))
    P.check("free_edges_var", P.lib.pump_free_edges_var(ptr(go), o.shape[0], ptr(goff),
                                                        ptr(gdisc), ptr(flags), E, None))
    assert np.array_equal(host(flags), orc["is_free_edges"])
    # freeEdges with the reference's own arguments (collisionCheck.cu:6-8):
    # numDisc in its slot, fixed numDisc + 1 rows per edge (mainPUMP.cu:448-451)
    for num_disc in (8, 1):
        W = num_disc + 1
        rows = fixed_rows(s, orc, W)
        off = np.arange(0, (E + 1) * W, W, dtype=np.int64)
        ref = np.zeros(E, np.uint8)
        O.po_free_edges(oracle.fp(o), o.shape[0], oracle.fp(rows), oracle.lp(off), E,
                        oracle.bp(ref), 4)
        grows = dev(rows)
        flags = dev(np.full(E, 9, np.uint8))
        P.check("free_edges", P.lib.pump_free_edges(ptr(go), o.shape[0], ptr(gs), None, num_disc,
                                                    ptr(grows), ptr(flags), E, None))
        got = host(flags)
        assert np.array_equal(got, ref), num_disc
        assert 0 < got.sum() < E
    sf = dev(np.zeros(s.shape[0], np.uint8))
    P.check("sample_free", P.lib.pump_sample_free(ptr(go), o.shape[0], ptr(gs), s.shape[0],
                                                  ptr(sf), None))
    assert np.array_equal(host(sf), orc["is_free_samples"])
    # the reference pair setupAsBs + cacheHalfspaces (no sentinel pass)
    max_hs = orc["max_hs"]
    hs = n_wp * max_hs
    as_, bs = dev(np.zeros((hs, 3), np.float32)), dev(np.zeros(hs, np.float32))
    cnt = dev(np.full(n_wp, -7, np.int32))
    emask = dev(np.full(E, 12345, np.uint64))
    P.check("setup_as_bs", P.lib.pump_setup_as_bs(hs, ptr(bs), ptr(as_), None))
    gfree = dev(orc["is_free_edges"])  # held until the launch has run
    P.check("cache_hs", P.lib.pump_cache_halfspaces(
        E, ptr(gdisc), ptr(gfree), o.shape[0], max_hs, ptr(go), 0, ptr(goff),
        ptr(as_), ptr(bs), ptr(cnt), orc["maxd2"], 0, ptr(emask), None))
    assert np.array_equal(host(cnt), orc["count_stored"])
    assert np.array_equal(host(bs), orc["bs"][:hs])
    assert np.array_equal(host(as_), orc["as_"][:hs])
    ref_mask = edge_mask_host(orc["wp_offsets"], orc["bs"][:hs], max_hs)
    assert ref_mask.any()
    assert np.array_equal(host(emask), ref_mask)
    rm.close()


@pytest.mark.parametrize("maxd2", [0.01, 0.05, float("nan"), -1.0])
def test_cache_halfspaces_many_boxes(maxd2):
    """cacheHalfspaces (hsmc.cu:10-99) on 600 boxes (past the 512 staged in
    LDS, ten candidate words), waypoints scattered over the unit cube in
    edges of 2-70 waypoints, some edges blocked, against the oracle: a
    neighbourhood holding many boxes (the first max_hs by index are kept), a
    NaN radius (every box passes the d2 > maxd2 skip) and a negative one."""
    rng = np.random.default_rng(23)
    n_obs, max_hs = 600, 5
    c = rng.uniform(0, 1, (n_obs, 3))
    h = rng.uniform(0.005, 0.04, (n_obs, 3))
    obs = np.zeros((n_obs, 12), np.float32)
    obs[:, :3], obs[:, 6:9] = c - h, c + h
    obs[:, 3:6], obs[:, 9:12] = -3.1, 3.1
    W = rng.integers(2, 71, 300)
    off = np.concatenate([[0], np.cumsum(W)]).astype(np.int64)
    n_wp, E = int(off[-1]), W.shape[0]
    disc = np.concatenate([rng.uniform(0, 1, (n_wp, 3)), rng.uniform(-0.5, 0.5, (n_wp, 3))],
                          1).astype(np.float32)
    disc[:40, 3:] = 0
    free = (rng.uniform(size=E) < 0.8).astype(np.uint8)
    md2 = np.float32(maxd2)
    hs = n_wp * max_hs
    ref_as, ref_bs = np.zeros((hs, 3), np.float32), np.zeros(hs, np.float32)
    ref_cnt = np.zeros(n_wp, np.int32)
    O.po_setup_as_bs(hs, oracle.fp(ref_bs), oracle.fp(ref_as))
    O.po_cache_halfspaces(E, oracle.fp(disc), oracle.lp(off), oracle.bp(free), n_obs, max_hs,
                          oracle.fp(obs), oracle.fp(ref_as), oracle.fp(ref_bs),
                          oracle.ip(ref_cnt), md2, 4)
    as_, bs = dev(np.zeros((hs, 3), np.float32)), dev(np.zeros(hs, np.float32))
    cnt, emask = dev(np.full(n_wp, -7, np.int32)), dev(np.zeros(E, np.uint64))
    gdisc, gfree, gobs, goff = dev(disc), dev(free), dev(obs), dev(off)
    P.check("cache_hs", P.lib.pump_cache_halfspaces(
        E, ptr(gdisc), ptr(gfree), n_obs, max_hs, ptr(gobs), 0, ptr(goff),
        ptr(as_), ptr(bs), ptr(cnt), md2, 1, ptr(emask), None))
    got_cnt = host(cnt)
    assert np.array_equal(got_cnt, ref_cnt)
    assert np.array_equal(host(bs).view(np.uint32), ref_bs.view(np.uint32))
    assert np.array_equal(host(as_).view(np.uint32), ref_as.view(np.uint32))
    assert np.array_equal(host(emask), edge_mask_host(off, ref_bs, max_hs))
    if maxd2 == 0.05 or maxd2 != maxd2:
        assert (got_cnt == max_hs).sum() > 100   # truncation by box index exercised
    if maxd2 < 0:
        assert not got_cnt.any()


def test_setup_valid_particles():
    v = dev(np.zeros((37, 101), np.uint8))
    cp = dev(np.zeros(37, np.float32))
    tt = dev(np.ones(37, np.float32))
    P.check("setup_valid", P.lib.pump_setup_valid_particles(ptr(v), 101, 37, ptr(cp), ptr(tt),
                                                            None))
    assert host(v).all() and np.all(host(cp) == 1) and np.all(host(tt) == 0)
    b = dev(np.zeros((5, 3), np.uint64))
    P.check("setup_bits", P.lib.pump_setup_valid_bits(ptr(b), 150, 5, None))
    hb = host(b)
    assert np.all(hb[:, :2] == np.uint64(0xFFFFFFFFFFFFFFFF))
    assert np.all(hb[:, 2] == np.uint64((1 << 22) - 1))


def test_select_rn_order_statistic():
    rng = np.random.default_rng(17)
    n = 3_000_017
    keys = rng.normal(10, 30, n).astype(np.float32)
    keys[:1000] = -0.0
    keys[1000:2000] = 0.0
    keys[2000:2100] = keys[5000]  # ties
    g = dev(keys)
    ws = dev(np.zeros(P.lib.pump_select_rn_workspace(n), np.uint8))
    srt = np.sort(keys)
    for k in (0, 1, 999, 1500, n // 3, 2222222, n - 1):
        rn = ctypes.c_float(0)
        P.check("select", P.lib.pump_select_rn(ptr(g), n, k, ctypes.byref(rn), ptr(ws),
                                               ws.nbytes, None))
        assert np.float32(rn.value) == srt[k] or (rn.value == 0 and srt[k] == 0)


def quartic_q(t, a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    dp_ = b[:3] - a[:3]
    D = (dp_ ** 2).sum()
    S = (dp_ * (a[3:] + b[3:])).sum()
    R = (a[3:] ** 2 + a[3:] * b[3:] + b[3:] ** 2).sum()
    k4 = 1 + float(np.float32(-9.81)) ** 2  # g is the float constant of 2pBVP.cu:3
    return k4 * t ** 4 - 4 * R * t ** 2 + 24 * S * t - 36 * D, (
        k4 * t ** 4 + 4 * R * t ** 2 + 24 * abs(S) * t + 36 * D)


def test_polish_improves_arrival_time():
    """FP64 polish (chosen form): each polished arrival time is a root of the
    stationarity quartic q(t) = t^4 c'(t) (or tmax with c'(tmax) < 0), at the
    local minimum the bisection found, and closer to it than the bisection."""
    cfg, s, o, ot, xc, rm, orc = build_pair(CASES[0], polish=1)
    t64 = rm.polished()
    _, t32 = rm.pairs()
    num = s.shape[0]
    rng = np.random.default_rng(0)
    idx = rng.integers(0, num * (num - 1), 3000)
    e32 = e64 = 0.0
    at_tmax = 0
    for q in idx:
        i = q // (num - 1)
        kk = q % (num - 1)
        j = kk + (kk >= i)
        a, b = s[i].copy(), s[j].copy()
        val, scale = quartic_q(t64[q], a, b)
        if t64[q] == 5.0:
            at_tmax += 1
            assert val < 0
            continue
        assert abs(val) <= 1e-11 * scale, (t32[q], t64[q], val, scale)
        ref = O.po64_topt_near(oracle.fp(a), oracle.fp(b), float(t32[q]), 5.0)
        assert abs(t64[q] - ref) <= 1e-9 * ref
        e32 = max(e32, abs(t32[q] - ref) / ref)
        e64 = max(e64, abs(t64[q] - ref) / ref)
    print("max rel err vs FP64 root: bisection", e32, "polished", e64, "kept tmax", at_tmax,
          flush=True)
    assert e64 < e32
    rm.close()
