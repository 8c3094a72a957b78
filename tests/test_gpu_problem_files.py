"""The reference's own problem files end to end: its input.txt and
indoor.txt (fixtures under tests/golden/problems), NUM = 2000 Halton samples
(createSamplesHalton, as run.sh builds the reference), the reference's run
parameters (numDisc 8, rnPerc 0.05, 128 HSMC particles, up to 5
half-spaces, cp target 0.01 and eta 2 as in its README, offsetMult from the
file), the generated noise tables; the roadmap, the search and the
smoothing against the oracle.  With these noise tables the 0.01 target is
not met on either file at NUM = 2000 (the search reports it, as the
reference would); a 0.05 target on input.txt gives a solution to smooth."""
import os

import numpy as np
import pytest

import pump_amd as P
from oracle import pipeline as OP

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
NT = min(16, os.cpu_count() or 1)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_close_f32(got, ref, rel=1e-5, max_mismatch_frac=1e-5):
    """Bit-equal except for at most a 1e-5 fraction of entries within 1e-5
    relative: the kernels use correctly rounded tau^3, tau^4 where the
    reference's glibc pow misrounds about 0.2% of them, which can move the
    final float by one ulp (one waypoint of 1.8M at indoor.txt)."""
    g, r = bits(got), bits(ref)
    assert g.shape == r.shape
    same = g == r
    assert 1 - same.mean() <= max_mismatch_frac, f"{(~same).sum()} of {same.size} differ"
    if not same.all():
        a = np.asarray(got, np.float64)[~same]
        b = np.asarray(ref, np.float64)[~same]
        assert np.all(np.abs(a - b) <= rel * np.abs(b))


@pytest.mark.parametrize("name,cp_target", [("input.txt", 0.01), ("indoor.txt", 0.01),
                                             ("input.txt", 0.05)])
def test_problem_file_end_to_end(name, cp_target):
    pf = P.read_problem(os.path.join(HERE, "golden", "problems", name))
    s = P.halton_samples(2000, pf["init"], pf["goal"], pf["lo"], pf["hi"])
    o = pf["obstacles"]
    prm = P.params(offset_mult=float(pf["offset_mult"]), cp_target=cp_target, cp_factor=2.0,
                   keep_pairs=1)
    ot, xc = P.make_noise(P.noise_spec(n_mc=2048))
    print(name, "boxes", o.shape[0], "offsetMult", pf["offset_mult"], flush=True)
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
    assert_close_f32(rm.waypoints(), orc["disc"])
    hs = rm.halfspaces()
    assert np.array_equal(hs["count_stored"], orc["count_stored"])
    res = rm.search()
    got = P.search_result_dict(res)
    ref = OP.search(orc, s, s[-1], o, ot, xc, prm, nthreads=NT, path_cap=prm.path_cap)
    print("search", {k: got[k] for k in ("status", "itrs", "num_paths", "soln_cost", "soln_cp")},
          flush=True)
    for key in ("status", "itrs", "num_paths", "num_considered", "num_goal_paths", "soln_nodes"):
        assert got[key] == ref[key], key
    for key in ("soln_cost", "soln_cp", "soln_hsmc_cp"):
        assert np.float32(got[key]) == np.float32(ref[key]), key
    sm, path, xs_s, disc = rm.smooth(res)
    sref = OP.smooth(got, s, o, xc, prm)
    print("smooth", sm.status, sm.itrs, sm.cost_before, sm.cost, sm.cp, flush=True)
    assert sm.status == sref["status"]
    if sm.status == 0:
        assert sm.itrs == sref["itrs"]
        assert np.float32(sm.cost) == np.float32(sref["cost"])
        assert np.float32(sm.cp) == np.float32(sref["cp"])
        assert np.array_equal(bits(path), bits(sref["path"]))
        assert np.array_equal(bits(disc), bits(sref["discretized"]))
    rm.close()
