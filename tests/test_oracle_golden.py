"""The CPU oracle against the golden vectors of the reference's own 2pBVP.cu
(tests/golden/*.npz, written by tests/golden/make_golden.py from
oracle/_ref), bit for bit."""
import os

import numpy as np
import pytest

import oracle
from oracle import fp

O = oracle.load()


def _load(golden_dir, name):
    return np.load(os.path.join(golden_dir, name))


def test_cost_dcost_bitexact(golden_dir):
    g = _load(golden_dir, "bvp_scalar.npz")
    x0, x1, tau = g["x0"], g["x1"], g["tau"]
    c = np.array([O.po_cost(tau[k], fp(x0[k].copy()), fp(x1[k].copy())) for k in range(len(tau))],
                 np.float32)
    d = np.array([O.po_dcost(tau[k], fp(x0[k].copy()), fp(x1[k].copy())) for k in range(len(tau))],
                 np.float32)
    assert np.array_equal(c.view(np.uint32), g["cost"].view(np.uint32))
    assert np.array_equal(d.view(np.uint32), g["dcost"].view(np.uint32))


def test_topt_bisection_bitexact(golden_dir):
    g = _load(golden_dir, "bvp_scalar.npz")
    t = np.array([O.po_topt_bisection(fp(a.copy()), fp(b.copy()), 5.0)
                  for a, b in zip(g["x0"], g["x1"])], np.float32)
    assert np.array_equal(t.view(np.uint32), g["topt"].view(np.uint32))


def test_path_point_bitexact(golden_dir):
    g = _load(golden_dir, "bvp_scalar.npz")
    out = np.zeros(6, np.float32)
    for k in range(len(g["t"])):
        O.po_path_point(g["t"][k], g["tau"][k], fp(g["x0"][k].copy()), fp(g["x1"][k].copy()),
                        fp(out))
        assert out.tobytes() == g["path_point"][k].tobytes(), k


@pytest.mark.parametrize("nd", [8, 19])
def test_find_discretized_path_bitexact(golden_dir, nd):
    g = _load(golden_dir, f"bvp_discretized_nd{nd}.npz")
    for k in range(len(g["topt"])):
        buf = np.zeros((nd + 1) * 6, np.float32)
        t = O.po_find_discretized_path(fp(buf), fp(g["x0"][k].copy()), fp(g["x1"][k].copy()), nd)
        assert np.float32(t).tobytes() == g["topt"][k].tobytes()
        assert buf.tobytes() == g["waypoints"][k].reshape(-1).tobytes(), k


def test_find_optimal_path_bitexact(golden_dir):
    import ctypes
    g = _load(golden_dir, "bvp_optimal_path.npz")
    for k in range(len(g["totals"])):
        nwp, plen = g["lens"][k]
        xs = np.ascontiguousarray(g["chains"][k][:nwp])
        buf = np.zeros((plen + 8) * 6, np.float32)
        ln = ctypes.c_int(0)
        tot = O.po_find_optimal_path(np.float32(0.05), fp(buf), fp(xs), int(nwp),
                                     ctypes.byref(ln), int(plen + 8))
        assert ln.value == plen
        assert np.float32(tot).tobytes() == g["totals"][k].tobytes()
        assert buf[: plen * 6].tobytes() == g["paths"][k][:plen].reshape(-1).tobytes(), k


@pytest.mark.parametrize("num", [64, 256])
def test_fill_copts_topts_bitexact(golden_dir, num):
    g = _load(golden_dir, f"bvp_pairs_n{num}.npz")
    s = np.ascontiguousarray(g["samples"])
    c = np.empty(num * (num - 1), np.float32)
    t = np.empty(num * (num - 1), np.float32)
    O.po_fill_copts_topts(fp(s), num, 5.0, fp(c), fp(t), 4)
    assert np.array_equal(t.view(np.uint32), g["topts"].view(np.uint32))
    assert np.array_equal(c.view(np.uint32), g["copts"].view(np.uint32))


def test_oracle_against_reference_build_random():
    """Where the reference build exists (this container), compare the oracle
    with it on fresh random inputs as well."""
    R = oracle.load_ref()
    if R is None:
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    rng = np.random.default_rng(99)
    for _ in range(3000):
        a = np.concatenate([rng.uniform(0, 1, 3), rng.uniform(-1, 1, 3)]).astype(np.float32)
        b = np.concatenate([rng.uniform(0, 1, 3), rng.uniform(-1, 1, 3)]).astype(np.float32)
        tau = np.float32(rng.uniform(0.001, 6))
        assert O.po_cost(tau, fp(a), fp(b)) == R.ref_cost(tau, fp(a), fp(b))
        t1 = O.po_topt_bisection(fp(a), fp(b), 5.0)
        t2 = R.ref_topt_bisection(fp(a), fp(b), 5.0)
        assert np.float32(t1).tobytes() == np.float32(t2).tobytes()
