"""Writes the golden fixtures in this folder from the reference's own 2pBVP.cu.

The fixtures are input/output vectors of the reference's host functions
(cost, dcost, toptBisection, pathPoint, findDiscretizedPath, findOptimalPath
and the fillCoptsTopts pair order), evaluated by oracle/_ref/libpump_ref.so,
which oracle/ref/Makefile builds from the reference sources where they lie.
Run it in a container that has the reference:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from oracle import fp, ip  # noqa: E402


def states(rng, n, pos=(0.0, 1.0), vel=(-0.5, 0.5)):
    x = np.empty((n, 6), np.float32)
    x[:, :3] = rng.uniform(pos[0], pos[1], (n, 3))
    x[:, 3:] = rng.uniform(vel[0], vel[1], (n, 3))
    return x


def main():
    R = oracle.load_ref()
    if R is None:
        raise SystemExit("oracle/_ref/libpump_ref.so missing: make -C oracle/ref")
    rng = np.random.default_rng(20170101)

    # ---- scalar functions: cost / dcost / pathPoint ----------------------
    n = 2000
    x0 = states(rng, n)
    x1 = states(rng, n)
    # special rows: zero velocities, coincident positions, wide velocities
    x0[:50, 3:] = 0
    x1[50:100, :3] = x0[50:100, :3]
    x0[100:150, 3:] = rng.uniform(-3, 3, (50, 3))
    tau = rng.uniform(0.005, 5.0, n).astype(np.float32)
    tau[:20] = np.float32(5.0)
    tau[20:40] = np.float32(0.01)
    t = (rng.uniform(0, 1, n) * tau).astype(np.float32)
    t[40:60] = 0
    t[60:80] = tau[60:80]
    cost = np.empty(n, np.float32)
    dcost = np.empty(n, np.float32)
    pp = np.empty((n, 6), np.float32)
    topt = np.empty(n, np.float32)
    for k in range(n):
        a, b = x0[k].copy(), x1[k].copy()
        cost[k] = R.ref_cost(tau[k], fp(a), fp(b))
        dcost[k] = R.ref_dcost(tau[k], fp(a), fp(b))
        out = np.zeros(6, np.float32)
        R.ref_path_point(t[k], tau[k], fp(a), fp(b), fp(out))
        pp[k] = out
        topt[k] = R.ref_topt_bisection(fp(a), fp(b), np.float32(5.0))
    np.savez_compressed(os.path.join(HERE, "bvp_scalar.npz"), x0=x0, x1=x1, tau=tau, t=t,
                        cost=cost, dcost=dcost, path_point=pp, topt=topt)

    # ---- findDiscretizedPath at the reference's numDisc=8 and W=20 -------
    m = 300
    d0 = states(rng, m)
    d1 = states(rng, m)
    for nd in (8, 19):
        wp = np.zeros((m, nd + 1, 6), np.float32)
        tt = np.empty(m, np.float32)
        for k in range(m):
            buf = np.zeros(((nd + 1) * 6,), np.float32)
            a, b = d0[k].copy(), d1[k].copy()
            tt[k] = R.ref_find_discretized_path(fp(buf), fp(a), fp(b), nd)
            wp[k] = buf.reshape(nd + 1, 6)
        np.savez_compressed(os.path.join(HERE, f"bvp_discretized_nd{nd}.npz"), x0=d0, x1=d1,
                            topt=tt, waypoints=wp)

    # ---- findOptimalPath over node chains --------------------------------
    chains, lens, totals, paths = [], [], [], []
    cap = 4000
    for k in range(60):
        nwp = int(rng.integers(2, 8))
        xs = states(rng, nwp)
        xs[0, 3:] = 0
        xs[-1, 3:] = 0
        buf = np.zeros(cap * 6, np.float32)
        plen = ctypes_int()
        tot = R.ref_find_optimal_path(np.float32(0.05), fp(buf), cap, fp(xs), nwp, ctypes_ref(plen))
        chains.append(np.pad(xs, ((0, 8 - nwp), (0, 0))))
        lens.append((nwp, plen.value))
        totals.append(tot)
        paths.append(buf.reshape(cap, 6)[: plen.value].copy())
    maxlen = max(p.shape[0] for p in paths)
    P = np.zeros((len(paths), maxlen, 6), np.float32)
    for k, p in enumerate(paths):
        P[k, : p.shape[0]] = p
    np.savez_compressed(os.path.join(HERE, "bvp_optimal_path.npz"), chains=np.stack(chains),
                        lens=np.array(lens, np.int32), totals=np.array(totals, np.float32),
                        paths=P)

    # ---- all-pairs tables in fillCoptsTopts order ------------------------
    for num, seed in ((64, 0), (256, 1)):
        r2 = np.random.default_rng(seed)
        s = states(r2, num)
        s[0] = (0.1, 0.1, 0.1, 0, 0, 0)
        s[-1] = (0.9, 0.9, 0.9, 0, 0, 0)
        s = np.ascontiguousarray(s)
        copts = np.empty(num * (num - 1), np.float32)
        topts = np.empty(num * (num - 1), np.float32)
        R.ref_fill_pairs(fp(s), num, np.float32(5.0), fp(copts), fp(topts))
        np.savez_compressed(os.path.join(HERE, f"bvp_pairs_n{num}.npz"), samples=s, copts=copts,
                            topts=topts)
    print("golden fixtures written to", HERE)


def ctypes_int():
    import ctypes
    return ctypes.c_int(0)


def ctypes_ref(x):
    import ctypes
    return ctypes.byref(x)


if __name__ == "__main__":
    main()
