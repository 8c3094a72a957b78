"""pump_propagate_cp_bits_noalloc (the batched HSMC step with its workspace
in a fixed device array) against pump_propagate_cp_bits with a caller
workspace and against the oracle: two calls in a row on the same array, the
second with a smaller wavefront, in FP32 and split-BF16 form."""
import numpy as np
import pytest

import pump_amd as P

from test_gpu_parity import (dev, edge_mask_host, host, oracle_propagate, pack_bits, ptr,
                             random_hsmc_inputs, split_table, unpack_bits)

pytestmark = pytest.mark.gpu


def run_pair(d, precision):
    n_p, npar, nn = d["n_p"], d["n_parent"], d["n_new"]
    g = {k: dev(d[k]) for k in ("off", "as_", "bs", "offs", "topts", "times", "wf_edge",
                                 "wf_prev")}
    g_mask = dev(edge_mask_host(d["off"], d["bs"], d["max_hs"]))
    g_split = split_table(g["offs"], d["T"], d["n_mc"]) if precision else None
    src = dev(pack_bits(d["valid"][:npar], n_p))
    out = []
    for noalloc in (False, True):
        dst = dev(np.zeros((nn, (n_p + 63) // 64), np.uint64))
        cp = dev(np.zeros(nn, np.float32))
        cnt = dev(np.zeros(2, np.uint64))
        args = (nn, 0, ptr(g["off"]), ptr(g["wf_edge"]), ptr(g["wf_prev"]), ptr(g["as_"]),
                ptr(g["bs"]), ptr(g_mask), ptr(g["times"]), ptr(g["topts"]), d["dt"],
                d["max_hs"], ptr(g["offs"]), d["T"], d["om"], d["n_mc"], ptr(src), ptr(dst), 0,
                n_p, ptr(cp), 0, precision, g_split.p if g_split else None, cnt.ptr + 8,
                ptr(cnt))
        if noalloc:
            print("noalloc call", nn, "paths, precision", precision, flush=True)
            P.check("pump_propagate_cp_bits_noalloc",
                    P.lib.pump_propagate_cp_bits_noalloc(*args, None))
        else:
            ws_bytes = P.lib.pump_propagate_workspace(nn, d["T"])
            ws = P.DeviceArray(shape=(ws_bytes,), dtype=np.uint8, fill=0)
            P.check("pump_propagate_cp_bits",
                    P.lib.pump_propagate_cp_bits(*args, ws.p, ws_bytes, None))
        P.sync()
        out.append((host(dst), host(cp), host(cnt)))
    return out


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_matches_workspace_form(precision):
    # first a larger wavefront, then a smaller one on the same device array
    for seed, n_new, T in ((5, 20000, 300), (6, 6000, 120)):
        d = random_hsmc_inputs(seed, n_edges=3000, W=9, max_hs=5, n_p=700, n_mc=1024, T=T,
                               n_new=n_new, n_parent=64)
        keep = np.random.default_rng(seed).uniform(size=d["off"].shape[0] - 1) < 0.2
        for e in np.nonzero(~keep)[0]:
            d["bs"][d["off"][e] * d["max_hs"]: d["off"][e + 1] * d["max_hs"]] = 1111
        ref_v, ref_cp, ref_cl = oracle_propagate(d)
        (v_ws, cp_ws, c_ws), (v_na, cp_na, c_na) = run_pair(d, precision)
        assert np.array_equal(v_na, v_ws) and np.array_equal(cp_na, cp_ws)
        assert np.array_equal(c_na, c_ws)
        npar = d["n_parent"]
        assert np.array_equal(unpack_bits(v_na, d["n_p"]), ref_v[npar:])
        assert np.array_equal(cp_na, ref_cp[npar:])
        assert int(c_na[0]) == ref_cl
        assert cp_na.max() > cp_na.min()
