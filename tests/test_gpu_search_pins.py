"""The GPU search on the hand-derived roadmaps of test_oracle_search.py.

Each case's tables go to the device through pump_roadmap_from_tables; the
search runs with its bookkeeping on the device and on the host, with byte and
bit validity, and (bits) with the split-BF16 half-space test.  Every form must
reproduce the hand-derived record: the result fields and the stored path
table (pump_search_copy_paths), literal for literal.
"""
import numpy as np
import pytest

import pump_amd as P

from test_oracle_search import CASES, build_case

pytestmark = pytest.mark.gpu

FORMS = [dict(search_device=1, validity_bits=0, precision=0),
         dict(search_device=1, validity_bits=1, precision=0),
         dict(search_device=1, validity_bits=1, precision=1),
         dict(search_device=0, validity_bits=0, precision=0),
         dict(search_device=0, validity_bits=1, precision=0)]

RESULT_KEYS = ("status", "itrs", "num_paths", "num_considered", "num_goal_paths",
               "soln_path_idx", "soln_nodes", "soln_cost", "soln_time", "soln_hsmc_cp",
               "soln_cp", "mccp_calls")
TABLE_KEYS = ("path_prev", "path_node", "path_cost", "path_time", "path_cp")


def gpu_params(over, form):
    base = dict(offset_mult=1.0, lambda_=0.5, dt=0.05, cp_target=0.25, cp_factor=2.0,
                eps_cost=1.0 / 16, eps_cp=1.0 / 64, num_buckets=2, n_p=8, max_itrs=20,
                max_path_count=1000, path_cap=4096)
    base.update(over)
    base.update(form)
    return P.params(**base)


@pytest.mark.parametrize("name,case,over,exp", CASES, ids=[c[0] for c in CASES])
def test_gpu_search_hand_derived(name, case, over, exp):
    rm_t, s, o, ot, xc = build_case(case)
    for form in FORMS:
        print("case", name, form, flush=True)
        prm = gpu_params(over, form)
        rm = P.Roadmap.from_tables(rm_t, s, o, ot, xc, prm)
        got = P.search_result_dict(rm.search(prm))
        for k in RESULT_KEYS:
            if k not in exp:
                continue
            g, v = got[k], exp[k]
            if isinstance(v, float):
                g, v = np.float32(g), np.float32(v)
            assert g == v, (name, form, k, g, v)
        if "path_prev" in exp:
            tab = rm.paths(got["num_paths"])
            for k in TABLE_KEYS:
                if k in exp:
                    assert np.array_equal(tab[k], np.asarray(exp[k], tab[k].dtype)), (name, form, k)
        rm.close()
