"""CPU oracle for PUMP's roadmap-edge and collision-probability path.

TEST INFRASTRUCTURE ONLY.  This package is the checker: tests/, the smoke()
entry point and bench.py's cpu_baseline leg load it; the product package
(pump_amd) never does.

``load()`` returns the ctypes handle of ``libpump_oracle.so`` (the C
restatement, built by ``make -C oracle``).  ``load_ref()`` returns the
handle of ``_ref/libpump_ref.so`` -- the reference's own 2pBVP.cu built
host-side -- or None where it has not been built (it needs the reference
sources, which exist only in the build container).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

f32p = ctypes.POINTER(ctypes.c_float)
f64p = ctypes.POINTER(ctypes.c_double)
i32p = ctypes.POINTER(ctypes.c_int)
i64p = ctypes.POINTER(ctypes.c_int64)
u8p = ctypes.POINTER(ctypes.c_uint8)
c_int, c_float, c_double, c_int64 = ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_int64


class SearchTrace(ctypes.Structure):
    _fields_ = [
        ("path_cap", c_int64), ("path_prev", i32p), ("path_node", i32p), ("path_bucket", i32p),
        ("path_cost", f32p), ("path_time", f32p), ("path_cp", f32p),
        ("itr_cap", c_int), ("itr_gidx", i32p), ("itr_n_new", i32p), ("itr_n_valid", i32p),
        ("itr_size_g", i32p), ("itr_threshold", f32p),
        ("wf_cap", c_int64), ("wf_prev", i32p), ("wf_edge", i32p), ("wf_fate", i32p),
        ("wf_len", c_int64),
    ]


class SearchIn(ctypes.Structure):
    _fields_ = [
        ("num", c_int), ("samples", f32p), ("goal", f32p), ("init_idx", c_int), ("goal_idx", c_int),
        ("row_start", i64p), ("edge_dst", i32p), ("topts_edge", f32p), ("copts_edge", f32p),
        ("h", f32p), ("is_free_edges", u8p), ("is_free_samples", u8p), ("wp_offsets", i64p),
        ("as_", f32p), ("bs", f32p), ("max_hs", c_int), ("offsets_time", f32p), ("xcomb", f32p),
        ("tmax_steps", c_int), ("n_mc", c_int), ("offset_mult", c_float), ("obstacles", f32p),
        ("n_obs", c_int), ("rn", c_float), ("lambda_", c_float), ("dt", c_float),
        ("cp_target", c_float), ("cp_factor", c_float), ("eps_cost", c_float), ("eps_cp", c_float),
        ("num_buckets", c_int), ("n_p", c_int), ("max_itrs", c_int), ("max_path_count", c_int64),
        ("path_cap", c_int), ("nthreads", c_int), ("trace", ctypes.POINTER(SearchTrace)),
    ]


class SearchOut(ctypes.Structure):
    _fields_ = [
        ("status", c_int), ("itrs", c_int), ("num_paths", c_int64), ("num_considered", c_int64),
        ("num_goal_paths", c_int), ("soln_path_idx", c_int), ("soln_cost", c_float),
        ("soln_cp", c_float), ("soln_time", c_float), ("soln_hsmc_cp", c_float),
        ("soln_len", c_int), ("soln_nodes", c_int * 256), ("clamped_time_idx", c_int64),
        ("particle_steps", c_int64), ("mccp_calls", c_int),
    ]


_SIGS = {
    "po_cost": (c_float, [c_float, f32p, f32p]),
    "po_dcost": (c_float, [c_float, f32p, f32p]),
    "po_topt_bisection": (c_float, [f32p, f32p, c_float]),
    "po_path_point": (None, [c_float, c_float, f32p, f32p, f32p]),
    "po_find_discretized_path": (c_float, [f32p, f32p, f32p, c_int]),
    "po_find_optimal_path": (c_float, [c_float, f32p, f32p, c_int, i32p, c_int]),
    "po_fill_copts_topts": (None, [f32p, c_int, c_float, f32p, f32p, c_int]),
    "po_rn_index": (c_int64, [c_int64, c_float]),
    "po_fill_rows": (None, [f32p, c_int, c_float, i32p, c_int, f32p, f32p, c_int]),
    "po_select_rn": (c_float, [f32p, c_int64, c_int64]),
    "po_count_edges": (c_int64, [f32p, c_int, c_float, i64p]),
    "po_fill_edges": (None, [f32p, f32p, c_int, c_float, i64p, i32p, f32p, f32p]),
    "po_max_nn_size": (c_int, [i64p, i32p, c_int64, c_int]),
    "po_nn_tables": (None, [i64p, i32p, f32p, c_int, c_int, i32p, i32p, f32p]),
    "po_heuristic": (None, [f32p, c_int, f32p, c_float, f32p]),
    "po_waypoints_for": (c_int, [c_float, c_int, c_float, c_int]),
    "po_discretize_edges": (None, [f32p, i64p, i32p, c_int, i64p, f32p, c_int]),
    "po_sample_free": (None, [f32p, c_int, f32p, c_int, u8p]),
    "po_is_motion_valid": (c_int, [f32p, f32p, f32p, f32p, c_int, f32p]),
    "po_free_edges": (None, [f32p, c_int, f32p, i64p, c_int64, u8p, c_int]),
    "po_setup_as_bs": (None, [c_int64, f32p, f32p]),
    "po_maxd2": (c_float, [c_float]),
    "po_cache_halfspaces": (None, [c_int64, f32p, i64p, u8p, c_int, c_int, f32p, f32p, f32p,
                                   i32p, c_float, c_int]),
    "po_setup_valid_particles": (None, [u8p, c_int, c_int64, f32p, f32p]),
    "po_propagate_valid_particles": (c_int64, [c_int, i64p, i32p, i32p, c_int64, f32p, f32p, f32p,
                                               f32p, c_float, c_int, f32p, c_int, c_float, c_int,
                                               u8p, c_int, c_int]),
    "po_calculate_cp_hsmc": (None, [c_int, c_int64, f32p, u8p, c_int]),
    "po_mccp_hits": (c_int, [f32p, c_int, f32p, c_int, c_int, c_float, f32p, c_int, i32p]),
    "po_collision_probability": (c_float, [f32p, c_int, f32p, c_int, c_int, c_float, f32p, c_int]),
    "po_search": (c_int, [ctypes.POINTER(SearchIn), ctypes.POINTER(SearchOut)]),
    "po_precomp_nn": (None, [f32p, c_int, c_float, c_int, i32p, f32p, i32p]),
    "po_gmt_wavefront": (c_int, [f32p, c_int, f32p, i32p, i32p, c_int, f32p, i64p, i64p, i32p,
                                 u8p, c_int, f32p, i32p, c_int, f32p, i32p]),
    "po_halton_samples": (c_int, [c_int, c_int, f32p, f32p, f32p, f32p, f32p]),
    "po_smooth": (c_int, [f32p, c_int, c_float, c_float, f32p, c_int, f32p, c_int, c_int, c_float,
                          c_float, c_float, c_int, f32p, i32p, f32p, f32p, c_int, f32p, i32p,
                          f32p, i32p]),
    "po64_cost": (c_double, [c_double, f32p, f32p]),
    "po64_topt": (c_double, [f32p, f32p, c_double]),
    "po64_topt_near": (c_double, [f32p, f32p, c_double, c_double]),
    "po64_mccp_hits": (c_int, [f32p, c_int, f32p, c_int, c_int, c_double, f64p, c_int]),
    "po64_find_optimal_path": (c_int, [c_double, f64p, f32p, c_int, c_int]),
    "po64_rows_cost_error": (None, [f32p, c_int, i32p, c_int, c_float, f32p, f32p, f64p, f64p,
                                    c_int]),
}

_REF_SIGS = {
    "ref_cost": (c_float, [c_float, f32p, f32p]),
    "ref_dcost": (c_float, [c_float, f32p, f32p]),
    "ref_topt_bisection": (c_float, [f32p, f32p, c_float]),
    "ref_path_point": (None, [c_float, c_float, f32p, f32p, f32p]),
    "ref_find_discretized_path": (c_float, [f32p, f32p, f32p, c_int]),
    "ref_find_optimal_path": (c_float, [c_float, f32p, c_int, f32p, c_int, i32p]),
    "ref_fill_pairs": (None, [f32p, c_int, c_float, f32p, f32p]),
}


def _bind(lib, sigs):
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def load():
    """The C restatement (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libpump_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/libpump_oracle.so is not built (run make -C oracle)")
        _LIB = _bind(ctypes.CDLL(path), _SIGS)
    return _LIB


def load_from(path):
    """A build of the restatement at another path (tests compile mutated
    copies of it to show that the pins fail them)."""
    return _bind(ctypes.CDLL(path), _SIGS)


def load_ref():
    """The reference's own 2pBVP.cu built host-side, or None if absent."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libpump_ref.so")
        if not os.path.exists(path):
            return None
        _REF = _bind(ctypes.CDLL(path), _REF_SIGS)
    return _REF


def ptr(a, kind=f32p):
    """ctypes pointer to a contiguous numpy array."""
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(kind)


def fp(a):
    return ptr(a, f32p)


def dp(a):
    return ptr(a, f64p)


def ip(a):
    return ptr(a, i32p)


def lp(a):
    return ptr(a, i64p)


def bp(a):
    return ptr(a, u8p)


def f32(x):
    return np.ascontiguousarray(x, dtype=np.float32)
