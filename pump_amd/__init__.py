"""pump_amd -- MI355X (gfx950) implementation of PUMP's roadmap-edge and
collision-probability path behind a C ABI (include/pump_abi.h) and a C++
host pipeline (include/pump_host.h).

This module loads ``libpump_amd.so`` (built in-tree by ``pump_amd.build``)
with ctypes.  There is no fallback: importing the package without the built
library raises, so nothing can silently run a non-native path.

Python mirrors of the reference entry points (fillCoptsTopts, freeEdges,
cacheHalfspaces, propagateValidParticles, calculateCP_HSMC,
collisionProbability, ...) are in ``pump_amd.ops`` (device tensors) and the
roadmap/search pipeline is ``pump_amd.Roadmap``.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpump_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is not built; run `python pump_amd/build.py` (hipcc, gfx950)")

lib = ctypes.CDLL(LIB_PATH)

c_int, c_float, c_double, c_int64 = ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_int64
c_uint64, c_uint32, c_size_t, c_void_p = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_size_t, ctypes.c_void_p
f32p = ctypes.POINTER(c_float)
f64p = ctypes.POINTER(c_double)
i32p = ctypes.POINTER(c_int)
i64p = ctypes.POINTER(c_int64)
u8p = ctypes.POINTER(ctypes.c_uint8)
u32p = ctypes.POINTER(c_uint32)
u64p = ctypes.POINTER(c_uint64)

PUMP_OK = 0
STATUS = {0: "ok", -1: "invalid argument", -2: "HIP runtime error", -3: "buffer too small",
          -4: "device allocation failed", -5: "argument outside the supported range"}


class PumpError(RuntimeError):
    def __init__(self, fn, code):
        detail = ""
        if code == -2:
            detail = " -- " + lib.pump_last_error().decode(errors="replace")
        super().__init__(f"{fn} failed: {code} ({STATUS.get(code, 'unknown')}){detail}")
        self.code = code


class ProblemSpec(ctypes.Structure):
    _fields_ = [
        ("num", c_int), ("seed", c_uint64), ("lo", c_float * 6), ("hi", c_float * 6),
        ("start", c_float * 6), ("goal", c_float * 6), ("n_obs", c_int), ("obstacle_kind", c_int),
        ("centre_lo", c_float), ("centre_hi", c_float), ("side_lo", c_float), ("side_hi", c_float),
        ("wall_thickness", c_float), ("wall_long_lo", c_float), ("wall_long_hi", c_float),
        ("wall_short_lo", c_float), ("wall_short_hi", c_float), ("vel_span", c_float),
    ]


class NoiseSpec(ctypes.Structure):
    _fields_ = [("n_mc", c_int), ("tmax_steps", c_int), ("dt", c_float), ("kp", c_double),
                ("kd", c_double), ("q", c_double), ("seed", c_uint64)]


class Params(ctypes.Structure):
    _fields_ = [
        ("tmax", c_float), ("rn_perc", c_float), ("w_fixed", c_int), ("wp_dt", c_float),
        ("w_cap", c_int), ("offset_mult", c_float), ("max_hs", c_int), ("dt", c_float),
        ("n_p", c_int), ("lambda_", c_float), ("num_buckets", c_int), ("eps_cost", c_float),
        ("eps_cp", c_float), ("cp_target", c_float), ("cp_factor", c_float), ("max_itrs", c_int),
        ("max_path_count", c_int64), ("validity_bits", c_int), ("precision", c_int),
        ("fused_edges", c_int), ("keep_pairs", c_int), ("polish", c_int), ("path_cap", c_int),
        ("mccp_batch", c_int), ("search_device", c_int),
    ]


class RoadmapInfo(ctypes.Structure):
    _fields_ = [
        ("num", c_int), ("n_obs", c_int), ("max_nn", c_int), ("max_hs", c_int), ("n_mc", c_int),
        ("tmax_steps", c_int), ("n_pairs", c_int64), ("rn_idx", c_int64), ("n_edges", c_int64),
        ("n_waypoints", c_int64), ("n_free_edges", c_int64), ("rn", c_float), ("maxd2", c_float),
        ("ms_pairs", c_float), ("ms_select", c_float), ("ms_edges_build", c_float),
        ("ms_nn", c_float), ("ms_sample_free", c_float), ("ms_edge_pass", c_float),
        ("ms_hs", c_float), ("ms_total", c_float), ("pair_steps", c_int64),
    ]


class RoadmapDev(ctypes.Structure):
    _fields_ = [(n, c_void_p) for n in (
        "samples", "obstacles", "copts", "topts", "copts_edge", "topts_edge", "h", "disc", "as_",
        "bs", "offsets_time", "xcomb", "row_start", "wp_offsets", "edge_src", "edge_dst",
        "count_stored", "is_free_edges", "is_free_samples", "edge_hs_mask", "offsets_split")]


class SearchResult(ctypes.Structure):
    _fields_ = [
        ("status", c_int), ("itrs", c_int), ("num_paths", c_int64), ("num_considered", c_int64),
        ("num_goal_paths", c_int), ("soln_path_idx", c_int), ("soln_cost", c_float),
        ("soln_cp", c_float), ("soln_time", c_float), ("soln_hsmc_cp", c_float),
        ("soln_len", c_int), ("soln_nodes", c_int * 256), ("clamped_time_idx", c_int64),
        ("particle_steps", c_int64), ("band_particles", c_int64), ("mccp_calls", c_int),
        ("ms_propagate", c_float), ("ms_total", c_float),
    ]


class RoadmapTables(ctypes.Structure):
    _fields_ = [("num", c_int), ("samples", f32p), ("obstacles", f32p), ("n_obs", c_int),
                ("n_edges", c_int64), ("row_start", i64p), ("edge_dst", i32p),
                ("copts_edge", f32p), ("topts_edge", f32p), ("h", f32p),
                ("is_free_edges", u8p), ("is_free_samples", u8p), ("wp_offsets", i64p),
                ("max_hs", c_int), ("as_", f32p), ("bs", f32p), ("rn", c_float)]


class BatchInstance(ctypes.Structure):
    _fields_ = [("samples", f32p), ("num", c_int), ("obstacles", f32p), ("n_obs", c_int),
                ("goal", f32p), ("prm", Params)]


class BatchInfo(ctypes.Structure):
    _fields_ = [("n_inst", c_int), ("pairs_blocks_per_cu", c_int), ("edges_per_block", c_int),
                ("n_pairs", c_int64), ("n_edges", c_int64), ("n_waypoints", c_int64),
                ("ms_pairs", c_float), ("ms_edges", c_float), ("ms_autotune", c_float),
                ("ms_total", c_float), ("edge_obs_tile", c_int), ("autotune_cached", c_int),
                ("mc_one_launch", c_int), ("mc_particles_per_wave", c_int)]


class ProblemFile(ctypes.Structure):
    _fields_ = [("init", c_float * 6), ("goal", c_float * 6), ("lo", c_float * 6),
                ("hi", c_float * 6), ("offset_mult", c_float), ("n_obs", c_int)]


class SmoothResult(ctypes.Structure):
    _fields_ = [("status", c_int), ("itrs", c_int), ("mccp_launches", c_int), ("n_nodes", c_int),
                ("path_rows", c_int), ("discretized_rows", c_int), ("cost_before", c_float),
                ("cost", c_float), ("cp", c_float), ("ms_total", c_float),
                ("trace", (c_float * 3) * 32)]


S = c_void_p  # stream / device pointer
_SIGS = {
    # pump_abi.h
    "pump_status_string": (ctypes.c_char_p, [c_int]),
    "pump_abi_version": (c_int, []),
    "pump_last_error": (ctypes.c_char_p, []),
    "pump_fill_copts_topts": (c_int, [S, c_int, c_float, S, S, S]),
    "pump_polish_topts": (c_int, [S, c_int, c_float, S, S, S]),
    "pump_select_rn_workspace": (c_size_t, [c_int64]),
    "pump_select_rn": (c_int, [S, c_int64, c_int64, f32p, S, c_size_t, S]),
    "pump_rn_index": (c_int64, [c_int64, c_float]),
    "pump_scan_workspace": (c_size_t, [c_int64]),
    "pump_count_edges": (c_int, [S, c_int, c_float, S, i64p, S, c_size_t, S]),
    "pump_fill_edges": (c_int, [S, S, c_int, c_float, S, S, S, S, S, S]),
    "pump_max_nn_size": (c_int, [S, S, c_int64, c_int, S, i32p, S]),
    "pump_nn_tables": (c_int, [S, S, S, S, c_int64, c_int, c_int, S, S, S, S, S]),
    "pump_heuristic": (c_int, [S, c_int, f32p, c_float, S, S]),
    "pump_waypoint_offsets": (c_int, [S, c_int64, c_int, c_float, c_int, S, i64p, S, c_size_t, S]),
    "pump_discretize_edges": (c_int, [S, S, S, S, c_int64, S, c_int, S, S]),
    "pump_sample_free": (c_int, [S, c_int, S, c_int, S, S]),
    "pump_free_edges": (c_int, [S, c_int, S, S, c_int, S, S, c_int64, S]),
    "pump_free_edges_var": (c_int, [S, c_int, S, S, S, c_int64, S]),
    "pump_edges_discretize_check": (c_int, [S, S, S, S, c_int64, S, c_int, S, c_int, S, S, S]),
    "pump_edges_discretize_check_fp32": (c_int, [S, S, S, S, c_int64, S, c_int, S, c_int, S, S,
                                                 S, S]),
    "pump_setup_as_bs": (c_int, [c_int64, S, S, S]),
    "pump_maxd2": (c_float, [c_float]),
    "pump_cache_halfspaces": (c_int, [c_int64, S, S, c_int, c_int, S, c_int, S, S, S, S, c_float,
                                      c_int, S, S]),
    "pump_setup_valid_particles": (c_int, [S, c_int, c_int64, S, S, S]),
    "pump_propagate_valid_particles": (c_int, [c_int, c_int, S, S, S, c_int64, S, S, S, S, S,
                                               c_float, c_int, S, c_int, c_float, c_int, S, c_int,
                                               S, S]),
    "pump_calculate_cp_hsmc": (c_int, [c_int, c_int64, S, S, c_int, S]),
    "pump_propagate_workspace": (c_size_t, [c_int, c_int]),
    "pump_propagate_workspace_pairs": (c_size_t, [c_int, c_int, c_int64]),
    "pump_propagate_cp": (c_int, [c_int, c_int, S, S, S, S, S, S, S, S, c_float, c_int, S, c_int,
                                  c_float, c_int, S, S, c_int64, c_int, S, c_int64, S, S, c_size_t,
                                  S]),
    "pump_gather_rows": (c_int, [S, S, c_int, S, c_int64, c_int64, S]),
    "pump_propagate_cp_bits": (c_int, [c_int, c_int, S, S, S, S, S, S, S, S, c_float, c_int, S,
                                       c_int, c_float, c_int, S, S, c_int64, c_int, S, c_int64,
                                       c_int, S, S, S, S, c_size_t, S]),
    "pump_propagate_cp_bits_noalloc": (c_int, [c_int, c_int, S, S, S, S, S, S, S, S, c_float,
                                               c_int, S, c_int, c_float, c_int, S, S, c_int64,
                                               c_int, S, c_int64, c_int, S, S, S, S]),
    "pump_propagate_cp_cow": (c_int, [c_int, c_int, c_int, S, S, S, S, S, S, S, S, c_float,
                                      c_int, S, c_int, c_float, c_int, S, S, S, c_int64, S,
                                      c_int, S, c_int64, c_int, S, S, S, S, c_size_t, S]),
    "pump_hsmc_autotune_cow": (c_int, [c_int, c_int, c_int, S, S, S, S, S, S, S, S, c_float,
                                       c_int, S, c_int, c_float, c_int, S, S, S, c_int64, S,
                                       c_int, S, c_int64, c_int, S, S, S, S, c_size_t, S,
                                       ctypes.POINTER(c_int)]),
    "pump_hsmc_particles_per_wave": (c_int, [c_int, c_int, c_int, c_int]),
    "pump_hsmc_set_particles_per_wave": (c_int, [c_int, c_int, c_int, c_int, c_int]),
    "pump_gather_rows_cow": (c_int, [S, S, S, S, c_int, S, c_int64, c_int64, c_int, S]),
    "pump_split_offsets_bytes": (c_size_t, [c_int, c_int]),
    "pump_split_offsets": (c_int, [S, c_int, c_int, S, S]),
    "pump_setup_valid_bits": (c_int, [S, c_int, c_int64, S]),
    "pump_collision_probability": (c_int, [S, c_int, S, c_int, c_int, c_float, S, c_int, f32p, S]),
    "pump_collision_probability_many": (c_int, [S, c_int, S, c_int, c_int, c_float, S, c_int64, S,
                                                c_int, S, c_int, S]),
    "pump_precomp_nn_count": (c_int, [S, c_int, c_float, S, S]),
    "pump_precomp_nn": (c_int, [S, c_int, c_float, c_int, S, S, S]),
    "pump_gmt_workspace": (c_size_t, [c_int]),
    "pump_gmt_wavefront": (c_int, [S, c_int, S, S, S, c_int, S, S, c_int, S, S, S, S, c_int, S, S,
                                   c_int, S, c_size_t, f32p, i32p, S]),
    # pump_host.h
    "pump_philox4x32": (None, [u32p, u32p, u32p]),
    "pump_gen_problem": (c_int, [ctypes.POINTER(ProblemSpec), f32p, f32p]),
    "pump_noise_defaults": (None, [ctypes.POINTER(NoiseSpec)]),
    "pump_make_noise": (c_int, [ctypes.POINTER(NoiseSpec), f32p, f32p, f32p]),
    "pump_make_noise_device": (c_int, [ctypes.POINTER(NoiseSpec), S, S, S, S, S]),
    "pump_params_defaults": (None, [ctypes.POINTER(Params)]),
    "pump_roadmap_build": (c_int, [f32p, c_int, f32p, c_int, f32p, f32p, f32p, c_int, c_int,
                                   ctypes.POINTER(Params), ctypes.POINTER(c_void_p)]),
    "pump_roadmap_from_tables": (c_int, [ctypes.POINTER(RoadmapTables), f32p, f32p, c_int,
                                         c_int, ctypes.POINTER(Params),
                                         ctypes.POINTER(c_void_p)]),
    "pump_search_copy_paths": (c_int, [c_void_p, c_int64, i32p, i32p, f32p, f32p, f32p]),
    "pump_roadmap_info_get": (c_int, [c_void_p, ctypes.POINTER(RoadmapInfo)]),
    "pump_roadmap_free": (None, [c_void_p]),
    "pump_roadmap_copy_pairs": (c_int, [c_void_p, f32p, f32p]),
    "pump_roadmap_copy_polished": (c_int, [c_void_p, f64p]),
    "pump_roadmap_copy_edges": (c_int, [c_void_p, i64p, i32p, i32p, f32p, f32p, f32p]),
    "pump_roadmap_copy_nn": (c_int, [c_void_p, i32p, i32p, f32p, i32p]),
    "pump_roadmap_copy_flags": (c_int, [c_void_p, i64p, u8p, u8p]),
    "pump_roadmap_copy_waypoints": (c_int, [c_void_p, f32p]),
    "pump_roadmap_copy_halfspaces": (c_int, [c_void_p, f32p, f32p, i32p]),
    "pump_roadmap_device": (c_int, [c_void_p, ctypes.POINTER(RoadmapDev)]),
    "pump_search_run": (c_int, [c_void_p, ctypes.POINTER(Params), ctypes.POINTER(SearchResult)]),
    "pump_read_problem": (c_int, [ctypes.c_char_p, ctypes.POINTER(ProblemFile), f32p, c_int]),
    "pump_halton_samples": (c_int, [c_int, c_int, f32p, f32p, f32p, f32p, f32p]),
    "pump_smooth_run": (c_int, [c_void_p, ctypes.POINTER(Params), ctypes.POINTER(SearchResult),
                                ctypes.POINTER(SmoothResult), f32p, c_int, f32p, f32p, c_int]),
    "pump_batch_build": (c_int, [c_int, ctypes.POINTER(BatchInstance), f32p, f32p, c_int, c_int,
                                 c_int, ctypes.POINTER(c_void_p)]),
    "pump_batch_info_get": (c_int, [c_void_p, ctypes.POINTER(BatchInfo)]),
    "pump_batch_roadmap": (c_int, [c_void_p, c_int, ctypes.POINTER(c_void_p)]),
    "pump_batch_replay": (c_int, [c_void_p, c_int, c_int, f32p]),
    "pump_batch_mc_build": (c_int, [c_void_p, c_int, c_int, c_uint64, f32p]),
    "pump_batch_mc_eager": (c_int, [c_void_p, c_int, f32p]),
    "pump_batch_mc_get": (c_int, [c_void_p, c_int, c_int, i32p, i32p, i32p, f32p, f32p, i32p]),
    "pump_batch_free": (None, [c_void_p]),
    "pump_find_optimal_path": (c_int, [c_float, f32p, c_int, f32p, c_int, f32p]),
    "pump_dev_alloc": (c_void_p, [c_size_t]),
    "pump_dev_free": (c_int, [c_void_p]),
    "pump_dev_upload": (c_int, [c_void_p, c_void_p, c_size_t]),
    "pump_dev_download": (c_int, [c_void_p, c_void_p, c_size_t]),
    "pump_dev_memset": (c_int, [c_void_p, c_int, c_size_t]),
    "pump_dev_copy": (c_int, [c_void_p, c_void_p, c_size_t]),
    "pump_dev_sync": (c_int, []),
    "pump_dev_count": (c_int, []),
    "pump_dev_name": (c_int, [ctypes.c_char_p, c_int]),
    "pump_stream_create": (c_int, [ctypes.POINTER(c_void_p)]),
    "pump_stream_destroy": (c_int, [c_void_p]),
    "pump_stream_sync": (c_int, [c_void_p]),
    "pump_timer_create": (c_void_p, [c_void_p]),
    "pump_timer_start": (c_int, [c_void_p]),
    "pump_timer_stop": (c_float, [c_void_p]),
    "pump_timer_destroy": (None, [c_void_p]),
    "pump_host_topt": (c_float, [f32p, f32p, c_float]),
    "pump_host_cost": (c_float, [c_float, f32p, f32p]),
    "pump_host_path_point": (None, [c_float, c_float, f32p, f32p, f32p]),
    "pump_host_div_rn": (None, [S, S, S, c_int64]),
    "pump_host_div_sq": (None, [S, S, S, c_int64]),
    "pump_host_div_magic": (None, [S, c_int64, ctypes.c_uint32, S]),
    "pump_host_edge_bbox": (None, [f32p, f32p, c_float, f32p, f32p]),
}

for _name, (_res, _args) in _SIGS.items():
    try:
        _fn = getattr(lib, _name)
    except AttributeError as exc:
        raise ImportError(f"{LIB_PATH} lacks {_name}; rebuild with `python pump_amd/build.py`") from exc
    _fn.restype = _res
    _fn.argtypes = _args

ABI_SYMBOLS = [n for n in _SIGS if not n.startswith("pump_philox") and n in _SIGS]


def check(fn, code):
    if code != PUMP_OK:
        raise PumpError(fn, code)


def _p(a, kind):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be contiguous"
    return a.ctypes.data_as(kind)


def fp(a):
    return _p(a, f32p)


def dp(a):
    return _p(a, f64p)


def ip(a):
    return _p(a, i32p)


def lp(a):
    return _p(a, i64p)


def bp(a):
    return _p(a, u8p)


# ---------------------------------------------------------------------------
# generators
# ---------------------------------------------------------------------------
def philox4x32(ctr, key):
    c = (c_uint32 * 4)(*ctr)
    k = (c_uint32 * 2)(*key)
    o = (c_uint32 * 4)()
    lib.pump_philox4x32(c, k, o)
    return list(o)


def gen_problem(spec):
    """(samples [num,6] f32, obstacles [n_obs,12] f32) for a ProblemSpec."""
    s = np.zeros((spec.num, 6), np.float32)
    o = np.zeros((max(spec.n_obs, 1), 12), np.float32)
    check("pump_gen_problem", lib.pump_gen_problem(ctypes.byref(spec), fp(s), fp(o)))
    return s, o[: spec.n_obs].copy()


def read_problem(path, obs_cap=4096):
    """The reference's problem file (mainPUMP.cu:152-230): dict of init,
    goal, lo, hi (float32[6]), offset_mult and obstacles [n_obs, 12]."""
    pf = ProblemFile()
    obs = np.zeros((obs_cap, 12), np.float32)
    check("pump_read_problem", lib.pump_read_problem(os.fsencode(path), ctypes.byref(pf),
                                                      fp(obs), obs_cap))
    arr = lambda a: np.array(list(a), np.float32)
    return dict(init=arr(pf.init), goal=arr(pf.goal), lo=arr(pf.lo), hi=arr(pf.hi),
                offset_mult=np.float32(pf.offset_mult), obstacles=obs[: pf.n_obs].copy())


def halton_samples(num, init, goal, lo, hi, skip=0):
    """createSamplesHalton (sampler.cu:20-57): [num, 6] float32."""
    out = np.zeros((num, 6), np.float32)
    args = [np.ascontiguousarray(a, np.float32) for a in (init, goal, lo, hi)]
    check("pump_halton_samples", lib.pump_halton_samples(num, skip, *[fp(a) for a in args],
                                                          fp(out)))
    return out


def precomp_nn(samples, r, nn_size=None):
    """Euclidean neighbours within r (precomp.cu:3-76) on the device:
    (nn_sizes [num], distances [num, nn_size], nn_edges [num, nn_size]).
    nn_size defaults to the largest count; rows keep their first nn_size
    neighbours in sample order and pad with -1 / 0."""
    s = DeviceArray(np.ascontiguousarray(samples, np.float32))
    num, r2 = s.shape[0], float(np.float32(r) * np.float32(r))
    sizes = DeviceArray(shape=(num,), dtype=np.int32)
    check("pump_precomp_nn_count", lib.pump_precomp_nn_count(s.p, num, r2, sizes.p, None))
    counts = sizes.numpy()
    if nn_size is None:
        nn_size = int(counts.max()) if num else 0
    dist = DeviceArray(shape=(num, nn_size), dtype=np.float32)
    edges = DeviceArray(shape=(num, nn_size), dtype=np.int32)
    check("pump_precomp_nn", lib.pump_precomp_nn(s.p, num, r2, nn_size, dist.p, edges.p, None))
    return counts, dist.numpy(), edges.numpy()


def gmt_wavefront(obstacles, nn_go_edges, nn_come_edges, distances_come, disc, is_free_samples,
                  goal_idx, wp_offsets=None, num_disc=0, nn_idxs=None, row_start=None,
                  edge_dst=None):
    """GMTwavefront (GMT.cu:3-261) on the device from host arrays: returns
    (costs [num], edges [num], cost_goal, itrs).  The connecting edge of a
    node is looked up in nn_idxs ([num, num], nnIdxs[to*NUM + from]) when
    given, else in the CSR rows (row_start, edge_dst); its waypoints are the
    rows wp_offsets[e]..wp_offsets[e+1]-1 of disc, or num_disc+1 rows per
    edge when wp_offsets is None."""
    go = np.ascontiguousarray(nn_go_edges, np.int32)
    num, m = go.shape
    hold = []

    def up(a, dt):
        if a is None:
            return None
        d = DeviceArray(np.ascontiguousarray(a, dt))
        hold.append(d)
        return d.p

    obs = np.ascontiguousarray(obstacles, np.float32).reshape(-1, 12)
    costs = DeviceArray(shape=(num,), dtype=np.float32)
    edges = DeviceArray(shape=(num,), dtype=np.int32)
    ws = DeviceArray(shape=(lib.pump_gmt_workspace(num),), dtype=np.uint8)
    cg, it = np.zeros(1, np.float32), np.zeros(1, np.int32)
    check("pump_gmt_wavefront", lib.pump_gmt_wavefront(
        up(obs, np.float32) if len(obs) else None, len(obs), up(distances_come, np.float32),
        up(go, np.int32), up(nn_come_edges, np.int32), m, up(disc, np.float32),
        up(wp_offsets, np.int64), num_disc, up(nn_idxs, np.int32), up(row_start, np.int64),
        up(edge_dst, np.int32), up(is_free_samples, np.uint8), num, costs.p, edges.p, goal_idx,
        ws.p, ws.nbytes, fp(cg), ip(it), None))
    sync()
    return costs.numpy(), edges.numpy(), float(cg[0]), int(it[0])


def noise_spec(**kw):
    ns = NoiseSpec()
    lib.pump_noise_defaults(ctypes.byref(ns))
    for k, v in kw.items():
        setattr(ns, k, v)
    return ns


def make_noise(ns, sigma=False):
    """(offsets_time [T, n_mc, 3], xcomb [n_mc, T, 3][, sigma_pos [T,3,3]])."""
    ot = np.zeros((ns.tmax_steps, ns.n_mc, 3), np.float32)
    xc = np.zeros((ns.n_mc, ns.tmax_steps, 3), np.float32)
    sg = np.zeros((ns.tmax_steps, 3, 3), np.float32) if sigma else None
    check("pump_make_noise", lib.pump_make_noise(ctypes.byref(ns), fp(ot), fp(xc), fp(sg)))
    return (ot, xc, sg) if sigma else (ot, xc)


def make_noise_device(ns):
    """The noise tables generated on the device: (offsets_time [T, n_mc, 3],
    xcomb [n_mc, T, 3], sigma_pos [T, 3, 3], L [6, 6] FP64 Cholesky of Q)."""
    ot = DeviceArray(shape=(ns.tmax_steps, ns.n_mc, 3), dtype=np.float32)
    xc = DeviceArray(shape=(ns.n_mc, ns.tmax_steps, 3), dtype=np.float32)
    sg = DeviceArray(shape=(ns.tmax_steps, 3, 3), dtype=np.float32)
    ch = DeviceArray(shape=(6, 6), dtype=np.float64)
    check("pump_make_noise_device", lib.pump_make_noise_device(
        ctypes.byref(ns), ot.p, xc.p, sg.p, ch.p, None))
    sync()
    return ot.numpy(), xc.numpy(), sg.numpy(), ch.numpy()


def params(**kw):
    p = Params()
    lib.pump_params_defaults(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, "lambda_" if k == "lambda" else k, v)
    return p


# ---------------------------------------------------------------------------
# roadmap + search pipeline
# ---------------------------------------------------------------------------
class Roadmap:
    """Roadmap precompute on the device (mainPUMP.cu:248-504) and the PUMP
    search (PUMP.cu:3-405) over it."""

    def __init__(self, samples, obstacles, goal, offsets_time, xcomb, prm):
        self.samples = np.ascontiguousarray(samples, np.float32)
        self.obstacles = np.ascontiguousarray(obstacles, np.float32).reshape(-1, 12)
        self.goal = np.ascontiguousarray(goal, np.float32)
        self.prm = prm
        n_mc = offsets_time.shape[1] if offsets_time is not None else 0
        tmax_steps = offsets_time.shape[0] if offsets_time is not None else 0
        self._h = c_void_p()
        obs = self.obstacles if self.obstacles.size else np.zeros((1, 12), np.float32)
        code = lib.pump_roadmap_build(
            fp(self.samples), self.samples.shape[0], fp(obs), self.obstacles.shape[0],
            fp(self.goal), fp(offsets_time), fp(xcomb), n_mc, tmax_steps, ctypes.byref(prm),
            ctypes.byref(self._h))
        check("pump_roadmap_build", code)
        self.info = RoadmapInfo()
        check("pump_roadmap_info_get", lib.pump_roadmap_info_get(self._h, ctypes.byref(self.info)))

    @classmethod
    def from_tables(cls, tables, samples, obstacles, offsets_time, xcomb, prm):
        """A roadmap from explicit search tables (pump_roadmap_from_tables):
        tables holds row_start, edge_dst, copts_edge, topts_edge, h,
        is_free_edges, is_free_samples, wp_offsets, as_, bs, max_hs, rn."""
        rm = cls.__new__(cls)
        rm.samples = np.ascontiguousarray(samples, np.float32)
        rm.obstacles = np.ascontiguousarray(obstacles, np.float32).reshape(-1, 12)
        rm.goal = rm.samples[-1].copy()
        rm.prm = prm
        keep = {k: np.ascontiguousarray(tables[k], dt) for k, dt in (
            ("row_start", np.int64), ("edge_dst", np.int32), ("copts_edge", np.float32),
            ("topts_edge", np.float32), ("h", np.float32), ("is_free_edges", np.uint8),
            ("is_free_samples", np.uint8), ("wp_offsets", np.int64), ("as_", np.float32),
            ("bs", np.float32))}
        obs = rm.obstacles if rm.obstacles.size else np.zeros((1, 12), np.float32)
        t = RoadmapTables()
        t.num = rm.samples.shape[0]
        t.samples = fp(rm.samples)
        t.obstacles = fp(obs)
        t.n_obs = rm.obstacles.shape[0]
        t.n_edges = keep["edge_dst"].shape[0]
        t.row_start, t.edge_dst = lp(keep["row_start"]), ip(keep["edge_dst"])
        t.copts_edge, t.topts_edge = fp(keep["copts_edge"]), fp(keep["topts_edge"])
        t.h = fp(keep["h"])
        t.is_free_edges, t.is_free_samples = bp(keep["is_free_edges"]), bp(keep["is_free_samples"])
        t.wp_offsets = lp(keep["wp_offsets"])
        t.max_hs = int(tables["max_hs"])
        t.as_, t.bs = fp(keep["as_"]), fp(keep["bs"])
        t.rn = float(tables["rn"])
        ot = np.ascontiguousarray(offsets_time, np.float32)
        xc = np.ascontiguousarray(xcomb, np.float32)
        rm._h = c_void_p()
        check("pump_roadmap_from_tables", lib.pump_roadmap_from_tables(
            ctypes.byref(t), fp(ot), fp(xc), ot.shape[1], ot.shape[0], ctypes.byref(prm),
            ctypes.byref(rm._h)))
        rm.info = RoadmapInfo()
        check("pump_roadmap_info_get", lib.pump_roadmap_info_get(rm._h, ctypes.byref(rm.info)))
        return rm

    @classmethod
    def _borrowed(cls, handle, owner, samples, obstacles, goal, prm):
        """A roadmap of a batch (freed with the batch)."""
        rm = cls.__new__(cls)
        rm.samples, rm.obstacles, rm.goal, rm.prm = samples, obstacles, goal, prm
        rm._h = handle
        rm._owner = owner
        rm.info = RoadmapInfo()
        check("pump_roadmap_info_get", lib.pump_roadmap_info_get(rm._h, ctypes.byref(rm.info)))
        return rm

    def close(self):
        if self._h and getattr(self, "_owner", None) is None:
            lib.pump_roadmap_free(self._h)
        self._h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def pairs(self):
        n = self.info.n_pairs
        c, t = np.empty(n, np.float32), np.empty(n, np.float32)
        check("pump_roadmap_copy_pairs", lib.pump_roadmap_copy_pairs(self._h, fp(c), fp(t)))
        return c, t

    def polished(self):
        t = np.empty(self.info.n_pairs, np.float64)
        check("pump_roadmap_copy_polished", lib.pump_roadmap_copy_polished(self._h, dp(t)))
        return t

    def edges(self):
        num, E = self.info.num, self.info.n_edges
        rs = np.empty(num + 1, np.int64)
        src, dst = np.empty(E, np.int32), np.empty(E, np.int32)
        ce, te, h = np.empty(E, np.float32), np.empty(E, np.float32), np.empty(num, np.float32)
        check("pump_roadmap_copy_edges",
              lib.pump_roadmap_copy_edges(self._h, lp(rs), ip(src), ip(dst), fp(ce), fp(te), fp(h)))
        return dict(row_start=rs, edge_src=src, edge_dst=dst, copts_edge=ce, topts_edge=te, h=h)

    def nn_tables(self, dense=False):
        num, m = self.info.num, self.info.max_nn
        go = np.empty((num, m), np.int32)
        come = np.empty((num, m), np.int32)
        dist = np.empty((num, m), np.float32)
        idx = np.empty((num, num), np.int32) if dense else None
        check("pump_roadmap_copy_nn",
              lib.pump_roadmap_copy_nn(self._h, ip(go), ip(come), fp(dist), ip(idx)))
        return dict(nn_go_edges=go, nn_come_edges=come, distances_come=dist, nn_idxs=idx)

    def flags(self):
        E, num = self.info.n_edges, self.info.num
        wo = np.empty(E + 1, np.int64)
        fe, fs = np.empty(E, np.uint8), np.empty(num, np.uint8)
        check("pump_roadmap_copy_flags", lib.pump_roadmap_copy_flags(self._h, lp(wo), bp(fe), bp(fs)))
        return dict(wp_offsets=wo, is_free_edges=fe, is_free_samples=fs)

    def waypoints(self):
        d = np.empty((self.info.n_waypoints, 6), np.float32)
        check("pump_roadmap_copy_waypoints", lib.pump_roadmap_copy_waypoints(self._h, fp(d)))
        return d

    def halfspaces(self):
        hs = self.info.n_waypoints * max(self.info.max_hs, 1)
        a = np.empty((hs, 3), np.float32)
        b = np.empty(hs, np.float32)
        c = np.empty(self.info.n_waypoints, np.int32)
        check("pump_roadmap_copy_halfspaces",
              lib.pump_roadmap_copy_halfspaces(self._h, fp(a), fp(b), ip(c)))
        return dict(as_=a, bs=b, count_stored=c)

    def device(self):
        d = RoadmapDev()
        check("pump_roadmap_device", lib.pump_roadmap_device(self._h, ctypes.byref(d)))
        return d

    def search(self, prm=None):
        r = SearchResult()
        check("pump_search_run",
              lib.pump_search_run(self._h, ctypes.byref(prm or self.prm), ctypes.byref(r)))
        return r

    def paths(self, n):
        """The last search's stored path table, rows [0, n)."""
        prev, node = np.empty(n, np.int32), np.empty(n, np.int32)
        cost, tm, cp = np.empty(n, np.float32), np.empty(n, np.float32), np.empty(n, np.float32)
        check("pump_search_copy_paths", lib.pump_search_copy_paths(
            self._h, n, ip(prev), ip(node), fp(cost), fp(tm), fp(cp)))
        return dict(path_prev=prev, path_node=node, path_cost=cost, path_time=tm, path_cp=cp)

    def smooth(self, search_result, prm=None, path_cap=4096, disc_cap=1 << 16):
        """Smoothing and controller discretisation (PUMP.cu:406-568) of a
        search result: (SmoothResult, kept path rows, last iterate's node
        states, discretised rows)."""
        out = SmoothResult()
        path = np.zeros((path_cap, 6), np.float32)
        xs = np.zeros((max(search_result.soln_len, 1), 6), np.float32)
        disc = np.zeros((disc_cap, 6), np.float32)
        check("pump_smooth_run", lib.pump_smooth_run(
            self._h, ctypes.byref(prm or self.prm), ctypes.byref(search_result),
            ctypes.byref(out), fp(path), path_cap, fp(xs), fp(disc), disc_cap))
        return out, path[: out.path_rows].copy(), xs, disc[: out.discretized_rows].copy()


class Batch:
    """Roadmaps of many independent problems built together: one launch per
    device pass for all instances, captured in hipGraphs (pump_batch_build).
    problems: list of (samples, obstacles, goal, Params)."""

    def __init__(self, problems, offsets_time, xcomb, autotune=True):
        self._keep = []
        arr = (BatchInstance * len(problems))()
        for k, (samples, obstacles, goal, prm) in enumerate(problems):
            smp = np.ascontiguousarray(samples, np.float32)
            obs = np.ascontiguousarray(obstacles, np.float32).reshape(-1, 12)
            obs_arg = obs if obs.size else np.zeros((1, 12), np.float32)
            g = np.ascontiguousarray(goal, np.float32)
            self._keep.append((smp, obs, obs_arg, g, prm))
            arr[k].samples, arr[k].num = fp(smp), smp.shape[0]
            arr[k].obstacles, arr[k].n_obs = fp(obs_arg), obs.shape[0]
            arr[k].goal, arr[k].prm = fp(g), prm
        n_mc = offsets_time.shape[1] if offsets_time is not None else 0
        tmax_steps = offsets_time.shape[0] if offsets_time is not None else 0
        self._h = c_void_p()
        check("pump_batch_build",
              lib.pump_batch_build(len(problems), arr, fp(offsets_time), fp(xcomb), n_mc,
                                   tmax_steps, 1 if autotune else 0, ctypes.byref(self._h)))
        self.info = BatchInfo()
        check("pump_batch_info_get", lib.pump_batch_info_get(self._h, ctypes.byref(self.info)))

    def __len__(self):
        return self.info.n_inst

    def roadmap(self, k):
        h = c_void_p()
        check("pump_batch_roadmap", lib.pump_batch_roadmap(self._h, k, ctypes.byref(h)))
        smp, obs, _, g, prm = self._keep[k]
        return Roadmap._borrowed(h, self, smp, obs, g, prm)

    def mc_build(self, n_parents=4096, max_paths=1 << 20, seed=1):
        """Captures the MC step of every instance (stage 2); ms of its first launch."""
        ms = c_float()
        check("pump_batch_mc_build", lib.pump_batch_mc_build(self._h, n_parents, max_paths, seed,
                                                             ctypes.byref(ms)))
        self._mc_parents = n_parents
        check("pump_batch_info_get", lib.pump_batch_info_get(self._h, ctypes.byref(self.info)))
        return ms.value

    def mc_eager(self, reps=1):
        """Mean ms of the same MC steps launched without the graph."""
        ms = c_float()
        check("pump_batch_mc_eager", lib.pump_batch_mc_eager(self._h, reps, ctypes.byref(ms)))
        return ms.value

    def mc_get(self, k):
        """Instance k's MC wavefront and results: dict of edge, prev, times, cp, row_src."""
        n = c_int()
        check("pump_batch_mc_get", lib.pump_batch_mc_get(self._h, k, 0, ctypes.byref(n), None,
                                                         None, None, None, None))
        n_new = n.value
        cap = max(n_new, self._mc_parents)
        edge, prev = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
        times, cp = np.zeros(cap, np.float32), np.zeros(cap, np.float32)
        rs = np.zeros(cap, np.int32)
        check("pump_batch_mc_get", lib.pump_batch_mc_get(self._h, k, cap, ctypes.byref(n), ip(edge),
                                                         ip(prev), fp(times), fp(cp), ip(rs)))
        return dict(edge=edge[:n_new], prev=prev[:n_new], times=times[:self._mc_parents],
                    cp=cp[:n_new], row_src=rs[:n_new])

    def replay(self, stage, reps=1):
        """Mean ms per replay of stage 0 (pair pass), 1 (edge pass + half-spaces)
        or 2 (the MC step of every instance)."""
        ms = c_float()
        check("pump_batch_replay", lib.pump_batch_replay(self._h, stage, reps, ctypes.byref(ms)))
        return ms.value

    def close(self):
        if self._h:
            lib.pump_batch_free(self._h)
            self._h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# device buffers / streams / timers (this library's HIP runtime only)
# ---------------------------------------------------------------------------
class DeviceArray:
    """A device allocation holding a copy of a numpy array's bytes."""

    def __init__(self, host=None, shape=None, dtype=None, fill=None):
        if host is not None:
            host = np.ascontiguousarray(host)
            shape, dtype = host.shape, host.dtype
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self.ptr = lib.pump_dev_alloc(max(self.nbytes, 1))
        if not self.ptr:
            raise PumpError("pump_dev_alloc", -4)
        if host is not None and self.nbytes:
            check("pump_dev_upload", lib.pump_dev_upload(self.ptr, host.ctypes.data, self.nbytes))
        elif fill is not None and self.nbytes:
            check("pump_dev_memset", lib.pump_dev_memset(self.ptr, int(fill), self.nbytes))

    @property
    def p(self):
        return c_void_p(self.ptr)

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check("pump_dev_download", lib.pump_dev_download(out.ctypes.data, self.ptr, self.nbytes))
        return out

    def upload(self, host):
        host = np.ascontiguousarray(host, self.dtype)
        assert host.nbytes == self.nbytes
        check("pump_dev_upload", lib.pump_dev_upload(self.ptr, host.ctypes.data, self.nbytes))

    def free(self):
        if self.ptr:
            lib.pump_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_name():
    buf = ctypes.create_string_buffer(256)
    check("pump_dev_name", lib.pump_dev_name(buf, 256))
    return buf.value.decode()


def sync():
    check("pump_dev_sync", lib.pump_dev_sync())


class Stream:
    def __init__(self):
        self.h = c_void_p()
        check("pump_stream_create", lib.pump_stream_create(ctypes.byref(self.h)))

    def sync(self):
        check("pump_stream_sync", lib.pump_stream_sync(self.h))

    def close(self):
        if self.h:
            lib.pump_stream_destroy(self.h)
            self.h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Timer:
    """hipEvent pair recorded on a stream (None = default stream)."""

    def __init__(self, stream=None):
        self.h = lib.pump_timer_create(stream.h if stream is not None else None)
        if not self.h:
            raise PumpError("pump_timer_create", -2)

    def start(self):
        check("pump_timer_start", lib.pump_timer_start(self.h))

    def stop(self):
        ms = lib.pump_timer_stop(self.h)
        if ms < 0:
            raise PumpError("pump_timer_stop", -2)
        return ms

    def __del__(self):
        try:
            lib.pump_timer_destroy(self.h)
        except Exception:
            pass


def search_result_dict(r):
    d = {k: getattr(r, k) for k, _ in SearchResult._fields_ if k != "soln_nodes"}
    d["soln_nodes"] = list(r.soln_nodes)[: r.soln_len]
    return d


def find_optimal_path(xs, dt=0.05, cap=4096):
    xs = np.ascontiguousarray(xs, np.float32)
    out = np.zeros((cap, 6), np.float32)
    tt = c_float(0)
    rows = lib.pump_find_optimal_path(c_float(dt), fp(xs), xs.shape[0], fp(out), cap,
                                      ctypes.byref(tt))
    if rows < 0:
        raise PumpError("pump_find_optimal_path", -3)
    return out[:rows].copy(), tt.value
