"""Workload definitions: the five benchmark configurations of the planning
path (NUM samples, obstacle worlds, waypoint policy, particles, noise) plus a
small case for tests.  Each returns a ProblemSpec, the search/roadmap Params
and the noise-table NoiseSpec.

Parameter mapping to the reference (mainPUMP.cu): NUM -> num, neighbour count
k -> rnPerc = k/(NUM-1) (mainPUMP.cu:261-263), waypoints W -> numDisc+1
(mainPUMP.cu:57), noise multiplier -> offsetMult (mainPUMP.cu:62,197-201),
particles -> numHSMCParticles (mainPUMP.cu:68), CP target / eta -> cpTarget /
cpFactor (mainPUMP.cu:110-111).
"""
import math

import numpy as np

from . import ProblemSpec, noise_spec, params

START = (0.1, 0.1, 0.1, 0.0, 0.0, 0.0)
GOAL = (0.9, 0.9, 0.9, 0.0, 0.0, 0.0)


def _spec(num, seed, n_obs, side=(0.05, 0.2), centre=(0.2, 0.8), kind=0, start=START,
          goal=GOAL):
    s = ProblemSpec()
    s.num = num
    s.seed = seed
    s.lo[:] = [0, 0, 0, -0.5, -0.5, -0.5]
    s.hi[:] = [1, 1, 1, 0.5, 0.5, 0.5]
    s.start[:] = list(start)
    s.goal[:] = list(goal)
    s.n_obs = n_obs
    s.obstacle_kind = kind
    s.centre_lo, s.centre_hi = centre
    s.side_lo, s.side_hi = side
    s.wall_thickness = 0.02
    s.wall_long_lo, s.wall_long_hi = 0.3, 0.8
    s.wall_short_lo, s.wall_short_hi = 0.05, 0.3
    s.vel_span = 3.1
    return s


def rn_perc(k, num):
    """rnPerc giving ~k neighbours per sample, as the float the reference uses."""
    return float(np.float32(k / (num - 1)))


def config(i, **over):
    """Configuration i (0-4) of the benchmark; keyword overrides apply to the
    Params (e.g. num=..., n_p=...)."""
    num = over.pop("num", None)
    if i == 0:
        num = num or 4000
        spec = _spec(num, 0, 10, side=(0.05, 0.2))
        prm = params(rn_perc=rn_perc(50, num), w_fixed=20, offset_mult=0.02, n_p=1000,
                     cp_target=0.01, cp_factor=2.0)
        ns = noise_spec(n_mc=2048)
        name = f"plain_num{num}_obs10_w20_p1000"
    elif i == 1:
        num = num or 16000
        spec = _spec(num, 0, 30, side=(0.03, 0.15))
        prm = params(rn_perc=rn_perc(80, num), w_fixed=0, wp_dt=0.05, w_cap=64, offset_mult=0.03,
                     n_p=2000, cp_target=0.01, cp_factor=2.0)
        ns = noise_spec(n_mc=2048)
        name = f"num{num}_obs30_wceil64_p2000"
    elif i == 3:
        num = num or 32000
        spec = _spec(num, 7, 200, kind=1, centre=(0.0, 1.0))
        prm = params(rn_perc=rn_perc(50, num), w_fixed=0, wp_dt=0.05, w_cap=64, offset_mult=0.02,
                     n_p=4096, cp_target=0.005, cp_factor=1.5, mccp_batch=64)
        ns = noise_spec(n_mc=4096)
        name = f"walls_num{num}_obs200_p4096"
    elif i == 4:
        num = num or 64000
        spec = _spec(num, 4, 100, side=(0.02, 0.1))
        prm = params(rn_perc=rn_perc(50, num), w_fixed=0, wp_dt=0.05, w_cap=128,
                     offset_mult=0.04, n_p=8192, cp_target=0.01, cp_factor=2.0)
        ns = noise_spec(n_mc=8192)
        name = f"num{num}_obs100_w128_p8192"
    else:
        raise ValueError(f"no configuration {i} (the batched set is batch_config())")
    for k, v in over.items():
        setattr(prm, "lambda_" if k == "lambda" else k, v)
    return dict(name=name, spec=spec, params=prm, noise=ns)


def batch_config(n_instances=256, seed=2):
    """Configuration 2: uneven independent instances (NUM log-uniform in
    [1000, 8000], 5-50 boxes, random start/goal corners, offsetMult in
    [0.01, 0.05], particles in {500, 1000, 2000})."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_instances):
        num = int(math.exp(rng.uniform(math.log(1000), math.log(8000))))
        n_obs = int(rng.integers(5, 51))
        start = tuple(rng.uniform(0, 0.2, 3)) + (0.0, 0.0, 0.0)
        goal = tuple(rng.uniform(0.8, 1.0, 3)) + (0.0, 0.0, 0.0)
        spec = _spec(num, 1000 + k, n_obs, side=(0.03, 0.2), start=start, goal=goal)
        n_p = int(rng.choice([500, 1000, 2000]))
        prm = params(rn_perc=rn_perc(50, num), w_fixed=0, wp_dt=0.05, w_cap=64,
                     offset_mult=float(rng.uniform(0.01, 0.05)), n_p=n_p, cp_target=0.01,
                     cp_factor=2.0)
        out.append(dict(name=f"batch{k}_num{num}", spec=spec, params=prm,
                        noise=noise_spec(n_mc=2048)))
    return out


def small(num=256, n_obs=10, seed=0, w_fixed=20, n_p=256, offset_mult=0.02, k=30, n_mc=512,
          tmax_steps=2001, **over):
    """A small instance for parity tests."""
    spec = _spec(num, seed, n_obs, side=(0.05, 0.2))
    prm = params(rn_perc=rn_perc(k, num), w_fixed=w_fixed, wp_dt=0.05, w_cap=64,
                 offset_mult=offset_mult, n_p=n_p, cp_target=0.01, cp_factor=2.0)
    for kk, v in over.items():
        setattr(prm, "lambda_" if kk == "lambda" else kk, v)
    return dict(name=f"small_num{num}", spec=spec, params=prm,
                noise=noise_spec(n_mc=n_mc, tmax_steps=tmax_steps))
