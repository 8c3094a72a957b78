"""Hand-derived pins of the oracle's PUMP search (PUMP.cu:3-405).

Each case is a small roadmap built by hand.  Its expected record -- the
wavefront of every iteration (parent path, edge) and the fate of each entry,
the per-iteration counters and bucket sizes, the stored path table with the
bucket of every stored path, and the returned path -- was worked out on
paper from the cited lines of PUMP.cu and is written here as literals; the
derivation is in the comments.  Nothing here is produced by running the
oracle.

Controlled CPs.  Every edge has two waypoints and one half-space slot per
waypoint.  The noise table puts particle p at offset (p, 0, 0) at every time
index, offsetMult is 1, and an edge "keeping c particles" stores the
half-space a = (1, 0, 0), b = c - 0.5 at its second waypoint.  The test of
PUMP.cu:647-651 is LHS*offsetMult < b with LHS = p exactly, so particles
0..c-1 stay valid and c..7 become invalid: with 8 particles the CP of a path
is (8 - min c over its edges) / 8, exactly.  Other slots hold the empty
sentinel (b = 1111, PUMP.cu:642).

Common parameters: rn = 2, lambda = 0.5 so dr = 1 (PUMP.cu:9); cpTarget
0.25 and cpFactor 2 so cpMinSoln = 0.125 and cpCutoff = 0.5 (PUMP.cu:10-11);
epsCost 1/16, epsCP 1/64 (both exact); 2 buckets; maxItrs 20.  Costs,
heuristics and arrival times are binary fractions, so every float sum below is
exact.  Without obstacles the full-MC CP of every path is 0.

The last test compiles mutated copies of the oracle (a flipped co-dominance
tie-break, a dropped "+1" in the bucket index, >= for the CP cutoff, < for the
goal test, no goal skip, no node-0 exclusion, a reversed goal-path tie order,
<= for the final CP test, no clamp of dCost at 0, no threshold step, staged
cut paths) and checks that every one of them fails at least one case.
"""
import os
import shutil
import subprocess
import tempfile
import types

import numpy as np
import pytest

import oracle
from oracle import pipeline as OP

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SRC = os.path.join(os.path.dirname(HERE), "oracle", "pump_oracle.c")

N_P = 8
T_STEPS = 256


def params(**over):
    p = types.SimpleNamespace(offset_mult=1.0, lambda_=0.5, dt=0.05, cp_target=0.25,
                              cp_factor=2.0, eps_cost=1.0 / 16, eps_cp=1.0 / 64, num_buckets=2,
                              n_p=N_P, max_itrs=20, max_path_count=1000)
    for k, v in over.items():
        setattr(p, k, v)
    return p


def line_samples(num):
    """Nodes on a line, zero velocity (only the full-MC CP reads them)."""
    s = np.zeros((num, 6), np.float32)
    s[:, 0] = 0.1 + 0.8 * np.arange(num) / max(num - 1, 1)
    s[:, 1] = 0.1
    s[:, 2] = 0.1
    return s


def make_case(num, edges, h, samples=None, free_samples=None, obstacles=None, rn=2.0):
    """edges: (src, dst, cost, tau, keep, free) in row-major order (edge id =
    list position); keep = particles kept by the edge's half-space or None."""
    E = len(edges)
    key = [(e[0], e[1]) for e in edges]
    assert key == sorted(key), "edges must be listed in the reference's numbering"
    src = np.array([e[0] for e in edges], np.int32)
    row_start = np.zeros(num + 1, np.int64)
    np.add.at(row_start, src + 1, 1)
    row_start = np.cumsum(row_start).astype(np.int64)
    as_ = np.full((2 * E, 3), -1.0, np.float32)
    bs = np.full(2 * E, 1111.0, np.float32)
    for k, e in enumerate(edges):
        if e[4] is not None:
            as_[2 * k + 1] = (1.0, 0.0, 0.0)
            bs[2 * k + 1] = e[4] - 0.5
    rm = dict(row_start=row_start,
              edge_dst=np.array([e[1] for e in edges], np.int32),
              copts_edge=np.array([e[2] for e in edges], np.float32),
              topts_edge=np.array([e[3] for e in edges], np.float32),
              h=np.array(h, np.float32),
              is_free_edges=np.array([1 if len(e) < 6 or e[5] else 0 for e in edges], np.uint8),
              is_free_samples=(np.ones(num, np.uint8) if free_samples is None
                               else np.array(free_samples, np.uint8)),
              wp_offsets=np.arange(0, 2 * E + 1, 2, dtype=np.int64),
              as_=as_, bs=bs, max_hs=1, rn=np.float32(rn), n_edges=E)
    ot = np.zeros((T_STEPS, N_P, 3), np.float32)
    ot[:, :, 0] = np.arange(N_P, dtype=np.float32)[None, :]
    xc = np.zeros((N_P, T_STEPS, 3), np.float32)
    s = line_samples(num) if samples is None else np.asarray(samples, np.float32)
    o = np.zeros((0, 12), np.float32) if obstacles is None else np.asarray(obstacles, np.float32)
    return rm, s, o, ot, xc


def run(case, prm, lib=None):
    rm, s, o, ot, xc = case
    return OP.search(rm, s, s[-1], o, ot, xc, prm, nthreads=1, path_cap=4096, trace=True,
                     lib=lib)


def check(got, exp):
    """Compare a search record with a hand-derived one (exact)."""
    for k, v in exp.items():
        g = got[k]
        if isinstance(v, (list, tuple)) and k not in ("soln_nodes",):
            g = np.asarray(g).tolist()
            v = np.asarray(v, np.asarray(got[k]).dtype).tolist()
        elif isinstance(v, float):
            g, v = np.float32(g), np.float32(v)
        assert g == v, f"{k}: got {g}, expected {v}"


# ---------------------------------------------------------------------------
# Case A -- buckets: (int)(dCost/dr) + Gidx + 1 (PUMP.cu:239-243), the clamp of
# dCost at 0 (:240), costThreshold += dr (:283), expansion of G[Gidx] only in
# stored order (:46-66), the goal test after the store (:258-266).
# Nodes 0 (init), 1, 2, 3, 4 (goal).  h = [2, 1.5, 0.25, 2.5, 0], tau 0.25.
CASE_A = dict(
    num=5,
    edges=[(0, 1, 0.5, 0.25, None),    # e0
           (0, 2, 2.5, 0.25, None),    # e1
           (0, 3, 1.0, 0.25, None),    # e2
           (1, 4, 1.5, 0.25, 6),       # e3  keeps 6 -> CP 2/8
           (2, 4, 0.75, 0.25, 6),      # e4  keeps 6 -> CP 2/8
           (3, 4, 0.125, 0.25, None)],  # e5
    h=[2.0, 1.5, 0.25, 2.5, 0.0])
# Iteration 1: Gidx 0, threshold h[0] = 2.  Path 0 expands e0, e1, e2 ->
#   new paths 1, 2, 3 at nodes 1, 2, 3; CP 0; cost 0.5, 2.5, 1.0; one path per
#   node, so none is dominated; stored in place (1, 2, 3).
#   dCost: 0.5+1.5-2 = 0 -> (0+0+1)%2 = 1; 2.5+0.25-2 = 0.75 -> (0+0+1)%2 = 1;
#   1.0+2.5-2 = 1.5 -> (1+0+1)%2 = 0.  sizeG = [1, 2], G[0] = [3], G[1] = [1, 2].
#   No goal path.  Gidx -> 1, threshold -> 3.
# Iteration 2: G[1] = [1, 2].  Path 1 (node 1) expands e3, path 2 (node 2)
#   expands e4 -> new paths 4, 5 at the goal: CP 2/8 each (<= cutoff 0.5, kept),
#   cost 0.5+1.5 = 2.0 and 2.5+0.75 = 3.25; goal paths skip the dominance test.
#   dCost: 2-3 = -1 -> clamp 0 -> (0+1+1)%2 = 0; 3.25-3 = 0.25 -> 0 -> 0.
#   sizeG = [3, 0] (G[0] = [3, 4, 5]).  Goal paths CP 0.25 > cpMinSoln 0.125.
#   Gidx -> 0, threshold -> 4.
# Iteration 3: G[0] = [3, 4, 5].  Path 3 (node 3) expands e5 -> path 6 at the
#   goal, cost 1.125, CP 0 (e5 has no half-space).  Paths 4, 5 sit at the goal,
#   which has no out-edges.  dCost 1.125-4 < 0 -> 0 -> (0+0+1)%2 = 1.
#   sizeG = [0, 1].  Goal path 6 has CP 0 <= 0.125: goal condition, stop.
# Selection (PUMP.cu:354-405): (1.125, 6) < (2.0, 4) < (3.25, 5); path 6's
#   full-MC CP is 0 < 0.25 -> returned, nodes [0, 3, 4], time 0.25+0.25.
EXP_A = dict(
    status=0, itrs=3, num_paths=7, num_considered=6, num_goal_paths=3, soln_path_idx=6,
    soln_nodes=[0, 3, 4], soln_cost=1.125, soln_time=0.5, soln_hsmc_cp=0.0, soln_cp=0.0,
    mccp_calls=1,
    itr_gidx=[0, 1, 0], itr_n_new=[3, 2, 1], itr_n_valid=[3, 2, 1],
    itr_threshold=[2.0, 3.0, 4.0], itr_size_g=[[1, 2], [3, 0], [0, 1]],
    wf_prev=[0, 0, 0, 1, 2, 3], wf_edge=[0, 1, 2, 3, 4, 5], wf_fate=[1, 2, 3, 4, 5, 6],
    path_prev=[-2, 0, 0, 0, 1, 2, 3], path_node=[0, 1, 2, 3, 4, 4, 4],
    path_bucket=[-1, 1, 1, 0, 0, 0, 1],
    path_cost=[0.0, 0.5, 2.5, 1.0, 2.0, 3.25, 1.125],
    path_cp=[0.0, 0.0, 0.0, 0.0, 0.25, 0.25, 0.0],
    path_time=[0.0, 0.25, 0.25, 0.25, 0.5, 0.5, 0.5])


# ---------------------------------------------------------------------------
# Case B -- co-dominance tie-break (PUMP.cu:188-218): two new paths at one
# node with equal cost and CP; the lower path index is kept.  Also the
# expansion filters (PUMP.cu:56-57): an edge flagged blocked and an edge into
# a sample flagged not free are skipped.
# Nodes 0, 1, 2, 3, 4 (not free), 5 (goal).  h = [3, 2, 2, 1, 0, 0].
CASE_B = dict(
    num=6,
    edges=[(0, 1, 1.0, 0.25, None),          # e0
           (0, 2, 1.0, 0.25, None),          # e1
           (0, 4, 0.5, 0.25, None),          # e2  into sample 4, not free
           (1, 3, 1.0, 0.25, None),          # e3
           (1, 5, 0.5, 0.25, None, False),   # e4  edge blocked
           (2, 3, 1.0, 0.25, None),          # e5
           (3, 5, 1.0, 0.25, None)],         # e6
    h=[3.0, 2.0, 2.0, 1.0, 0.0, 0.0],
    free_samples=[1, 1, 1, 1, 0, 1])
# Iteration 1: Gidx 0, threshold 3.  Path 0: e0 -> node 1, e1 -> node 2, e2
#   skipped (sample 4 not free).  New paths 1, 2, cost 1.0 each, CP 0, stored.
#   dCost 1+2-3 = 0 -> bucket 1 for both.  sizeG = [0, 2].  Gidx 1, threshold 4.
# Iteration 2: G[1] = [1, 2].  Path 1: e3 -> node 3, e4 skipped (blocked).
#   Path 2: e5 -> node 3.  New paths 3 (from 1) and 4 (from 2) at node 3, both
#   cost 2.0 and CP 0.  P[3] = [3, 4] staged.
#   Path 3 against 4: 0 < 1/64+0 and 2 < 1/16+2 (dominance form); co-dominance
#   0+1/64 >= 0, 2+1/16 >= 2 and 3 < 4 -> continue: path 3 is kept.
#   Path 4 against 3: dominance form holds, co-dominance fails (4 < 3 is
#   false) -> path 4 is removed (fate -2).
#   Path 3 stored at 3: dCost 2+1-4 = -1 -> 0 -> (0+1+1)%2 = 0.  sizeG = [1, 0].
#   Gidx 0, threshold 5.
# Iteration 3: G[0] = [3].  Path 3: e6 -> goal, new path 4, cost 3, CP 0.
#   Stored at 4, dCost 3-5 < 0 -> (0+0+1)%2 = 1.  Goal condition (CP 0).
# Selection: only path 4, full-MC CP 0 -> nodes [0, 1, 3, 5].
EXP_B = dict(
    status=0, itrs=3, num_paths=5, num_considered=5, num_goal_paths=1, soln_path_idx=4,
    soln_nodes=[0, 1, 3, 5], soln_cost=3.0, soln_time=0.75, soln_hsmc_cp=0.0, mccp_calls=1,
    itr_gidx=[0, 1, 0], itr_n_new=[2, 2, 1], itr_n_valid=[2, 1, 1],
    itr_threshold=[3.0, 4.0, 5.0], itr_size_g=[[0, 2], [1, 0], [0, 1]],
    wf_prev=[0, 0, 1, 2, 3], wf_edge=[0, 1, 3, 5, 6], wf_fate=[1, 2, 3, -2, 4],
    path_prev=[-2, 0, 0, 1, 3], path_node=[0, 1, 2, 3, 5], path_bucket=[-1, 1, 1, 0, 1],
    path_cost=[0.0, 1.0, 1.0, 2.0, 3.0], path_cp=[0.0] * 5,
    path_time=[0.0, 0.25, 0.25, 0.5, 0.75])


# ---------------------------------------------------------------------------
# Case C -- the CP cutoff is strict (pathCP > cpCutoff, PUMP.cu:178) and a cut
# path is not staged, so it cannot remove another; strict dominance by a path
# stored in an earlier iteration; the goal test is pathCP <= cpMinSoln
# (PUMP.cu:261).  epsCP = 1/4 here, so the cut path (CP 5/8, cost 1) would
# dominate the kept path at its node (CP 1/2, cost 2) if it were staged:
# 5/8 < 1/4 + 1/2 and 1 < 1/16 + 2, while 1 + 1/16 >= 2 fails.
# Nodes 0, 1, 2, 3, 4 (goal).  h = [3, 2, 1, 1, 0].
CASE_C = dict(
    num=5,
    edges=[(0, 1, 0.5, 0.25, None),    # e0
           (0, 3, 1.0, 0.25, None),    # e1
           (1, 2, 0.5, 0.25, 3),       # e2  keeps 3 -> CP 5/8 (cut)
           (1, 3, 2.5, 0.25, None),    # e3
           (2, 4, 1.0, 0.25, 7),       # e4
           (3, 2, 1.0, 0.25, 4),       # e5  keeps 4 -> CP 4/8 = cutoff (kept)
           (3, 4, 2.0, 0.25, 7)],      # e6  keeps 7 -> CP 1/8 = cpMinSoln
    h=[3.0, 2.0, 1.0, 1.0, 0.0])
PRM_C = dict(eps_cp=0.25)
# Iteration 1: Gidx 0, threshold 3.  Path 0: e0 -> node 1, e1 -> node 3; new
#   paths 1, 2 (cost 0.5, 1.0, CP 0) stored in place; dCost 0.5+2-3 < 0 and
#   1+1-3 < 0 -> bucket 1 both.  sizeG = [0, 2].  Gidx 1, threshold 4.
# Iteration 2: G[1] = [1, 2].  Path 1 (node 1): e2 -> node 2 (new 3), e3 ->
#   node 3 (new 4).  Path 2 (node 3): e5 -> node 2 (new 5), e6 -> goal (new 6).
#   CP: new 3 keeps 3 -> 5/8; new 4 -> 0; new 5 keeps 4 -> 4/8; new 6 keeps 7 -> 1/8.
#   Cost: 1.0, 3.0, 2.0, 3.0.
#   Cutoff: 5/8 > 0.5 -> new 3 cut (fate -1), not staged; 4/8 > 0.5 is false ->
#   new 5 staged.  P[3] = [2, 4], P[2] = [5], P[4] = [6].
#   New 4 (node 3) against stored path 2 (CP 0, cost 1): 0 < 1/4 + 0 and
#   1 < 1/16 + 3; co-dominance needs 1 + 1/16 >= 3, false -> removed (fate -2).
#   New 5 (node 2): only itself at node 2 -> kept.  New 6: goal, not tested.
#   Stored: new 5 -> 3 (dCost 2+1-4 = -1 -> 0 -> bucket 0), new 6 -> 4 (dCost
#   3+0-4 < 0 -> bucket 0).  sizeG = [2, 0].
#   Goal: path 4 has CP 1/8 <= 1/8 -> goal condition, stop after 2 iterations.
# Selection: path 4, nodes [0, 3, 4], full-MC CP 0.
EXP_C = dict(
    status=0, itrs=2, num_paths=5, num_considered=6, num_goal_paths=1, soln_path_idx=4,
    soln_nodes=[0, 3, 4], soln_cost=3.0, soln_time=0.5, soln_hsmc_cp=0.125, mccp_calls=1,
    itr_gidx=[0, 1], itr_n_new=[2, 4], itr_n_valid=[2, 2], itr_threshold=[3.0, 4.0],
    itr_size_g=[[0, 2], [2, 0]],
    wf_prev=[0, 0, 1, 1, 2, 2], wf_edge=[0, 1, 2, 3, 5, 6], wf_fate=[1, 2, -1, -2, 3, 4],
    path_prev=[-2, 0, 0, 2, 2], path_node=[0, 1, 3, 2, 4], path_bucket=[-1, 1, 1, 0, 0],
    path_cost=[0.0, 0.5, 1.0, 2.0, 3.0], path_cp=[0.0, 0.0, 0.0, 0.5, 0.125],
    path_time=[0.0, 0.25, 0.25, 0.5, 0.5])


# ---------------------------------------------------------------------------
# Case D -- goal paths skip the dominance test (PUMP.cu:191) and goal-path
# selection sorts (cost, path index) and returns the first whose full-MC CP is
# below cpTarget (PUMP.cu:354-405).
# All velocities are 0, so each hop of the nominal path (2pBVP.cu:5-36) runs
# on the straight segment between node positions.  One box spans x in
# [0.4, 0.6], y in [0.4, 1], z in [-1, 1].  Nodes: 0 = (0.1, 0.1, 0.1),
# 1 = (0.5, 0.7, 0.1), 2 = (0.5, 0.1, 0.1), 3 = (0.5, 0.2, 0.1),
# 4 = goal (0.9, 0.1, 0.1).  The hop 0 -> 1 crosses the plane x = 0.4 at
# y = 0.1 + 0.6 * 0.3/0.4 = 0.55 inside the face (z = 0.1): with zero noise
# every particle's segment across that plane hits the face (collisionCheck.cu:
# 114-143), in the first hop, far from the skipped last segment, so the full-MC
# CP of any path through node 1 is 8/8.  Through 2 or 3, y stays <= 0.2 < 0.4
# and the broad phase clears every segment: CP 0.
CASE_D = dict(
    num=5,
    edges=[(0, 1, 1.0, 0.25, None),    # e0
           (0, 2, 1.0, 0.25, None),    # e1
           (0, 3, 1.5, 0.25, None),    # e2
           (1, 4, 1.0, 0.25, None),    # e3
           (2, 4, 1.5, 0.25, None),    # e4
           (3, 4, 1.0, 0.25, None)],   # e5
    h=[2.0, 1.0, 1.0, 1.0, 0.0],
    samples=[[0.1, 0.1, 0.1, 0, 0, 0], [0.5, 0.7, 0.1, 0, 0, 0], [0.5, 0.1, 0.1, 0, 0, 0],
             [0.5, 0.2, 0.1, 0, 0, 0], [0.9, 0.1, 0.1, 0, 0, 0]],
    obstacles=[[0.4, 0.4, -1.0, -10.0, -10.0, -10.0, 0.6, 1.0, 1.0, 10.0, 10.0, 10.0]])
# Iteration 1: threshold 2.  New paths 1, 2, 3 (nodes 1, 2, 3; cost 1, 1, 1.5),
#   dCost 0, 0, 0.5 -> bucket 1 each.  sizeG = [0, 3].  Gidx 1, threshold 3.
# Iteration 2: G[1] = [1, 2, 3] expand e3, e4, e5 -> new paths 4, 5, 6, all at
#   the goal, cost 2.0, 2.5, 2.5, CP 0.  At any other node path 4 would remove
#   5 and 6 (2 < 1/16 + 2.5 while 2 + 1/16 >= 2.5 fails), and 5 would remove 6
#   by the index rule; at the goal none is tested, so all three are stored
#   (dCost <= 0 -> bucket 0).  Goal condition (CP 0).
# Selection: (2.0, 4), (2.5, 5), (2.5, 6).  Path 4 = [0, 1, 4]: full-MC CP 1,
#   not < 0.25.  Path 5 = [0, 2, 4]: CP 0 -> returned after 2 full-MC runs.
EXP_D = dict(
    status=0, itrs=2, num_paths=7, num_considered=6, num_goal_paths=3, soln_path_idx=5,
    soln_nodes=[0, 2, 4], soln_cost=2.5, soln_cp=0.0, mccp_calls=2,
    itr_n_new=[3, 3], itr_n_valid=[3, 3], itr_size_g=[[0, 3], [3, 0]],
    wf_fate=[1, 2, 3, 4, 5, 6],
    path_prev=[-2, 0, 0, 0, 1, 2, 3], path_node=[0, 1, 2, 3, 4, 4, 4],
    path_bucket=[-1, 1, 1, 1, 0, 0, 0], path_cost=[0.0, 1.0, 1.0, 1.5, 2.0, 2.5, 2.5])


# ---------------------------------------------------------------------------
# Case E -- a new path into node 0 is dropped at the store (PUMP.cu:227,
# "node is at the init") even when no path dominates it; costs can be
# negative (the gravity term of 2pBVP.cu:104).  The search ends when a
# wavefront is empty (PUMP.cu:68-71).
# Nodes 0, 1, 2, 3 (goal).  h = [1.5, 1, 0.5, 0].
CASE_E = dict(
    num=4,
    edges=[(0, 1, 0.5, 0.25, None),    # e0
           (1, 0, -2.0, 0.25, None),   # e1  back to node 0, negative cost
           (1, 2, 0.5, 0.25, None),    # e2
           (2, 3, 0.5, 0.25, 4)],      # e3  keeps 4 -> CP 4/8
    h=[1.5, 1.0, 0.5, 0.0])
# Iteration 1: threshold 1.5.  New path 1 (node 1, cost 0.5): dCost 0 ->
#   bucket 1.  Gidx 1, threshold 2.5.
# Iteration 2: path 1 expands e1 -> node 0 (new 2, cost -1.5) and e2 -> node 2
#   (new 3, cost 1.0).  New 2 against path 0 (CP 0, cost 0): 0 < 1/16 - 1.5 is
#   false -> not dominated; at the store its node is 0 -> dropped (fate -3).
#   New 3 stored at 2: dCost 1+0.5-2.5 = -1 -> 0 -> (0+1+1)%2 = 0.
#   sizeG = [1, 0].  Gidx 0, threshold 3.5.
# Iteration 3: path 2 (node 2) expands e3 -> goal (new 3, cost 1.5, CP 4/8 =
#   cutoff, kept).  Stored at 3, dCost < 0 -> bucket (0+0+1)%2 = 1.  CP 0.5 >
#   0.125: no goal condition.  sizeG = [0, 1].  Gidx 1, threshold 4.5.
# Iteration 4: path 3 sits at the goal, which has no out-edges: no new paths,
#   the loop ends (itrs = 4, nothing recorded for this iteration).
# Selection: path 3 = [0, 1, 2, 3], full-MC CP 0 < 0.25 -> returned.
EXP_E = dict(
    status=0, itrs=4, num_paths=4, num_considered=4, num_goal_paths=1, soln_path_idx=3,
    soln_nodes=[0, 1, 2, 3], soln_cost=1.5, soln_time=0.75, soln_hsmc_cp=0.5,
    itr_gidx=[0, 1, 0], itr_n_new=[1, 2, 1], itr_n_valid=[1, 1, 1],
    itr_threshold=[1.5, 2.5, 3.5], itr_size_g=[[0, 1], [1, 0], [0, 1]],
    wf_prev=[0, 1, 1, 2], wf_edge=[0, 1, 2, 3], wf_fate=[1, -3, 2, 3],
    path_prev=[-2, 0, 1, 2], path_node=[0, 1, 2, 3], path_bucket=[-1, 1, 0, 1],
    path_cost=[0.0, 0.5, 1.0, 1.5], path_cp=[0.0, 0.0, 0.0, 0.5])


# ---------------------------------------------------------------------------
# Case F -- an empty open set ends the search (PUMP.cu:270-276): the only new
# path is cut by CP, no bucket holds a path, no goal path exists.
CASE_F = dict(num=3, edges=[(0, 1, 0.5, 0.25, 3), (1, 2, 0.5, 0.25, None)], h=[1.0, 0.5, 0.0])
# Iteration 1: new path 1 keeps 3 particles -> CP 5/8 > 0.5: cut (fate -1).
#   Nothing stored; both buckets empty -> stop.  No goal path: no solution.
EXP_F = dict(status=1, itrs=1, num_paths=1, num_considered=1, num_goal_paths=0,
             soln_path_idx=-1, mccp_calls=0, itr_n_new=[1], itr_n_valid=[0],
             itr_size_g=[[0, 0]], wf_fate=[-1])

# Case G -- maxItrs (PUMP.cu:34-38): a chain 0 -> 1 -> ... -> 6 (goal) with
# maxItrs 3 stops after three iterations, one new path each, before the goal.
CASE_G = dict(num=7, edges=[(k, k + 1, 0.5, 0.25, None) for k in range(6)],
              h=[3.0, 2.5, 2.0, 1.5, 1.0, 0.5, 0.0])
PRM_G = dict(max_itrs=3)
# Iteration k (threshold 3 + (k-1)): new path k at node k, cost 0.5k, dCost
#   0.5k + (3 - 0.5k) - (3 + k - 1) = 1 - k <= 0 -> bucket (k-1+1)%2 = k%2.
EXP_G = dict(status=1, itrs=3, num_paths=4, num_considered=3, num_goal_paths=0,
             soln_path_idx=-1, itr_gidx=[0, 1, 0], itr_size_g=[[0, 1], [1, 0], [0, 1]],
             path_bucket=[-1, 1, 0, 1], path_cost=[0.0, 0.5, 1.0, 1.5])

# Case H -- the final test is full-MC CP < cpTarget, strict (PUMP.cu:400).
# Case D's world with cpTarget 1 (cpMinSoln 0.5, cpCutoff 2): the only goal
# path goes through node 1, whose full-MC CP is exactly 1 -> not below 1 ->
# no solution, though the path is reported as the last one tried.
CASE_H = dict(num=3, edges=[(0, 1, 1.0, 0.25, None), (1, 2, 1.0, 0.25, None)],
              h=[2.0, 1.0, 0.0],
              samples=[[0.1, 0.1, 0.1, 0, 0, 0], [0.5, 0.7, 0.1, 0, 0, 0],
                       [0.9, 0.1, 0.1, 0, 0, 0]],
              obstacles=CASE_D["obstacles"])
PRM_H = dict(cp_target=1.0)
EXP_H = dict(status=1, itrs=2, num_paths=3, num_goal_paths=1, soln_path_idx=2,
             soln_nodes=[0, 1, 2], soln_cp=1.0, mccp_calls=1)

CASES = [("A", CASE_A, {}, EXP_A), ("B", CASE_B, {}, EXP_B), ("C", CASE_C, PRM_C, EXP_C),
         ("D", CASE_D, {}, EXP_D), ("E", CASE_E, {}, EXP_E), ("F", CASE_F, {}, EXP_F),
         ("G", CASE_G, PRM_G, EXP_G), ("H", CASE_H, PRM_H, EXP_H)]


def build_case(case):
    return make_case(case["num"], case["edges"], case["h"], samples=case.get("samples"),
                     free_samples=case.get("free_samples"), obstacles=case.get("obstacles"))


@pytest.mark.parametrize("name,case,over,exp", CASES, ids=[c[0] for c in CASES])
def test_search_hand_derived(name, case, over, exp):
    got = run(build_case(case), params(**over))
    check(got, exp)


def _mutants():
    # (name, text in pump_oracle.c, replacement): each a plausible slip in
    # restating PUMP.cu
    return [
        ("tie-break flipped", "path_cost[pidx] && pidx < cmp)", "path_cost[pidx] && pidx > cmp)"),
        ("bucket +1 dropped", "(((int)(dcost / dr)) + gidx + 1) % nb",
         "(((int)(dcost / dr)) + gidx) % nb"),
        ("cutoff >=", "if (path_cp[pidx] > cp_cutoff) {", "if (path_cp[pidx] >= cp_cutoff) {"),
        ("goal test <", "if (path_cp[pidx] <= cp_min_soln) {",
         "if (path_cp[pidx] < cp_min_soln) {"),
        ("goal not skipped", "if (node == -1 || node == in->goal_idx)", "if (node == -1)"),
        ("node 0 kept", "if (node == -1 || node == 0) {", "if (node == -1) {"),
        ("goal ties by descending index", "return (a->idx > b->idx) - (a->idx < b->idx);",
         "return (a->idx < b->idx) - (a->idx > b->idx);"),
        ("final test <=", "if (cp < in->cp_target) {", "if (cp <= in->cp_target) {"),
        ("dCost not clamped", "dcost = dcost > 0 ? dcost : 0;", "(void)0;"),
        ("threshold not stepped", "cost_threshold += dr;", "(void)dr;"),
        ("cut paths staged", "fate[i] = -1;\n                continue;",
         "fate[i] = -1;\n                ivec_push(&P[node], (int)pidx);\n"
         "                pnew[node]++;\n                continue;"),
    ]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc to build mutated copies")
def test_mutants_fail_a_case():
    """Every plausible slip in the restated search changes the record of at
    least one hand-derived case."""
    src = open(ORACLE_SRC).read()
    hdr = os.path.join(os.path.dirname(ORACLE_SRC), "pump_oracle.h")
    survivors = []
    with tempfile.TemporaryDirectory() as tmp:
        shutil.copy(hdr, tmp)
        for k, (name, old, new) in enumerate(_mutants()):
            assert src.count(old) == 1, f"mutation site not unique: {name}"
            c = os.path.join(tmp, f"m{k}.c")
            so = os.path.join(tmp, f"libm{k}.so")
            with open(c, "w") as f:
                f.write(src.replace(old, new))
            subprocess.run(["gcc", "-O1", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off",
                            "-fno-fast-math", c, "-o", so, "-lm"], check=True)
            lib = oracle.load_from(so)
            caught = []
            for cname, case, over, exp in CASES:
                try:
                    check(run(build_case(case), params(**over), lib=lib), exp)
                except AssertionError:
                    caught.append(cname)
            if not caught:
                survivors.append(name)
    assert not survivors, f"mutants that pass every case: {survivors}"
