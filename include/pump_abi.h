/*
 * pump_abi.h -- C ABI of the MI355X (gfx950) implementation of PUMP's
 * roadmap-edge and collision-probability path.
 *
 * Every entry point replaces one reference kernel launch or host function
 * (cited file:line in the reference StanfordASL/PUMP) and keeps its argument
 * meaning and array layouts:
 *   samples[i*6 + d]                       (roadMap.cuh:16)
 *   obstacles[o*12 + {lo 0..5, hi 6..11}]  (mainPUMP.cu:214-221)
 *   discMotions[(wp)*6 + d], wp = e*W + k  (collisionCheck.cu:20-23)
 *   as[(wp*maxHS + h)*3 + d], bs[wp*maxHS + h], countStoredHSs[wp]
 *                                          (hsmc.cu:49,86-89; PUMP.cu:642-648)
 *   offsetsTime[(t*nMC + p)*3 + d]         (hardCoded.cuh:18-19)
 *   xcomb[(p*Tmax + t)*3 + d]              (hardCoded.cuh:15-16)
 *   validParticles[path*nP + p] (1 byte)   (PUMP.cu:630,654)
 *
 * Conventions (replacing the reference's print-and-continue / exit(1)):
 *   - device pointers are caller-allocated, as in the reference;
 *   - every function returns 0 (PUMP_OK) or a negative pump_status;
 *   - work is enqueued on `stream` (a hipStream_t; NULL = default stream);
 *     functions that return a host scalar synchronise that stream;
 *   - all sizes are runtime arguments; pair/edge/waypoint counts are 64-bit;
 *   - a variable number of waypoints per edge is given by wp_offsets[E+1]
 *     (int64, exclusive prefix of W_e); pass NULL to use the reference's
 *     fixed count W = waypoints_count for every edge.
 * Nothing here prints or exits.  No global state: the reference's globals
 * offsetMult / maxd2 (mainPUMP.cu:62-63) are arguments.
 */
#ifndef PUMP_ABI_H
#define PUMP_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *pump_stream_t; /* hipStream_t */

enum pump_status {
    PUMP_OK = 0,
    PUMP_E_INVALID = -1,  /* null pointer / bad size */
    PUMP_E_HIP = -2,      /* HIP runtime error (launch or copy) */
    PUMP_E_CAPACITY = -3, /* a caller buffer is too small */
    PUMP_E_NOMEM = -4,    /* device allocation failed */
    PUMP_E_RANGE = -5     /* an argument is outside the supported range */
};

const char *pump_status_string(int status);
int pump_abi_version(void);
/* The HIP error and call site behind the last PUMP_E_HIP on this thread. */
const char *pump_last_error(void);

/* ---------------------------------------------------------------------------
 * Roadmap precompute ("roadMap/precomp")
 * ------------------------------------------------------------------------ */

/* fillCoptsTopts -- 2pBVP.cu:133-150, launched at mainPUMP.cu:293-299.
 * For every ordered pair i != j: topts[idx] = toptBisection(x_i, x_j, tmax),
 * copts[idx] = cost(topts[idx], x_i, x_j), idx = i*(num-1) + j - [i<j].
 * Arithmetic is the reference's host semantics (cost evaluated in double and
 * rounded to float; float bisection). */
int pump_fill_copts_topts(const float *d_samples, int num, float tmax, float *d_copts,
                          float *d_topts, pump_stream_t stream);

/* Chosen form (no reference counterpart): FP64 Newton polish of each bisected
 * arrival time on the stationarity quartic of cost; d_topts64[idx]. */
int pump_polish_topts(const float *d_samples, int num, float tmax, const float *d_topts,
                      double *d_topts64, pump_stream_t stream);

/* rn = sort(copts)[rn_idx] -- mainPUMP.cu:303-304 (thrust::sort + index).
 * Exact k-th order statistic by device radix select; writes *rn_host and
 * synchronises `stream`.  Workspace of pump_select_rn_workspace() bytes. */
size_t pump_select_rn_workspace(int64_t n_pairs);
int pump_select_rn(const float *d_copts, int64_t n_pairs, int64_t rn_idx, float *rn_host,
                   void *d_workspace, size_t workspace_bytes, pump_stream_t stream);

/* rnIdx -- mainPUMP.cu:262: (int)(pairs * rnPerc) evaluated in float. */
int64_t pump_rn_index(int64_t num, float rn_perc);

/* Edge numbering -- mainPUMP.cu:327-346: edges are the pairs with copts < rn
 * (strict), numbered row-major (source i ascending, then destination j).
 * pump_count_edges writes d_row_start[num+1] (out-edges of node i are
 * [row_start[i], row_start[i+1])) and *n_edges_host (synchronises).
 * Workspace: pump_scan_workspace(num+1) bytes. */
size_t pump_scan_workspace(int64_t n);
int pump_count_edges(const float *d_copts, int num, float rn, int64_t *d_row_start,
                     int64_t *n_edges_host, void *d_workspace, size_t workspace_bytes,
                     pump_stream_t stream);
/* coptsEdge/toptsEdge (mainPUMP.cu:334-335) plus the edge endpoints; the
 * reference's nnIdxs[j*NUM+i] = e is the inverse of (edge_src, edge_dst). */
int pump_fill_edges(const float *d_copts, const float *d_topts, int num, float rn,
                    const int64_t *d_row_start, int *d_edge_src, int *d_edge_dst,
                    float *d_copts_edge, float *d_topts_edge, pump_stream_t stream);
/* maxNNSize (mainPUMP.cu:362-368) and the padded tables nnGoEdges,
 * nnComeEdges, distancesCome [num x max_nn] (mainPUMP.cu:371-392).
 * pump_max_nn_size synchronises.  d_nn_idxs (dense [num*num], default -3,
 * mainPUMP.cu:315,338) is optional (NULL to skip). */
int pump_max_nn_size(const int64_t *d_row_start, const int *d_edge_dst, int64_t n_edges, int num,
                     int *d_in_degree, int *max_nn_host, pump_stream_t stream);
int pump_nn_tables(const int64_t *d_row_start, const int *d_edge_src, const int *d_edge_dst,
                   const float *d_copts_edge, int64_t n_edges, int num, int max_nn,
                   int *d_nn_go_edges, int *d_nn_come_edges, float *d_distances_come,
                   int *d_nn_idxs, pump_stream_t stream);
/* h[i] = cost(toptBisection(x_i, goal, tmax), x_i, goal) for i < num-1,
 * h[num-1] = 0 -- mainPUMP.cu:351-355.  goal_host: 6 floats (host). */
int pump_heuristic(const float *d_samples, int num, const float *goal_host, float tmax, float *d_h,
                   pump_stream_t stream);

/* Waypoints per edge: W_e = w_fixed if w_fixed > 0 (the reference's
 * numDisc+1, mainPUMP.cu:57), else clamp(ceil(tau_e/wp_dt), 2, w_cap).
 * Every edge entry point takes 2 <= W_e < 2^23 (the reference's numDisc >= 1;
 * an edge with fewer waypoints has no defined discretisation).
 * Writes d_wp_offsets[n_edges+1] and *total_host (synchronises).
 * Workspace: pump_scan_workspace(n_edges+1). */
int pump_waypoint_offsets(const float *d_topts_edge, int64_t n_edges, int w_fixed, float wp_dt,
                          int w_cap, int64_t *d_wp_offsets, int64_t *total_host,
                          void *d_workspace, size_t workspace_bytes, pump_stream_t stream);

/* Edge discretization -- the host loop mainPUMP.cu:327-346 calling
 * findDiscretizedPath (2pBVP.cu:53-69) per edge, moved to the device.
 * The arrival time is taken from d_topts_edge: it is the same toptBisection
 * of the same states that the reference re-solves on the host. */
int pump_discretize_edges(const float *d_samples, const int *d_edge_src, const int *d_edge_dst,
                          const float *d_topts_edge, int64_t n_edges, const int64_t *d_wp_offsets,
                          int waypoints_count, float *d_disc_motions, pump_stream_t stream);

/* ---------------------------------------------------------------------------
 * Collision checking ("collisionCheck")
 * ------------------------------------------------------------------------ */

/* sampleFree -- sampler.cu:61-88, launched at mainPUMP.cu:430-436. */
int pump_sample_free(const float *d_obstacles, int n_obs, const float *d_samples, int num,
                     uint8_t *d_is_free_samples, pump_stream_t stream);

/* freeEdges -- collisionCheck.cu:5-37 (signature collisionCheck.cu:6-8),
 * launched at mainPUMP.cu:444-452.  The reference's arguments in its order:
 * num_disc segments per edge, edge e's num_disc + 1 waypoints at rows
 * e*(num_disc+1) .. of discMotions (collisionCheck.cu:20-23).  samples and
 * isFreeSamples are accepted and unused, as in the reference; debugOutput is
 * dropped and the stream added.  num_disc = 0 checks no segment (every edge
 * free), as the reference's loop does. */
int pump_free_edges(const float *d_obstacles, int n_obs, const float *d_samples,
                    const uint8_t *d_is_free_samples, int num_disc, const float *d_disc_motions,
                    uint8_t *d_is_free_edges, int64_t n_edges, pump_stream_t stream);

/* freeEdges over a variable number of waypoints per edge: edge e's
 * waypoints are rows wp_offsets[e] .. wp_offsets[e+1]-1 (at least one). */
int pump_free_edges_var(const float *d_obstacles, int n_obs, const int64_t *d_wp_offsets,
                        const float *d_disc_motions, uint8_t *d_is_free_edges, int64_t n_edges,
                        pump_stream_t stream);

/* Fused edge pass (kernel b): waypoint generation (findDiscretizedPath) and
 * the segment/AABB test (freeEdges) in one launch; writes discMotions and
 * isFreeEdges, identical to pump_discretize_edges + pump_free_edges. */
int pump_edges_discretize_check(const float *d_samples, const int *d_edge_src,
                                const int *d_edge_dst, const float *d_topts_edge,
                                int64_t n_edges, const int64_t *d_wp_offsets, int waypoints_count,
                                const float *d_obstacles, int n_obs, float *d_disc_motions,
                                uint8_t *d_is_free_edges, pump_stream_t stream);

/* The fused edge pass with FP32 waypoints (chosen form): the same call as
 * pump_edges_discretize_check, with every interior waypoint evaluated in
 * FP32 from per-edge FP32 coefficients of the cubic Hermite segment
 * (Horner's rule at the reference's t) instead of the reference's FP64
 * expression tree.  Each position component is within eps = 12 u S of the
 * reference's waypoint (u = 2^-24, S = |p0| + |v0| tau + |c2| tau^2 +
 * |c3| tau^3 of its axis); end states are exact copies.  The flags are the
 * reference's bit for bit: a segment whose outcome against a box could
 * change within those bounds (counting the reference's own FP32 roundings)
 * is re-tested with the exact FP64 waypoints, counted in *d_retests
 * (optional). */
int pump_edges_discretize_check_fp32(const float *d_samples, const int *d_edge_src,
                                     const int *d_edge_dst, const float *d_topts_edge,
                                     int64_t n_edges, const int64_t *d_wp_offsets,
                                     int waypoints_count, const float *d_obstacles, int n_obs,
                                     float *d_disc_motions, uint8_t *d_is_free_edges,
                                     unsigned long long *d_retests, pump_stream_t stream);

/* ---------------------------------------------------------------------------
 * Half-spaces and HSMC ("hsmc")
 * ------------------------------------------------------------------------ */

/* setupAsBs -- hsmc.cu:201-211: as = -1, bs = 1111 (b >= 1000: empty). */
int pump_setup_as_bs(int64_t hs_count, float *d_bs, float *d_as, pump_stream_t stream);

/* maxd2 = 0.1*offsetMult^2 (double, narrowed to float) -- mainPUMP.cu:63,201 */
float pump_maxd2(float offset_mult);

/* cacheHalfspaces -- hsmc.cu:10-99 (+ bvlsShortcut, bvls.cu:259-273),
 * launched at mainPUMP.cu:484-498.  Any obstacle count is supported (the
 * reference stages only 1024 floats of obstacles, hsmc.cu:20-22).
 * countStoredHSs of waypoints on blocked edges is written as 0 (the
 * reference leaves it unwritten).  write_sentinels != 0 additionally writes
 * the empty sentinel into every unused slot, which makes a preceding
 * pump_setup_as_bs unnecessary (same final contents).  d_edge_hs_mask, if
 * non-NULL (n_edges words), receives per edge a waypoint summary: bit k
 * (k < 63) set iff waypoint k stores a half-space with b not >= 1000, bit 63
 * iff some waypoint >= 63 does.  It describes d_bs only when the unused
 * slots hold the empty sentinel (pump_setup_as_bs or write_sentinels). */
int pump_cache_halfspaces(int64_t n_edges, const float *d_disc_motions,
                          const uint8_t *d_is_free_edges, int n_obs, int max_hs,
                          const float *d_obstacles, int waypoints_count,
                          const int64_t *d_wp_offsets, float *d_as, float *d_bs,
                          int *d_count_stored, float maxd2, int write_sentinels,
                          uint64_t *d_edge_hs_mask, pump_stream_t stream);

/* setupValidParticles -- PUMP.cu:766-777, launched at mainPUMP.cu:663-668. */
int pump_setup_valid_particles(uint8_t *d_valid, int n_p, int64_t max_path_count,
                               float *d_path_cp, float *d_path_time, pump_stream_t stream);

/* propagateValidParticles (+ ...Many) -- PUMP.cu:613-705, launched at
 * PUMP.cu:94-121.  Child path path_base+w inherits validity from
 * wf_prev[w] and ANDs (offset.a)*offsetMult < b over every stored
 * half-space of every waypoint of edge wf_edge[w].  tmax_steps is the row
 * count of offsets_time; time indices past it are clamped (the reference
 * reads past its table) and counted into *d_clamped if non-NULL. */
int pump_propagate_valid_particles(int num_new_paths, int waypoints_count,
                                   const int64_t *d_wp_offsets, const int *d_wf_edge,
                                   const int *d_wf_prev, int64_t path_base, const float *d_as,
                                   const float *d_bs, const int *d_count_stored,
                                   const float *d_times, const float *d_topts, float dt,
                                   int max_hs, const float *d_offsets_time, int tmax_steps,
                                   float offset_mult, int num_mc_particles, uint8_t *d_valid,
                                   int n_p, unsigned long long *d_clamped, pump_stream_t stream);

/* calculateCP_HSMC -- PUMP.cu:745-758, launched at PUMP.cu:136-141. */
int pump_calculate_cp_hsmc(int num_new_paths, int64_t path_base, float *d_path_cp,
                           const uint8_t *d_valid, int n_p, pump_stream_t stream);

/* Propagate + CP with separate parent/child validity arrays (child rows
 * dst_base+w of d_valid_dst).  Same results as the two calls above; used by
 * the search consumer so that compaction needs no aliasing copies
 * (PUMP.cu:236-237 issues one D2D memcpy per surviving path).  With a
 * workspace of pump_propagate_workspace(num_new_paths, tmax_steps) bytes,
 * large wavefronts run in a batched form (classify, copy, and the
 * half-space tests binned by time index); pass NULL, 0 for a single launch.  d_edge_hs_mask: the per-edge waypoint
 * summary written by pump_cache_halfspaces for these d_bs, or NULL (then
 * every waypoint's slots are scanned, as the reference does). */
size_t pump_propagate_workspace(int num_new_paths, int tmax_steps);
/* A workspace whose binned form holds marked_pairs (path, waypoint) pairs
 * with a stored half-space (at least the default 4 per path + 65536): a
 * wavefront with more such pairs than its workspace holds runs the slower
 * group form (same results).  Any larger workspace passed to the calls
 * below is used the same way.  The capacity stops at 2^30 pairs.  No
 * counterpart in the reference. */
size_t pump_propagate_workspace_pairs(int num_new_paths, int tmax_steps, int64_t marked_pairs);
int pump_propagate_cp(int num_new_paths, int waypoints_count, const int64_t *d_wp_offsets,
                      const int *d_wf_edge, const int *d_wf_prev, const float *d_as,
                      const float *d_bs, const uint64_t *d_edge_hs_mask, const float *d_times,
                      const float *d_topts, float dt, int max_hs, const float *d_offsets_time, int tmax_steps, float offset_mult,
                      int num_mc_particles, const uint8_t *d_valid_src, uint8_t *d_valid_dst,
                      int64_t dst_base, int n_p, float *d_path_cp, int64_t cp_base,
                      unsigned long long *d_clamped, void *d_workspace, size_t workspace_bytes,
                      pump_stream_t stream);

/* Compaction of surviving paths -- the per-path D2D copies of
 * PUMP.cu:236-237 as one gather: dst row dst_base+k = src row src_rows[k]. */
int pump_gather_rows(const uint8_t *d_src, const int *d_src_rows, int n_rows, uint8_t *d_dst,
                     int64_t dst_base, int64_t row_bytes, pump_stream_t stream);

/* Chosen form: validity as bit masks (bit p%64 of word path*ceil(nP/64) +
 * p/64), child masks written by ballots; CP from popcounts.  Identical
 * per-particle decisions to pump_propagate_cp.  precision: 0 = FP32 (the
 * reference's operation order); 1 = split-BF16: the half-space tests of the
 * batched form's time-binned pass run as bf16 hi/lo products on MFMA with
 * FP32 accumulation, and every test whose |offset.a - b/offsetMult| lies
 * within the margin band 6e-5*max|o|*sum|a| + 2.5e-5*|b/offsetMult| is redone
 * in FP32 in the reference's order (counted into *d_band_count when
 * non-NULL), so the decisions stay identical.  precision 1 needs the
 * particle-vector table of pump_split_offsets (d_offsets_split); other
 * stages, offsetMult <= 0 or max_hs > 8 run in FP32. */
int pump_propagate_cp_bits(int num_new_paths, int waypoints_count, const int64_t *d_wp_offsets,
                           const int *d_wf_edge, const int *d_wf_prev, const float *d_as,
                           const float *d_bs, const uint64_t *d_edge_hs_mask,
                           const float *d_times, const float *d_topts, float dt, int max_hs,
                           const float *d_offsets_time, int tmax_steps, float offset_mult,
                           int num_mc_particles, const uint64_t *d_valid_src,
                           uint64_t *d_valid_dst, int64_t dst_base, int n_p, float *d_path_cp,
                           int64_t cp_base, int precision, const void *d_offsets_split,
                           unsigned long long *d_band_count, unsigned long long *d_clamped,
                           void *d_workspace, size_t workspace_bytes, pump_stream_t stream);

/* Particle-vector table of the split-BF16 test: for each (time index,
 * particle) of offsetsTime, 16 bf16 (bf16 hi/lo of the three offsets in MFMA
 * operand order, and max|o| rounded up).  Build once per noise table;
 * d_split must be 16-byte aligned and pump_split_offsets_bytes() long. */
size_t pump_split_offsets_bytes(int tmax_steps, int n_mc);
int pump_split_offsets(const float *d_offsets_time, int tmax_steps, int n_mc, void *d_split,
                       pump_stream_t stream);

/* The batched step with copy-on-write child rows (byte or bit validity):
 * the rows of new paths that meet no stored half-space equal their parents'
 * rows and are not written.  d_row_src[w] = w when child w's row was written
 * (row dst_base + w of d_valid_dst), -1 - prev when it is parent prev's row
 * (of d_valid_src); such a child's CP is d_parent_cp[prev] (the same count,
 * the same float).  Other arguments as pump_propagate_cp(_bits); precision 1
 * needs bit validity.  The rows a caller keeps are materialised with
 * pump_gather_rows_cow -- the reference copies its survivors' rows at that
 * point anyway (PUMP.cu:236-237). */
int pump_propagate_cp_cow(int validity_bits, int num_new_paths, int waypoints_count,
                          const int64_t *d_wp_offsets, const int *d_wf_edge, const int *d_wf_prev,
                          const float *d_as, const float *d_bs, const uint64_t *d_edge_hs_mask,
                          const float *d_times, const float *d_topts, float dt, int max_hs,
                          const float *d_offsets_time, int tmax_steps, float offset_mult,
                          int num_mc_particles, const void *d_valid_src, const float *d_parent_cp,
                          void *d_valid_dst, int64_t dst_base, int *d_row_src, int n_p,
                          float *d_path_cp, int64_t cp_base, int precision,
                          const void *d_offsets_split, unsigned long long *d_band_count,
                          unsigned long long *d_clamped, void *d_workspace,
                          size_t workspace_bytes, pump_stream_t stream);

/* Autotuner of the batched MC pass (no counterpart in the reference, whose
 * step has no launch shape to choose): runs pump_propagate_cp_cow with the
 * same arguments once per candidate count of particles per wave of the
 * half-space test (64, 128, 256, 512), times three repetitions of each on
 * the stream, keeps the fastest for the process under the configuration
 * (validity form, precision, n_p, max_hs) -- every later batched pass of
 * that configuration uses it -- and finishes with one call in the chosen
 * shape, whose outputs (and counters) are the call's results.  The
 * candidates' outputs are identical; *particles_per_wave receives the
 * choice.  Synchronizes the stream. */
int pump_hsmc_autotune_cow(int validity_bits, int num_new_paths, int waypoints_count,
                           const int64_t *d_wp_offsets, const int *d_wf_edge,
                           const int *d_wf_prev, const float *d_as, const float *d_bs,
                           const uint64_t *d_edge_hs_mask, const float *d_times,
                           const float *d_topts, float dt, int max_hs,
                           const float *d_offsets_time, int tmax_steps, float offset_mult,
                           int num_mc_particles, const void *d_valid_src,
                           const float *d_parent_cp, void *d_valid_dst, int64_t dst_base,
                           int *d_row_src, int n_p, float *d_path_cp, int64_t cp_base,
                           int precision, const void *d_offsets_split,
                           unsigned long long *d_band_count, unsigned long long *d_clamped,
                           void *d_workspace, size_t workspace_bytes, pump_stream_t stream,
                           int *particles_per_wave);

/* Particles per wave the batched MC pass uses for a configuration: the
 * tuned choice, or 256 where none was made. */
int pump_hsmc_particles_per_wave(int validity_bits, int precision, int n_p, int max_hs);

/* Sets the choice for a configuration by hand (64, 128, 256 or 512; 0
 * forgets it).  No counterpart in the reference. */
int pump_hsmc_set_particles_per_wave(int validity_bits, int precision, int n_p, int max_hs,
                                     int particles_per_wave);

/* Child sel[j] (j < n_sel; sel NULL: j itself) of a copy-on-write step to
 * row dst_base + j of d_dst, from d_children (rows relative to the step's
 * dst_base) or d_parents as d_row_src says; byte rows (validity_bits 0) are
 * normalised to 0/1.  Parent rows must not lie in the destination rows. */
int pump_gather_rows_cow(const void *d_children, const void *d_parents, const int *d_row_src,
                         const int *d_sel, int n_sel, void *d_dst, int64_t dst_base,
                         int64_t row_bytes, int validity_bits, pump_stream_t stream);

/* pump_propagate_cp_bits without a workspace argument and without any
 * device allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The batched form's workspace is a
 * fixed 64 MiB device array held by each GPU the library runs on until the
 * process ends; a wavefront whose pump_propagate_workspace(num_new_paths,
 * tmax_steps) exceeds it (566,564 paths at 2001 time indices) returns
 * PUMP_E_CAPACITY before anything runs.  All calls on one GPU share that
 * GPU's array: calls that overlap on one GPU race (calls on one stream are
 * ordered).  Same results as pump_propagate_cp_bits. */
int pump_propagate_cp_bits_noalloc(
    int num_new_paths, int waypoints_count, const int64_t *d_wp_offsets, const int *d_wf_edge,
    const int *d_wf_prev, const float *d_as, const float *d_bs, const uint64_t *d_edge_hs_mask,
    const float *d_times, const float *d_topts, float dt, int max_hs,
    const float *d_offsets_time, int tmax_steps, float offset_mult, int num_mc_particles,
    const uint64_t *d_valid_src, uint64_t *d_valid_dst, int64_t dst_base, int n_p,
    float *d_path_cp, int64_t cp_base, int precision, const void *d_offsets_split,
    unsigned long long *d_band_count, unsigned long long *d_clamped, pump_stream_t stream);

/* Initial bit masks for pump_propagate_cp_bits: all n_p particles valid in
 * each of n_paths rows (the bit form of setupValidParticles). */
int pump_setup_valid_bits(uint64_t *d_valid, int n_p, int64_t n_paths, pump_stream_t stream);

/* ---------------------------------------------------------------------------
 * Full Monte Carlo CP ("collisionProbability")
 * ------------------------------------------------------------------------ */

/* collisionProbability + MCCP -- collisionProbability.cu:4-92.  Writes
 * cp = hits/n_mc to *cp_host (synchronises).  Allocates a 4-byte hit
 * counter per call, as the reference allocates its per-call scratch. */
int pump_collision_probability(const float *d_obstacles, int n_obs, const float *d_xcomb,
                               int tmax_steps, int n_mc, float offset_mult, const float *d_path,
                               int T, float *cp_host, pump_stream_t stream);

/* Batched form: n_paths paths, path k at d_paths + k*path_stride_rows*6
 * with T = d_T[k]; hit counts to d_hits[k] (caller-zeroed or zeroed here
 * when zero_hits != 0). */
int pump_collision_probability_many(const float *d_obstacles, int n_obs, const float *d_xcomb,
                                    int tmax_steps, int n_mc, float offset_mult,
                                    const float *d_paths, int64_t path_stride_rows,
                                    const int *d_T, int n_paths, int *d_hits, int zero_hits,
                                    pump_stream_t stream);

/* ---------------------------------------------------------------------------
 * Group Marching Tree and Euclidean neighbours (not in the reference's
 * build; listed API surface)
 * ------------------------------------------------------------------------ */

/* calculateNumberNN (precomp.cu:3-30): d_nn_sizes[i] = #{j != i :
 * |x_j - x_i|^2 < r2} over all 6 state dimensions. */
int pump_precomp_nn_count(const float *d_samples, int num, float r2, int *d_nn_sizes,
                          pump_stream_t stream);
/* setupPrecompNN + precompNN (precomp.cu:32-76): d_nn_edges/d_distances
 * [num x nn_size], neighbours in sample order with their distance, -1 / 0
 * past the last one (at most nn_size kept per node). */
int pump_precomp_nn(const float *d_samples, int num, float r2, int nn_size, float *d_distances,
                    int *d_nn_edges, pump_stream_t stream);

/* GMTwavefront (GMT.cu:3-115, lambda = 1) over the padded neighbour tables
 * of pump_nn_tables (or precomp): writes the tree's costs and parents
 * (d_costs, d_edges [num]; -11 / -2 on blocked samples, 0 / -1 where not
 * reached), the goal's cost (0 if not reached) and the iterations.  The
 * connecting edge from parent to node is found through d_nn_idxs (dense
 * [to*num + from], mainPUMP.cu:338) when given, else by the CSR rows
 * d_row_start / d_edge_dst; its waypoints are d_disc at d_wp_offsets (or
 * numDisc+1 per edge when NULL).  Node 0 starts the wavefront
 * (GMT.cu:128-134).  Workspace: pump_gmt_workspace(num) bytes.
 * Synchronises once per iteration. */
size_t pump_gmt_workspace(int num);
int pump_gmt_wavefront(const float *d_obstacles, int n_obs, const float *d_distances_come,
                       const int *d_nn_go_edges, const int *d_nn_come_edges, int nn_size,
                       const float *d_disc, const int64_t *d_wp_offsets, int num_disc,
                       const int *d_nn_idxs, const int64_t *d_row_start, const int *d_edge_dst,
                       const uint8_t *d_is_free_samples, int num, float *d_costs, int *d_edges,
                       int goal_idx, void *d_ws, size_t ws_bytes, float *cost_goal_host,
                       int *itrs_host, pump_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* PUMP_ABI_H */
