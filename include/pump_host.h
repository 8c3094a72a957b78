/*
 * pump_host.h -- host-side C++ pipeline above the C ABI (pump_abi.h):
 * problem and noise-table generation, the roadmap precompute of
 * mainPUMP.cu:248-504,621-681 and the PUMP search consumer of
 * PUMP.cu:3-405, driven entirely through the pump_* device entry points.
 * Exposed as C for the Python tests and the benchmark.
 */
#ifndef PUMP_HOST_H
#define PUMP_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "pump_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- device memory, streams and timers ----------------------------------
 * Minimal helpers so that a host program can drive the C ABI with only this
 * library's HIP runtime in the process. */
void *pump_dev_alloc(size_t bytes); /* NULL on failure (see pump_last_error) */
int pump_dev_free(void *p);
int pump_dev_upload(void *dst, const void *src, size_t bytes);
int pump_dev_download(void *dst, const void *src, size_t bytes);
int pump_dev_memset(void *dst, int value, size_t bytes);
int pump_dev_copy(void *dst, const void *src, size_t bytes); /* device to device */
int pump_dev_sync(void);
int pump_dev_count(void);
int pump_dev_name(char *buf, int len);
int pump_stream_create(pump_stream_t *s);
int pump_stream_destroy(pump_stream_t s);
int pump_stream_sync(pump_stream_t s);
typedef struct pump_timer pump_timer; /* hipEvent pair on one stream */
pump_timer *pump_timer_create(pump_stream_t s);
int pump_timer_start(pump_timer *t);
float pump_timer_stop(pump_timer *t); /* ms between start and stop, -1 on error */
void pump_timer_destroy(pump_timer *t);

/* ---- synthetic problems (counter-based Philox4x32-10 draws) ------------- */
enum pump_obstacle_kind { PUMP_OBS_BOXES = 0, PUMP_OBS_WALLS = 1 };

typedef struct pump_problem_spec {
    int num;             /* samples; sample 0 = start, sample num-1 = goal */
    uint64_t seed;
    float lo[6], hi[6];  /* sampling bounds per state dimension */
    float start[6], goal[6];
    int n_obs;
    int obstacle_kind;   /* pump_obstacle_kind */
    float centre_lo, centre_hi; /* box centres uniform in [centre_lo, centre_hi]^3 */
    float side_lo, side_hi;     /* box sides uniform (boxes) */
    float wall_thickness, wall_long_lo, wall_long_hi, wall_short_lo, wall_short_hi;
    float vel_span;      /* velocity bounds of every box: [-vel_span, vel_span] */
} pump_problem_spec;

/* Philox4x32-10 (Salmon et al., SC'11): 4 output words for counter ctr[4]
 * and key key[2]. */
void pump_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

int pump_gen_problem(const pump_problem_spec *spec, float *samples_host, float *obstacles_host);

/* ---- the reference's problem files and sampler -------------------------
 * pump_read_problem: the problem file of mainPUMP.cu:152-230 (the
 * reference's input.txt / indoor.txt): comma-separated init, goal, lo, hi
 * (6 values each), offsetMult, the obstacle count, then 12 values per
 * obstacle (lowest corner, highest corner); the rest of the file is
 * ignored.  Values parse as `stream >> float` (strtof).  A missing or
 * unreadable value returns PUMP_E_INVALID (the reference keeps 0 there);
 * more than obs_cap obstacles returns PUMP_E_CAPACITY.
 * pump_halton_samples: createSamplesHalton (sampler.cu:20-57), NUM samples
 * from the Halton bases bases[d + skip] (2, 3, 5, ...), sample 0 = init,
 * sample num-1 = goal. */
typedef struct pump_problem_file {
    float init[6], goal[6], lo[6], hi[6];
    float offset_mult;
    int n_obs;
} pump_problem_file;

int pump_read_problem(const char *path, pump_problem_file *out, float *obstacles_host,
                      int obs_cap);
int pump_halton_samples(int num, int skip, const float *init, const float *goal, const float *lo,
                        const float *hi, float *samples_host);

/* ---- noise tables ------------------------------------------------------
 * Closed-loop double-integrator offsets from the nominal path: per axis
 * x_{t+1} = A_cl x_t + L z_t with A_cl the dt-discretised double integrator
 * under u = -kp p - kd v, Q = q [[dt^3/3, dt^2/2], [dt^2/2, dt]] per axis,
 * L = chol(Q) of the 6x6 Q in FP64, z_t ~ N(0, I) from Philox + Box-Muller,
 * x_0 = 0.  Positions are stored (unit scale; offsetMult scales them at
 * use) in both layouts: offsets_time[(t*n_mc + p)*3 + d] and
 * xcomb[(p*tmax_steps + t)*3 + d].  sigma_pos (optional,
 * [tmax_steps*3*3]) receives the FP32-propagated position covariance. */
typedef struct pump_noise_spec {
    int n_mc, tmax_steps;
    float dt;
    double kp, kd, q;
    uint64_t seed;
} pump_noise_spec;

void pump_noise_defaults(pump_noise_spec *spec);
int pump_make_noise(const pump_noise_spec *spec, float *offsets_time_host, float *xcomb_host,
                    float *sigma_pos_host);
/* The same tables generated on the device (stream-ordered): d_chol [36]
 * receives the FP64 Cholesky factor L of Q (required, it also carries L to
 * the particle kernel); d_offsets_time / d_xcomb / d_sigma_pos as above,
 * NULL to skip.  L and sigma_pos are bit-identical to the host's; the tables
 * differ only where the device's FP64 log/sin/cos round differently. */
int pump_make_noise_device(const pump_noise_spec *spec, float *d_offsets_time, float *d_xcomb,
                           float *d_sigma_pos, double *d_chol, pump_stream_t stream);

/* ---- roadmap precompute + search --------------------------------------- */
typedef struct pump_params {
    float tmax;          /* 2pBVP upper bracket (mainPUMP.cu:266) */
    float rn_perc;       /* neighbour fraction (mainPUMP.cu:261) */
    int w_fixed;         /* waypoints per edge (numDisc+1), or <= 0 for ceil(tau/wp_dt) */
    float wp_dt;
    int w_cap;
    float offset_mult;   /* noise multiplier (mainPUMP.cu:62) */
    int max_hs;          /* min(nObs, 5) in the reference (mainPUMP.cu:241) */
    float dt;            /* noise table time step (mainPUMP.cu:60) */
    int n_p;             /* HSMC particles per path */
    float lambda;        /* mainPUMP.cu:66 */
    int num_buckets;     /* mainPUMP.cu:67 */
    float eps_cost, eps_cp;
    float cp_target, cp_factor;
    int max_itrs;        /* PUMP.cu:34 */
    int64_t max_path_count; /* <= 0: 2*NUM*sqrt(NUM) (mainPUMP.cu:623) */
    int validity_bits;   /* 0: byte validity (reference layout), 1: bit masks */
    int precision;       /* 0: FP32 half-space products, 1: split-BF16 MFMA */
    int fused_edges;     /* 1: pump_edges_discretize_check, 0: discretize + freeEdges */
    int keep_pairs;      /* keep copts/topts on the device after the build */
    int polish;          /* also compute FP64-polished arrival times (chosen form) */
    int path_cap;        /* rows of a candidate path for full MC */
    int mccp_batch;      /* goal candidates evaluated per full-MC launch */
    int search_device;   /* 1: search bookkeeping on the device, 0: on the host */
} pump_params;

void pump_params_defaults(pump_params *p);

typedef struct pump_roadmap pump_roadmap;

typedef struct pump_roadmap_info {
    int num, n_obs, max_nn, max_hs, n_mc, tmax_steps;
    int64_t n_pairs, rn_idx, n_edges, n_waypoints, n_free_edges;
    float rn, maxd2;
    /* stage times (ms, hipEvent) of the last build */
    float ms_pairs, ms_select, ms_edges_build, ms_nn, ms_sample_free, ms_edge_pass, ms_hs,
        ms_total;
    /* bisection steps over all pairs of the last build (each step is two
     * cost evaluations); 0 below 65,536 pairs, where they are not counted */
    int64_t pair_steps;
} pump_roadmap_info;

int pump_roadmap_build(const float *samples_host, int num, const float *obstacles_host, int n_obs,
                       const float *goal_host, const float *offsets_time_host,
                       const float *xcomb_host, int n_mc, int tmax_steps, const pump_params *prm,
                       pump_roadmap **out);
int pump_roadmap_info_get(const pump_roadmap *rm, pump_roadmap_info *info);
void pump_roadmap_free(pump_roadmap *rm);

/* A roadmap from explicit tables instead of a precompute: the arrays the
 * search consumer reads (PUMP.cu:3-405; mainPUMP.cu:313-504 builds them), in
 * the reference's numbering (edges row-major by source, destinations
 * ascending within a row, out-edge nn of node i = edge row_start[i] + nn).
 * Sample 0 is the init, sample num-1 the goal.  Half-space slots are
 * as[(wp*max_hs + h)*3 + d], bs[wp*max_hs + h] with wp the global waypoint
 * index of wp_offsets; b >= 1000 is the empty sentinel.  The per-edge
 * waypoint summary is derived from bs.  Waypoint rows are not held, so
 * pump_roadmap_copy_waypoints and the edge passes do not apply. */
typedef struct pump_roadmap_tables {
    int num;
    const float *samples;         /* [num*6] */
    const float *obstacles;       /* [n_obs*12] */
    int n_obs;
    int64_t n_edges;
    const int64_t *row_start;     /* [num+1] */
    const int *edge_dst;          /* [E] */
    const float *copts_edge, *topts_edge; /* [E] */
    const float *h;               /* [num] */
    const uint8_t *is_free_edges; /* [E] */
    const uint8_t *is_free_samples; /* [num] */
    const int64_t *wp_offsets;    /* [E+1] */
    int max_hs;
    const float *as, *bs;         /* [n_wp*max_hs*3], [n_wp*max_hs] */
    float rn;
} pump_roadmap_tables;

int pump_roadmap_from_tables(const pump_roadmap_tables *t, const float *offsets_time_host,
                             const float *xcomb_host, int n_mc, int tmax_steps,
                             const pump_params *prm, pump_roadmap **out);

/* host copies of the roadmap arrays (NULL pointers are skipped) */
int pump_roadmap_copy_pairs(const pump_roadmap *rm, float *copts, float *topts);
int pump_roadmap_copy_polished(const pump_roadmap *rm, double *topts64);
int pump_roadmap_copy_edges(const pump_roadmap *rm, int64_t *row_start, int *edge_src,
                            int *edge_dst, float *copts_edge, float *topts_edge, float *h);
int pump_roadmap_copy_nn(const pump_roadmap *rm, int *nn_go_edges, int *nn_come_edges,
                         float *distances_come, int *nn_idxs);
int pump_roadmap_copy_flags(const pump_roadmap *rm, int64_t *wp_offsets, uint8_t *is_free_edges,
                            uint8_t *is_free_samples);
int pump_roadmap_copy_waypoints(const pump_roadmap *rm, float *disc_motions);
int pump_roadmap_copy_halfspaces(const pump_roadmap *rm, float *as, float *bs, int *count_stored);

/* device pointers, for callers that drive the entry points themselves */
typedef struct pump_roadmap_dev {
    float *samples, *obstacles, *copts, *topts, *copts_edge, *topts_edge, *h, *disc, *as, *bs;
    float *offsets_time, *xcomb;
    int64_t *row_start, *wp_offsets;
    int *edge_src, *edge_dst, *count_stored;
    uint8_t *is_free_edges, *is_free_samples;
    uint64_t *edge_hs_mask; /* per-edge waypoint summary of the stored half-spaces */
    void *offsets_split;    /* particle vectors of the split-BF16 test */
} pump_roadmap_dev;
int pump_roadmap_device(const pump_roadmap *rm, pump_roadmap_dev *dev);

/* Batched roadmap build: the reference's per-problem precompute
 * (mainPUMP.cu:248-504) run for many independent problems at once.  The
 * reference has no batched form; this is the multi-instance workload of
 * 256 problems with NUM in [1000, 8000].  The pair pass of every instance is one persistent
 * launch, the fused edge pass plus the half-space cache of every instance are
 * two more, each captured in a hipGraph and replayable.  Launch shapes are
 * chosen by timing the candidates on the device when autotune != 0: resident
 * workgroups per CU of the pair pass, edges per workgroup and the obstacle
 * tile (boxes staged in LDS) of the edge pass.  The choice is kept for the
 * process, keyed by the configuration's shape (instance count, pair, edge
 * and waypoint totals, largest box count), and a later build of the same
 * shape reuses it without timing. */
typedef struct pump_batch pump_batch;

typedef struct pump_batch_instance {
    const float *samples;   /* host, [num * 6] */
    int num;
    const float *obstacles; /* host, [n_obs * 12] */
    int n_obs;
    const float *goal;      /* host, [6] */
    pump_params prm;        /* per-instance parameters (fused_edges is implied) */
} pump_batch_instance;

typedef struct pump_batch_info {
    int n_inst;
    int pairs_blocks_per_cu;  /* resident 256-lane workgroups per CU of the pair pass */
    int edges_per_block;      /* edges per workgroup of the edge pass (64, 128, 256) */
    int64_t n_pairs, n_edges, n_waypoints; /* totals over the instances */
    float ms_pairs;           /* one launch of the pair-pass graph */
    float ms_edges;           /* one launch of the edge-pass + half-space graph */
    float ms_autotune;        /* device time spent timing candidates */
    float ms_total;           /* wall time of the build */
    int edge_obs_tile;        /* boxes staged in LDS (with candidate masks) by the edge pass */
    int autotune_cached;      /* bit 0: pair-pass shape, bit 1: edge-pass shape reused from an
                                 earlier build of the same configuration in this process */
    int mc_one_launch;        /* 1: the MC stage (pump_batch_mc_build) runs one launch per
                                 stage for all instances; 0: one step per instance */
    int mc_particles_per_wave; /* the one-launch MC stage's particles per wave (timed) */
} pump_batch_info;

int pump_batch_build(int n_inst, const pump_batch_instance *inst, const float *offsets_time_host,
                     const float *xcomb_host, int n_mc, int tmax_steps, int autotune,
                     pump_batch **out);
int pump_batch_info_get(const pump_batch *b, pump_batch_info *info);
/* the roadmap of instance k, owned by the batch (pump_roadmap_free ignores it) */
int pump_batch_roadmap(pump_batch *b, int k, pump_roadmap **rm);
/* replays a captured stage reps times: 0 = pair pass (needs keep_pairs on
 * every instance), 1 = edge pass + half-space cache, 2 = the MC step of every
 * instance (pump_batch_mc_build); ms = mean per replay */
int pump_batch_replay(pump_batch *b, int stage, int reps, float *ms);

/* The HSMC step (PUMP.cu:94-160, copy-on-write form, bit validity, FP32
 * tests) of every instance, captured as the batch's third graph.  Each
 * instance's wavefront is its free edges in edge order (at most max_paths),
 * child i expanded from parent i % n_parents; the parents are all-valid
 * rows with CP 0 and times uniform in [0, 2) s drawn from seed (splitmix64).
 * The steps of all instances run as one launch per stage (workspace
 * clearing, classification, time-index scan, row copies, entry emission,
 * binned tests, overflow group form, CP), each workgroup serving one
 * instance's share of the stage (pump_batch_info.mc_one_launch = 1), with
 * the particles per wave timed over the whole step before the capture
 * (pump_batch_info.mc_particles_per_wave).  Where an instance's step cannot
 * take that form, the particles per wave of each MC configuration (particle
 * count, half-space slots) are tuned once (pump_hsmc_autotune_cow) on its
 * first instance and the steps fork from the batch stream onto eight side
 * streams of default priority (instance k on stream k mod 8) and join back.
 * ms receives the graph's first launch.  Needs the batch's noise tables. */
int pump_batch_mc_build(pump_batch *b, int n_parents, int max_paths, uint64_t seed, float *ms);
/* the same steps launched eagerly, reps times; ms = mean per pass */
int pump_batch_mc_eager(pump_batch *b, int reps, float *ms);
/* instance k's wavefront and results (the first cap entries; times: the
 * first cap parent times): children's CPs and row sources (-1 - parent for
 * a child that met no half-space) */
int pump_batch_mc_get(pump_batch *b, int k, int cap, int *n_new, int *edge, int *prev,
                      float *times, float *cp, int *row_src);
void pump_batch_free(pump_batch *b);

typedef struct pump_search_result {
    int status;           /* 0 success, 1 no goal path below target, <0 aborted */
    int itrs;
    int64_t num_paths, num_considered;
    int num_goal_paths;
    int soln_path_idx;
    float soln_cost, soln_cp, soln_time, soln_hsmc_cp;
    int soln_len;
    int soln_nodes[256];
    int64_t clamped_time_idx;
    int64_t particle_steps;
    int64_t band_particles; /* split-BF16 margin-band re-tests */
    int mccp_calls;
    float ms_propagate;    /* device time in the HSMC step (hipEvent) */
    float ms_total;        /* wall time of the search */
} pump_search_result;

/* The roadmap keeps the search's device buffers from its first search on
 * (released by pump_roadmap_free); searches on one roadmap must not overlap. */
int pump_search_run(pump_roadmap *rm, const pump_params *prm, pump_search_result *res);

/* The stored path table of the roadmap's last search, rows [0, n) with n <=
 * its num_paths: pathPrev, pathNode, pathCost, pathTime and pathCP of
 * PUMP.cu (path 0 is the init path, prev -2).  NULL outputs are skipped;
 * PUMP_E_RANGE if n exceeds the rows the search stored. */
int pump_search_copy_paths(const pump_roadmap *rm, int64_t n, int *prev, int *node, float *cost,
                           float *time, float *cp);

/* Smoothing of the search's solution path and its controller discretisation
 * (PUMP.cu:406-568).  The solution's node states are blended towards the
 * obstacle-free optimal path's states, alpha bisected on the full-MC CP (at
 * most 30 steps, kSmoothDepth = 4 steps evaluated per batched full-MC
 * launch, with the sequential loop's alphas and decisions); the kept path
 * is the one with CP below the target closest to it.  Outputs: the kept path
 * rows at dt (path.txt; path_cap rows), the last iterate's node states
 * (n_nodes x 6), and the controller discretisation of the last iterate at
 * 0.05 s (discretizedPath.txt; disc_cap rows).  NULL outputs are skipped. */
typedef struct pump_smooth_result {
    int status;           /* 0 smoothed; 1 the search has no solution; <0 a path exceeds the buffers */
    int itrs;             /* bisection steps (full-MC evaluations of the sequential loop) */
    int mccp_launches;    /* batched full-MC launches that evaluated them */
    int n_nodes;          /* states of the solution path */
    int path_rows;        /* rows of the kept path */
    int discretized_rows; /* rows of the controller discretisation */
    float cost_before, cost, cp;
    float ms_total;
    float trace[32][3];   /* alpha, CP, cost of each step */
} pump_smooth_result;

int pump_smooth_run(pump_roadmap *rm, const pump_params *prm, const pump_search_result *sres,
                    pump_smooth_result *out, float *path_host, int path_cap,
                    float *xs_smooth_host, float *discretized_host, int disc_cap);

/* findOptimalPath (2pBVP.cu:5-36) on the host with the library's arithmetic;
 * returns rows written or -1 if more than cap rows would be needed. */
int pump_find_optimal_path(float dt, const float *xs, int n_wp, float *split, int cap,
                           float *total_time);
/* single-pair helpers (host) for tests */
float pump_host_topt(const float *x0, const float *x1, float tmax);
float pump_host_cost(float tau, const float *x0, const float *x1);
/* pathPoint (2pBVP.cu:117-131) with the kernels' arithmetic */
void pump_host_path_point(float t, float tau, const float *x0, const float *x1, float *x);
/* the kernels' FP64 division a/b (reciprocal + residual steps), n pairs */
void pump_host_div_rn(const double *a, const double *b, double *q, int64_t n);
void pump_host_div_sq(const double *a, const float *tau, double *q, int64_t n);
/* the pair pass's row of a pair index: n[i] / d by multiply-high with d's
 * magic number (d >= 1), count dividends */
void pump_host_div_magic(const uint32_t *n, int64_t count, uint32_t d, uint32_t *q);
/* the edge pass's bounding box of an edge's waypoint positions (lo[3], hi[3]) */
void pump_host_edge_bbox(const float *x0, const float *x1, float tau, float *lo, float *hi);

#ifdef __cplusplus
}
#endif
#endif /* PUMP_HOST_H */
