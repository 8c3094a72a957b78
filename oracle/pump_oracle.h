/*
 * pump_oracle.h -- CPU restatement of PUMP's roadmap-edge and
 * collision-probability path (reference: StanfordASL/PUMP).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * by nothing else.  The product path (pump_amd/) never links or calls it.
 *
 * Semantics follow the reference's host build (g++/clang on x86-64):
 *   - cost() and pathPoint() evaluate pow(float,int) in double and round
 *     the result to float (2pBVP.cu:100-105, 119-130);
 *   - every other expression is IEEE binary32 with no FP contraction and
 *     correctly rounded division.
 * Each function cites the reference file:line it restates.
 */
#ifndef PUMP_ORACLE_H
#define PUMP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PO_DIM 6

/* ---- 2pBVP (2pBVP.cu) ------------------------------------------------- */
float po_cost(float tau, const float *x0, const float *x1);
float po_dcost(float tau, const float *x0, const float *x1);
float po_topt_bisection(const float *x0, const float *x1, float tmax);
void  po_path_point(float t, float tau, const float *x0, const float *x1, float *x);
float po_find_discretized_path(float *wp, const float *x0, const float *x1, int num_disc);
/* returns total time; *len = number of path rows written; returns -1 with
 * *len = -1 if more than cap rows would be written */
float po_find_optimal_path(float dt, float *split, const float *xs, int n_wp, int *len, int cap);

/* ---- roadmap precompute (2pBVP.cu:133-150, mainPUMP.cu:257-392) ------- */
void    po_fill_copts_topts(const float *samples, int num, float tmax,
                            float *copts, float *topts, int nthreads);
void    po_fill_rows(const float *samples, int num, float tmax, const int *rows, int n_rows,
                     float *copts, float *topts, int nthreads);
int64_t po_rn_index(int64_t num, float rn_perc);
float   po_select_rn(const float *copts, int64_t n_pairs, int64_t rn_idx);
/* count edges (copts < rn); fills row_start[num+1] (int64) */
int64_t po_count_edges(const float *copts, int num, float rn, int64_t *row_start);
/* fill edge arrays in reference numbering (row-major i, then j) */
void    po_fill_edges(const float *copts, const float *topts, int num, float rn,
                      const int64_t *row_start, int *edge_dst,
                      float *copts_edge, float *topts_edge);
/* padded neighbour tables (mainPUMP.cu:362-392); returns maxNNSize */
int     po_max_nn_size(const int64_t *row_start, const int *edge_dst, int64_t n_edges, int num);
void    po_nn_tables(const int64_t *row_start, const int *edge_dst, const float *copts_edge,
                     int num, int max_nn, int *nn_go_edges, int *nn_come_edges,
                     float *distances_come);
void    po_heuristic(const float *samples, int num, const float *goal, float tmax, float *h);

/* waypoint counts: fixed W when w_fixed > 0, else clamp(ceil(tau/wp_dt), 2, w_cap) */
int     po_waypoints_for(float tau, int w_fixed, float wp_dt, int w_cap);
/* discretize every edge (mainPUMP.cu:327-346 + findDiscretizedPath);
 * wp_offsets[n_edges+1] must already be filled. nthreads for OpenMP. */
void    po_discretize_edges(const float *samples, const int64_t *row_start, const int *edge_dst,
                            int num, const int64_t *wp_offsets, float *disc_motions, int nthreads);

/* ---- collision checking (sampler.cu, collisionCheck.cu) ---------------- */
void po_sample_free(const float *obstacles, int n_obs, const float *samples, int num,
                    uint8_t *is_free);
int  po_is_motion_valid(const float *v, const float *w, const float *bb_min,
                        const float *bb_max, int n_obs, const float *obstacles);
void po_free_edges(const float *obstacles, int n_obs, const float *disc_motions,
                   const int64_t *wp_offsets, int64_t n_edges, uint8_t *is_free_edges,
                   int nthreads);

/* ---- half-spaces (hsmc.cu, bvls.cu) ----------------------------------- */
void  po_setup_as_bs(int64_t hs_count, float *bs, float *as);
float po_maxd2(float offset_mult);
void  po_cache_halfspaces(int64_t n_edges, const float *disc_motions, const int64_t *wp_offsets,
                          const uint8_t *is_free_edges, int n_obs, int max_hs,
                          const float *obstacles, float *as, float *bs, int *count_stored,
                          float maxd2, int nthreads);

/* ---- HSMC (PUMP.cu:613-777) ------------------------------------------- */
void po_setup_valid_particles(uint8_t *valid, int n_p, int64_t max_paths,
                              float *path_cp, float *path_time);
/* returns the number of (path, waypoint) time indices that fell outside
 * [0, tmax_steps) and were clamped (the reference reads out of bounds) */
int64_t po_propagate_valid_particles(int n_new, const int64_t *wp_offsets, const int *wf_edge,
                                     const int *wf_prev, int64_t path_base, const float *as,
                                     const float *bs, const float *times, const float *topts,
                                     float dt, int max_hs, const float *offsets_time,
                                     int tmax_steps, float offset_mult, int n_mc,
                                     uint8_t *valid, int n_p, int nthreads);
void po_calculate_cp_hsmc(int n_new, int64_t path_base, float *path_cp, const uint8_t *valid,
                          int n_p);

/* ---- full MC (collisionProbability.cu) --------------------------------- */
int   po_mccp_hits(const float *obstacles, int n_obs, const float *xcomb, int tmax_steps,
                   int n_mc, float offset_mult, const float *path, int T, int *is_collision);
float po_collision_probability(const float *obstacles, int n_obs, const float *xcomb,
                               int tmax_steps, int n_mc, float offset_mult,
                               const float *path, int T);

/* ---- PUMP search consumer (PUMP.cu:3-405) ------------------------------ */

/* Optional record of a search, for tests: the stored path table, each
 * iteration's counters, and the fate of every wavefront entry.  A zero cap
 * (or a NULL array) skips that part; entries past a cap are not written. */
typedef struct po_search_trace {
    int64_t path_cap;               /* rows of the path arrays */
    int *path_prev, *path_node;     /* pathPrev, pathNode */
    int *path_bucket;               /* bucket a stored path was appended to (-1 for path 0) */
    float *path_cost, *path_time, *path_cp;
    int itr_cap;                    /* rows of the per-iteration arrays */
    int *itr_gidx;                  /* bucket expanded (Gidx) */
    int *itr_n_new, *itr_n_valid;   /* numNewPaths, numValidNewPaths */
    int *itr_size_g;                /* [itr][num_buckets] sizeG after the iteration */
    float *itr_threshold;           /* costThreshold used by the iteration */
    int64_t wf_cap;                 /* entries of the wavefront log */
    int *wf_prev, *wf_edge;         /* wavefrontPathPrev, wavefrontEdge, iterations concatenated */
    int *wf_fate;                   /* stored index, or -1 CP cutoff, -2 dominated, -3 node 0 */
    int64_t wf_len;                 /* out: entries the search produced (may exceed wf_cap) */
} po_search_trace;

typedef struct po_search_in {
    int num;
    const float *samples;           /* [num*6] */
    const float *goal;              /* [6] */
    int init_idx, goal_idx;
    const int64_t *row_start;       /* [num+1] out-edges of node i are row_start[i].. */
    const int *edge_dst;            /* [E] */
    const float *topts_edge, *copts_edge;
    const float *h;                 /* [num] */
    const uint8_t *is_free_edges;   /* [E] */
    const uint8_t *is_free_samples; /* [num] */
    const int64_t *wp_offsets;      /* [E+1] */
    const float *as, *bs;
    int max_hs;
    const float *offsets_time;      /* [tmax_steps][n_mc][3] */
    const float *xcomb;             /* [n_mc][tmax_steps][3] */
    int tmax_steps, n_mc;
    float offset_mult;
    const float *obstacles;
    int n_obs;
    float rn, lambda, dt, cp_target, cp_factor, eps_cost, eps_cp;
    int num_buckets, n_p, max_itrs;
    int64_t max_path_count;
    int path_cap;                   /* row capacity of the MCCP path buffer */
    int nthreads;
    po_search_trace *trace;         /* optional (NULL) */
} po_search_in;

typedef struct po_search_out {
    int status;           /* 0 ok/success, 1 no path below target, <0 aborted */
    int itrs;
    int64_t num_paths;
    int64_t num_considered;
    int num_goal_paths;
    int soln_path_idx;    /* -1 if none */
    float soln_cost, soln_cp, soln_time, soln_hsmc_cp;
    int soln_len;         /* number of nodes on the returned path */
    int soln_nodes[256];  /* init .. goal */
    int64_t clamped_time_idx;
    int64_t particle_steps;
    int mccp_calls;
} po_search_out;

int po_search(const po_search_in *in, po_search_out *out);

/* ---- FP64 oracle ------------------------------------------------------ */
double po64_cost(double tau, const float *x0, const float *x1);
void po64_rows_cost_error(const float *samples, int num, const int *rows, int nrows, float tmax,
                          const float *topts32, const float *copts32, const double *topts_pol,
                          double *stats, int nthreads);
/* tau* by exact-arithmetic restatement of the bisection bracket in FP64 on
 * the quartic (1+g^2)tau^4 - 4R tau^2 + 24 S tau - 36 D = 0 */
double po64_topt(const float *x0, const float *x1, double tmax);
/* the FP64 root of the same quartic nearest to t_guess */
double po64_topt_near(const float *x0, const float *x1, double t_guess, double tmax);
int    po64_mccp_hits(const float *obstacles, int n_obs, const float *xcomb, int tmax_steps,
                      int n_mc, double offset_mult, const double *path, int T);
/* path sampled in FP64 at the same dt schedule as findOptimalPath */
int    po64_find_optimal_path(double dt, double *split, const float *xs, int n_wp, int cap);

/* precomp.cu:3-76 and GMT.cu:3-261 */
void po_precomp_nn(const float *samples, int num, float r2, int nn_size, int *nn_sizes,
                   float *distances, int *nn_edges);
int po_gmt_wavefront(const float *obstacles, int n_obs, const float *distances_come,
                     const int *nn_go, const int *nn_come, int nn_size, const float *disc,
                     const int64_t *wp_offsets, const int64_t *row_start, const int *edge_dst,
                     const uint8_t *is_free_samples, int num, float *costs, int *edges, int goal,
                     float *cost_goal, int *itrs_out);
/* createSamplesHalton (sampler.cu:20-57) */
int po_halton_samples(int num, int skip, const float *init, const float *goal, const float *lo,
                      const float *hi, float *samples);
/* smoothing + controller discretisation (PUMP.cu:406-568) */
int po_smooth(const float *xs, int n, float soln_cost, float soln_cp, const float *obstacles,
              int n_obs, const float *xcomb, int tmax_steps, int n_mc, float offset_mult, float dt,
              float cp_target, int path_cap, float *best_path, int *best_rows, float *best_cost_cp,
              float *xs_smooth, int disc_cap, float *disc, int *disc_rows, float *trace, int *itrs);

#ifdef __cplusplus
}
#endif
#endif
