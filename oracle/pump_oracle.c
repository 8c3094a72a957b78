/*
 * pump_oracle.c -- CPU restatement of PUMP's roadmap-edge and
 * collision-probability path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product path.
 *
 * Arithmetic model (the reference's host build on x86-64):
 *   * cost() / pathPoint(): pow(float,int) promotes to double, so those
 *     expressions mix float and double sub-expressions.  Each statement below
 *     keeps the exact left-to-right grouping and the C type of every
 *     sub-expression of the reference expression, so the double sum and its
 *     final rounding to float are reproduced bit for bit.
 *   * Everything else: binary32, no contraction (built with
 *     -ffp-contract=off), correctly rounded division.
 *
 * Parity pinning (which check fixes which function) is listed per function.
 */
#include "pump_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define DIM PO_DIM
#define HALF 3

/* 2pBVP.cu:3 -- gravity constant (float) */
static const float PO_G = -9.81f;

/* ======================================================================
 * 2pBVP
 * ==================================================================== */

/*
 * cost(tau, x0, x1) -- 2pBVP.cu:97-107.
 * Pinned by: golden vectors from the reference's own 2pBVP.cu (oracle/_ref,
 * bit-exact); the closed form 12D/t^3 - 12S/t^2 + 4R/t + 2g(vz0-vz1) + (1+g^2)t
 * (tests/test_oracle_math.py).
 *
 * The numerator is a 30-term sum.  Terms written with pow() are double; the
 * terms built only from floats, int literals and tau are float products that
 * are rounded to float before they are added.  The sum runs strictly left to
 * right in the reference's term order.
 */
float po_cost(float tau, const float *a, const float *b)
{
    const float g = PO_G;
    double s;
    /* squared positions (double) and the one float cross product */
    s = 12 * pow(a[0], 2);
    s = s + 12 * pow(a[2], 2);
    s = s + 12 * pow(a[1], 2);
    s = s + 12 * pow(b[0], 2);
    s = s + 12 * pow(b[2], 2);
    s = s - (double)(24 * a[1] * b[1]);
    s = s + 12 * pow(b[1], 2);
    /* position x velocity x tau cross terms (float products) */
    s = s + (double)(12 * a[4] * a[1] * tau);
    s = s + (double)(12 * a[1] * b[4] * tau);
    s = s - (double)(12 * a[3] * b[0] * tau);
    s = s - (double)(12 * b[3] * b[0] * tau);
    s = s - (double)(12 * a[5] * b[2] * tau);
    s = s - (double)(12 * b[5] * b[2] * tau);
    s = s - (double)(12 * a[4] * b[1] * tau);
    s = s - (double)(12 * b[4] * b[1] * tau);
    /* velocity energy x tau^2 (double; cross products are float first) */
    s = s + 4 * pow(a[4], 2) * pow(tau, 2);
    s = s + 4 * pow(a[3], 2) * pow(tau, 2);
    s = s + 4 * pow(a[5], 2) * pow(tau, 2);
    s = s + (double)(4 * a[4] * b[4]) * pow(tau, 2);
    s = s + 4 * pow(b[4], 2) * pow(tau, 2);
    s = s + (double)(4 * a[3] * b[3]) * pow(tau, 2);
    s = s + 4 * pow(b[3], 2) * pow(tau, 2);
    s = s + (double)(4 * a[5] * b[5]) * pow(tau, 2);
    s = s + 4 * pow(b[5], 2) * pow(tau, 2);
    /* gravity and time terms */
    s = s + (double)(2 * g * a[5]) * pow(tau, 3);
    s = s - (double)(2 * g * b[5]) * pow(tau, 3);
    s = s + pow(tau, 4);
    s = s + pow(g, 2) * pow(tau, 4);
    /* x and z position cross terms, entirely float */
    s = s + (double)(12 * a[0] * (-2 * b[0] + (a[3] + b[3]) * tau));
    s = s + (double)(12 * a[2] * (-2 * b[2] + (a[5] + b[5]) * tau));
    return (float)((1 / pow(tau, 3)) * s);
}

/* dcost -- 2pBVP.cu:109-115: float central difference of two float costs.
 * Pinned by golden vectors (oracle/_ref). */
float po_dcost(float tau, const float *x0, const float *x1)
{
    const float h = 0.000001; /* double literal narrowed to float, as in the source */
    float up = po_cost(tau + h / 2, x0, x1);
    float dn = po_cost(tau - h / 2, x0, x1);
    return (up - dn) / h;
}

/* toptBisection -- 2pBVP.cu:71-95.
 * Pinned by golden vectors (oracle/_ref) and by the FP64 quartic root
 * (tests/test_oracle_math.py).
 * The only departure: the lower-bracket halving loop stops after 400
 * halvings (tl has then underflowed to 0); the reference loops forever in
 * that degenerate case (identical states). */
float po_topt_bisection(const float *x0, const float *x1, float tmax)
{
    float tu = tmax;
    if (po_dcost(tmax, x0, x1) < 0)
        return tmax;
    float tl = 0.01; /* 9.99999978e-3f */
    for (int k = 0; k < 400 && po_dcost(tl, x0, x1) > 0; ++k)
        tl = tl / 2;
    float topt = 0;
    float dcval = 1;
    int itr = 0;
    /* abs(float) is the float overload; the comparison is done in double */
    while ((double)fabsf(dcval) > 0.0001 && itr < 50) {
        ++itr;
        topt = (tu + tl) / 2;
        dcval = po_dcost(topt, x0, x1);
        if (dcval > 0)
            tu = topt;
        else
            tl = topt;
    }
    return topt;
}

/* x/y position row of pathPoint (2pBVP.cu:119-122): float linear part, then
 * two double quotients. */
static float pp_pos_xy(float t, float tau, float p0, float v0, float p1, float v1)
{
    float lin = t * v0 + p0;
    float c3 = 2 * p0 - 2 * p1 + (v0 + v1) * tau;
    float c2 = 3 * p0 - 3 * p1 + (2 * v0 + v1) * tau;
    return (float)((double)lin + (pow(t, 3) * c3) / pow(tau, 3) - (pow(t, 2) * c2) / pow(tau, 2));
}

/* x/y velocity row of pathPoint (2pBVP.cu:125-128). */
static float pp_vel_xy(float t, float tau, float p0, float v0, float p1, float v1)
{
    float e = 3 * p0 - 3 * p1 + v1 * tau;
    float f = -(2 * t * e);
    float c = 2 * p0 - 2 * p1 + v1 * tau;
    double q = 3 * pow(t, 2) - (double)(4 * t * tau) + pow(tau, 2);
    return (float)(f / pow(tau, 2) + (3 * pow(t, 2) * c) / pow(tau, 3) + (v0 * q) / pow(tau, 2));
}

/*
 * pathPoint -- 2pBVP.cu:117-131: the cubic (gravity-free) state at time t
 * along the optimal edge of duration tau.  The z rows (2pBVP.cu:123-124,
 * 129-130) use double literals, so their brackets are double throughout.
 * Pinned by golden vectors (oracle/_ref) and the endpoint identities
 * x(0)=x0, x(tau)=x1 (tests/test_oracle_math.py).
 */
void po_path_point(float t, float tau, const float *x0, const float *x1, float *x)
{
    x[0] = pp_pos_xy(t, tau, x0[0], x0[3], x1[0], x1[3]);
    x[1] = pp_pos_xy(t, tau, x0[1], x0[4], x1[1], x1[4]);
    {
        double a = (double)(t * x0[5]) * pow(tau, 3);
        double b = x0[2] * pow(tau, 3);
        double c = pow(t, 2) * tau *
                   (-3. * x0[2] + 3. * x1[2] - 2. * x0[5] * tau - 1. * x1[5] * tau);
        double d = pow(t, 3) * (2. * x0[2] - 2. * x1[2] + (double)((x0[5] + x1[5]) * tau));
        x[2] = (float)((a + b + c + d) / pow(tau, 3));
    }
    x[3] = pp_vel_xy(t, tau, x0[0], x0[3], x1[0], x1[3]);
    x[4] = pp_vel_xy(t, tau, x0[1], x0[4], x1[1], x1[4]);
    {
        double a = x0[5] * pow(tau, 3);
        double b = (double)(t * tau) *
                   (-6. * x0[2] + 6. * x1[2] - 4. * x0[5] * tau - 2. * x1[5] * tau);
        double c = pow(t, 2) * (6. * x0[2] - 6. * x1[2] + 3. * x0[5] * tau + 3. * x1[5] * tau);
        x[5] = (float)((a + b + c) / pow(tau, 3));
    }
}

/* findDiscretizedPath -- 2pBVP.cu:53-69: re-bisects tau (tmax 5), copies
 * the endpoints, interior waypoints at t = (tau/numDisc)*i.
 * Pinned by golden vectors (oracle/_ref). */
float po_find_discretized_path(float *wp, const float *x0, const float *x1, int num_disc)
{
    float topt = po_topt_bisection(x0, x1, 5);
    float dt = topt / num_disc;
    for (int d = 0; d < DIM; ++d) {
        wp[d] = x0[d];
        wp[d + num_disc * DIM] = x1[d];
    }
    for (int i = 1; i < num_disc; ++i)
        po_path_point(dt * i, topt, x0, x1, &wp[i * DIM]);
    return topt;
}

/* findOptimalPath -- 2pBVP.cu:5-36: concatenates the optimal edges through
 * the node states xs at a fixed dt, carrying the time offset across edges.
 * Pinned by golden vectors (oracle/_ref).  cap bounds the output rows (the
 * reference writes into a fixed Tmax-row array without a check). */
float po_find_optimal_path(float dt, float *split, const float *xs, int n_wp, int *len, int cap)
{
    float tmax = 5;
    float toff = 0, topt = 0, t = 0, ttotal = 0;
    int rows = 0;
    for (int i = 0; i < n_wp - 1; ++i) {
        const float *x0 = &xs[i * DIM];
        const float *x1 = &xs[(i + 1) * DIM];
        topt = po_topt_bisection(x0, x1, tmax);
        int T = (int)((topt - toff) / dt + 1);
        if (rows + (T > 0 ? T : 0) > cap) {
            *len = -1;
            return -1;
        }
        for (int k = 0; k < T; ++k) {
            t = dt * k + toff;
            po_path_point(t, topt, x0, x1, &split[(rows + k) * DIM]);
        }
        rows += T;
        toff = (t - topt) + dt;
        ttotal += topt;
    }
    *len = rows;
    return ttotal;
}

/* ======================================================================
 * Roadmap precompute
 * ==================================================================== */

/* fillCoptsTopts -- 2pBVP.cu:133-150: all ordered pairs i != j,
 * idx = i*(NUM-1) + j - [i<j].  Pinned by golden vectors. */
void po_fill_copts_topts(const float *samples, int num, float tmax, float *copts, float *topts,
                         int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4) if (nthreads != 1)
#endif
    for (int i = 0; i < num; ++i) {
        for (int j = 0; j < num; ++j) {
            if (i == j)
                continue;
            int64_t idx = (int64_t)i * (num - 1) + j - (i < j ? 1 : 0);
            float tau = po_topt_bisection(&samples[i * DIM], &samples[j * DIM], tmax);
            topts[idx] = tau;
            copts[idx] = po_cost(tau, &samples[i * DIM], &samples[j * DIM]);
        }
    }
    (void)nthreads;
}

/* fillCoptsTopts restricted to the source rows rows[0..n_rows): row k of
 * the outputs holds the NUM-1 entries of row rows[k], in the layout of
 * 2pBVP.cu:144-149 (j != i, ascending).  Same per-pair arithmetic as
 * po_fill_copts_topts; used to check sampled rows of full-size tables. */
void po_fill_rows(const float *samples, int num, float tmax, const int *rows, int n_rows,
                  float *copts, float *topts, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) if (nthreads != 1)
#endif
    for (int k = 0; k < n_rows; ++k) {
        const int i = rows[k];
        for (int j = 0; j < num; ++j) {
            if (i == j)
                continue;
            int64_t idx = (int64_t)k * (num - 1) + j - (i < j ? 1 : 0);
            float tau = po_topt_bisection(&samples[i * DIM], &samples[j * DIM], tmax);
            topts[idx] = tau;
            copts[idx] = po_cost(tau, &samples[i * DIM], &samples[j * DIM]);
        }
    }
    (void)nthreads;
}

/* rnIdx -- mainPUMP.cu:262: (int)((NUM-1)*NUM*rnPerc); the int pair count
 * is converted to float and multiplied in float.  (The int product is
 * formed in 64 bits here; the reference overflows for NUM >= 46341.) */
int64_t po_rn_index(int64_t num, float rn_perc)
{
    float prod = (float)((num - 1) * num) * rn_perc;
    return (int64_t)prod;
}

/* rn -- mainPUMP.cu:303-304: rn = sort(copts)[rnIdx].
 * Pinned by the order-statistic definition (numpy's full sort). */

/* order-preserving key of a non-NaN float (-0 orders just below +0) */
static uint32_t po_key(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

/* sorted(copts)[rn_idx] (mainPUMP.cu:303-304: thrust::sort, then the
 * rnIdx-th entry) as an exact selection: the rank falls in one bin of the
 * keys' upper 16 bits, then in one value of the lower 16 bits among that
 * bin's entries; two passes instead of a full sort. */
float po_select_rn(const float *copts, int64_t n_pairs, int64_t rn_idx)
{
    int64_t *hist = (int64_t *)calloc(65536, sizeof(int64_t));
    for (int64_t q = 0; q < n_pairs; ++q)
        hist[po_key(copts[q]) >> 16]++;
    int64_t below = 0;
    uint32_t hi = 0;
    while (below + hist[hi] <= rn_idx)
        below += hist[hi++];
    memset(hist, 0, 65536 * sizeof(int64_t));
    for (int64_t q = 0; q < n_pairs; ++q) {
        uint32_t k = po_key(copts[q]);
        if ((k >> 16) == hi)
            hist[k & 0xffffu]++;
    }
    uint32_t lo = 0;
    while (below + hist[lo] <= rn_idx)
        below += hist[lo++];
    free(hist);
    uint32_t key = (hi << 16) | lo;
    uint32_t u = (key & 0x80000000u) ? (key & 0x7fffffffu) : ~key;
    float rn;
    memcpy(&rn, &u, 4);
    return rn;
}

/* edge count per source row (mainPUMP.cu:327-346, strict copts < rn) */
int64_t po_count_edges(const float *copts, int num, float rn, int64_t *row_start)
{
    int64_t total = 0;
    for (int i = 0; i < num; ++i) {
        row_start[i] = total;
        const float *row = &copts[(int64_t)i * (num - 1)];
        for (int k = 0; k < num - 1; ++k)
            total += row[k] < rn;
    }
    row_start[num] = total;
    return total;
}

/* edge arrays in the reference's numbering (mainPUMP.cu:327-346):
 * coptsEdge/toptsEdge[e], and the destination j of edge e (the reference
 * stores the inverse map nnIdxs[j*NUM+i] = e). */
void po_fill_edges(const float *copts, const float *topts, int num, float rn,
                   const int64_t *row_start, int *edge_dst, float *copts_edge,
                   float *topts_edge)
{
    for (int i = 0; i < num; ++i) {
        int64_t e = row_start[i];
        for (int j = 0; j < num; ++j) {
            if (j == i)
                continue;
            int64_t idx = (int64_t)i * (num - 1) + j - (i < j ? 1 : 0);
            if (copts[idx] < rn) {
                edge_dst[e] = j;
                copts_edge[e] = copts[idx];
                topts_edge[e] = topts[idx];
                ++e;
            }
        }
    }
}

/* maxNNSize -- mainPUMP.cu:362-368: max over nodes of out- and in-degree */
int po_max_nn_size(const int64_t *row_start, const int *edge_dst, int64_t n_edges, int num)
{
    int *come = (int *)calloc((size_t)num, sizeof(int));
    int m = 0;
    for (int64_t e = 0; e < n_edges; ++e)
        come[edge_dst[e]]++;
    for (int i = 0; i < num; ++i) {
        int go = (int)(row_start[i + 1] - row_start[i]);
        if (m < go)
            m = go;
        if (m < come[i])
            m = come[i];
    }
    free(come);
    return m;
}

/* nnGoEdges / nnComeEdges / distancesCome -- mainPUMP.cu:371-392 */
void po_nn_tables(const int64_t *row_start, const int *edge_dst, const float *copts_edge, int num,
                  int max_nn, int *nn_go_edges, int *nn_come_edges, float *distances_come)
{
    int *come = (int *)calloc((size_t)num, sizeof(int));
    for (int64_t k = 0; k < (int64_t)num * max_nn; ++k) {
        nn_go_edges[k] = -1;
        nn_come_edges[k] = -1;
        distances_come[k] = 0;
    }
    for (int i = 0; i < num; ++i) {
        for (int64_t e = row_start[i]; e < row_start[i + 1]; ++e) {
            int j = edge_dst[e];
            nn_go_edges[(int64_t)i * max_nn + (e - row_start[i])] = j;
            nn_come_edges[(int64_t)j * max_nn + come[j]] = i;
            distances_come[(int64_t)j * max_nn + come[j]] = copts_edge[e];
            come[j]++;
        }
    }
    free(come);
}

/* heuristic h -- mainPUMP.cu:351-355: cost-to-goal for i < NUM-1, h[NUM-1]=0 */
void po_heuristic(const float *samples, int num, const float *goal, float tmax, float *h)
{
    for (int i = 0; i < num - 1; ++i) {
        float topt = po_topt_bisection(&samples[i * DIM], goal, tmax);
        h[i] = po_cost(topt, &samples[i * DIM], goal);
    }
    h[num - 1] = 0;
}

/* Waypoint count per edge.  Fixed W reproduces the reference's global
 * numDisc+1 (mainPUMP.cu:57); the variable form is the workload's
 * ceil(tau/dt) rule clamped to [2, cap]. */
int po_waypoints_for(float tau, int w_fixed, float wp_dt, int w_cap)
{
    if (w_fixed > 0)
        return w_fixed;
    int w = (int)ceilf(tau / wp_dt);
    if (w < 2)
        w = 2;
    if (w > w_cap)
        w = w_cap;
    return w;
}

/* Edge discretization loop -- mainPUMP.cu:327-346 with findDiscretizedPath
 * per edge (numDisc = W_e - 1). */
void po_discretize_edges(const float *samples, const int64_t *row_start, const int *edge_dst,
                         int num, const int64_t *wp_offsets, float *disc_motions, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 8) if (nthreads != 1)
#endif
    for (int i = 0; i < num; ++i) {
        for (int64_t e = row_start[i]; e < row_start[i + 1]; ++e) {
            int nd = (int)(wp_offsets[e + 1] - wp_offsets[e]) - 1;
            po_find_discretized_path(&disc_motions[wp_offsets[e] * DIM], &samples[i * DIM],
                                     &samples[edge_dst[e] * DIM], nd);
        }
    }
    (void)nthreads;
}

/* ======================================================================
 * Collision checking
 * ==================================================================== */

/* sampleFree -- sampler.cu:61-88: a sample is blocked iff its position is
 * strictly inside some box in all three position axes.
 * Pinned by the closed form (strict point-in-open-box, boundary cases). */
void po_sample_free(const float *obstacles, int n_obs, const float *samples, int num,
                    uint8_t *is_free)
{
    for (int n = 0; n < num; ++n) {
        const float *p = &samples[n * DIM];
        uint8_t free_ = 1;
        for (int o = 0; o < n_obs && free_; ++o) {
            const float *lo = &obstacles[o * 2 * DIM], *hi = lo + DIM;
            if (p[0] > lo[0] && p[0] < hi[0] && p[1] > lo[1] && p[1] < hi[1] && p[2] > lo[2] &&
                p[2] < hi[2])
                free_ = 0;
        }
        is_free[n] = free_;
    }
}

/* faceContainsProjection -- collisionCheck.cu:133-143 (closed intervals) */
static int face_contains(const float *v, const float *w, float lambda, int j, const float *lo,
                         const float *hi)
{
    for (int d = 0; d < HALF; ++d) {
        float proj = v[d] + (w[d] - v[d]) * lambda;
        if (d != j && !(lo[d] <= proj && proj <= hi[d]))
            return 0;
    }
    return 1;
}

/* motionValidQ -- collisionCheck.cu:114-131: per axis, the face on v's side
 * (lo if v<lo else hi), lambda unclamped and unguarded. */
static int motion_valid_q(const float *v, const float *w, const float *lo, const float *hi)
{
    for (int d = 0; d < HALF; ++d) {
        float vw = w[d] - v[d];
        float lambda = (v[d] < lo[d]) ? (lo[d] - v[d]) / vw : (hi[d] - v[d]) / vw;
        if (face_contains(v, w, lambda, d, lo, hi))
            return 0;
    }
    return 1;
}

/* isMotionValid + broadphaseValidQ -- collisionCheck.cu:65-112.
 * Pinned by brute force on the segment/box geometry (an entering segment
 * must be reported) and separating-axis cases (tests/test_oracle_math.py). */
int po_is_motion_valid(const float *v, const float *w, const float *bb_min, const float *bb_max,
                       int n_obs, const float *obstacles)
{
    for (int o = 0; o < n_obs; ++o) {
        const float *lo = &obstacles[o * 2 * DIM], *hi = lo + DIM;
        int separated = 0;
        for (int d = 0; d < HALF; ++d) {
            if (bb_max[d] <= lo[d] || hi[d] <= bb_min[d]) {
                separated = 1;
                break;
            }
        }
        if (!separated && !motion_valid_q(v, w, lo, hi))
            return 0;
    }
    return 1;
}

/* freeEdges -- collisionCheck.cu:5-37: AND over the edge's segments. */
void po_free_edges(const float *obstacles, int n_obs, const float *disc_motions,
                   const int64_t *wp_offsets, int64_t n_edges, uint8_t *is_free_edges,
                   int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256) if (nthreads != 1)
#endif
    for (int64_t e = 0; e < n_edges; ++e) {
        int nseg = (int)(wp_offsets[e + 1] - wp_offsets[e]) - 1;
        const float *wp = &disc_motions[wp_offsets[e] * DIM];
        int ok = 1;
        for (int i = 0; i < nseg && ok; ++i) {
            const float *v = &wp[i * DIM], *w = &wp[(i + 1) * DIM];
            float bmin[DIM], bmax[DIM];
            for (int d = 0; d < DIM; ++d) {
                if (v[d] > w[d]) {
                    bmin[d] = w[d];
                    bmax[d] = v[d];
                } else {
                    bmin[d] = v[d];
                    bmax[d] = w[d];
                }
            }
            ok = po_is_motion_valid(v, w, bmin, bmax, n_obs, obstacles);
        }
        is_free_edges[e] = (uint8_t)ok;
    }
    (void)nthreads;
}

/* ======================================================================
 * Half-spaces
 * ==================================================================== */

/* setupAsBs -- hsmc.cu:201-211: a = -1, b = 1111 (b >= 1000 is "empty") */
void po_setup_as_bs(int64_t hs_count, float *bs, float *as)
{
    for (int64_t k = 0; k < hs_count; ++k) {
        as[k * 3 + 0] = -1;
        as[k * 3 + 1] = -1;
        as[k * 3 + 2] = -1;
        bs[k] = 1111;
    }
}

/* maxd2 -- mainPUMP.cu:63,201: 0.1*offsetMult*offsetMult in double,
 * narrowed to float on assignment. */
float po_maxd2(float offset_mult)
{
    return (float)(0.1 * offset_mult * offset_mult);
}

/*
 * cacheHalfspaces + bvlsShortcut -- hsmc.cu:10-99, bvls.cu:259-273.
 * For each waypoint of a free edge, walk the boxes in index order: closest
 * box point x (clamp), skip if |x - pos|^2 > maxd2, else remove the velocity
 * component and store a = x, b = a.a; stop after maxHS stored (first by
 * index, not nearest).  Waypoints of blocked edges get count 0 here (the
 * reference leaves that slot unwritten).
 * Pinned by: closest-point-on-box closed form, b = |a|^2 and a.v = 0
 * invariants, first-k-by-index rule (tests/test_oracle_math.py).
 */
void po_cache_halfspaces(int64_t n_edges, const float *disc_motions, const int64_t *wp_offsets,
                         const uint8_t *is_free_edges, int n_obs, int max_hs,
                         const float *obstacles, float *as, float *bs, int *count_stored,
                         float maxd2, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256) if (nthreads != 1)
#endif
    for (int64_t e = 0; e < n_edges; ++e) {
        for (int64_t wpi = wp_offsets[e]; wpi < wp_offsets[e + 1]; ++wpi) {
            if (!is_free_edges[e]) {
                count_stored[wpi] = 0;
                continue;
            }
            const float *pos = &disc_motions[wpi * DIM];
            const float *vel = pos + HALF;
            int64_t hs = wpi * max_hs;
            int stored = 0;
            for (int o = 0; o < n_obs && stored < max_hs; ++o) {
                const float *lo = &obstacles[o * 2 * DIM], *hi = lo + DIM;
                float x[HALF];
                for (int d = 0; d < HALF; ++d)
                    x[d] = (pos[d] <= lo[d]) ? lo[d] : (pos[d] >= hi[d]) ? hi[d] : pos[d];
                for (int d = 0; d < HALF; ++d)
                    x[d] = x[d] - pos[d];
                float d2 = 0;
                for (int d = 0; d < HALF; ++d)
                    d2 = d2 + x[d] * x[d];
                if (d2 > maxd2)
                    continue;
                float xdv = 0, vdv = 0;
                for (int d = 0; d < HALF; ++d) {
                    xdv = xdv + x[d] * vel[d];
                    vdv = vdv + vel[d] * vel[d];
                }
                if (vdv == 0)
                    vdv = 1e-5;
                float ratio = xdv / vdv;
                for (int d = 0; d < HALF; ++d)
                    x[d] = x[d] - ratio * vel[d];
                float xx = 0;
                for (int d = 0; d < HALF; ++d) {
                    as[hs * HALF + d] = x[d];
                    xx = xx + x[d] * x[d];
                }
                bs[hs] = xx;
                ++stored;
                ++hs;
            }
            count_stored[wpi] = stored;
        }
    }
    (void)nthreads;
}

/* ======================================================================
 * HSMC
 * ==================================================================== */

/* setupValidParticles -- PUMP.cu:766-777 */
void po_setup_valid_particles(uint8_t *valid, int n_p, int64_t max_paths, float *path_cp,
                              float *path_time)
{
    memset(valid, 1, (size_t)(max_paths * n_p));
    for (int64_t k = 0; k < max_paths; ++k) {
        path_cp[k] = 1;
        path_time[k] = 0;
    }
}

/*
 * propagateValidParticles -- PUMP.cu:613-655.
 * valid(child, p) = valid(parent, p) AND, for every waypoint i of the edge
 * and every stored half-space (b < 1000), (offset(t_i, p).a * offsetMult < b)
 * with t_i = times[parent] + i * (tau_e / (W_e - 1)), dtIdx = (int)(t_i/dt).
 * The 3-term dot product accumulates left to right from 0 in float.
 * Pinned by: the half-space inequality evaluated independently in FP64 away
 * from the boundary, monotone validity, zero-offset particles
 * (tests/test_oracle_math.py).
 * Out-of-table time indices are clamped and counted (the reference reads
 * past its table).
 */
int64_t po_propagate_valid_particles(int n_new, const int64_t *wp_offsets, const int *wf_edge,
                                     const int *wf_prev, int64_t path_base, const float *as,
                                     const float *bs, const float *times, const float *topts,
                                     float dt, int max_hs, const float *offsets_time,
                                     int tmax_steps, float offset_mult, int n_mc,
                                     uint8_t *valid, int n_p, int nthreads)
{
    int64_t clamped = 0;
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : clamped) if (nthreads != 1)
#endif
    for (int w = 0; w < n_new; ++w) {
        int e = wf_edge[w];
        int prev = wf_prev[w];
        int W = (int)(wp_offsets[e + 1] - wp_offsets[e]);
        int64_t hs_head = wp_offsets[e] * max_hs;
        float dt_local = topts[e] / (W - 1);
        const uint8_t *src = &valid[(int64_t)prev * n_p];
        uint8_t *dst = &valid[(path_base + w) * n_p];
        /* per waypoint time index (uniform over particles) */
        for (int p = 0; p < n_p; ++p)
            dst[p] = src[p];
        for (int i = 0; i < W; ++i) {
            float time = times[prev] + i * dt_local;
            int ti = (int)(time / dt);
            int any = 0;
            for (int o = 0; o < max_hs; ++o)
                any |= bs[hs_head + (int64_t)i * max_hs + o] < 1000;
            if (any && (ti < 0 || ti >= tmax_steps)) {
                clamped++;
                ti = ti < 0 ? 0 : tmax_steps - 1;
            }
            for (int o = 0; o < max_hs; ++o) {
                float b = bs[hs_head + (int64_t)i * max_hs + o];
                if (b >= 1000)
                    continue;
                const float *a = &as[(hs_head + (int64_t)i * max_hs + o) * HALF];
                const float *off = &offsets_time[(int64_t)ti * n_mc * HALF];
                for (int p = 0; p < n_p; ++p) {
                    float lhs = 0;
                    for (int d = 0; d < HALF; ++d)
                        lhs = lhs + off[p * HALF + d] * a[d];
                    dst[p] = (uint8_t)(dst[p] & (lhs * offset_mult < b));
                }
            }
        }
    }
    (void)nthreads;
    return clamped;
}

/* calculateCP_HSMC -- PUMP.cu:745-758: (float)count / nP */
void po_calculate_cp_hsmc(int n_new, int64_t path_base, float *path_cp, const uint8_t *valid,
                          int n_p)
{
    for (int w = 0; w < n_new; ++w) {
        int hits = 0;
        const uint8_t *v = &valid[(path_base + w) * n_p];
        for (int p = 0; p < n_p; ++p)
            hits += !v[p];
        path_cp[path_base + w] = ((float)hits) / n_p;
    }
}

/* ======================================================================
 * Full Monte Carlo CP
 * ==================================================================== */

/* MCCP -- collisionProbability.cu:39-92: particle p perturbs the nominal
 * path by xcomb*offsetMult (positions only) and checks segments t < T-2
 * (the last segment is skipped, as in the source).
 * Pinned by: a zero-noise table reduces to the nominal path's segment
 * check; the FP64 twin below away from boundaries. */
int po_mccp_hits(const float *obstacles, int n_obs, const float *xcomb, int tmax_steps, int n_mc,
                 float offset_mult, const float *path, int T, int *is_collision)
{
    int total = 0;
    for (int p = 0; p < n_mc; ++p) {
        int hit = 0;
        for (int t = 0; t < T - 2 && !hit; ++t) {
            float v[HALF], w[HALF], bmin[HALF], bmax[HALF];
            for (int d = 0; d < HALF; ++d) {
                v[d] = path[t * DIM + d] +
                       xcomb[(int64_t)p * HALF * tmax_steps + t * HALF + d] * offset_mult;
                w[d] = path[(t + 1) * DIM + d] +
                       xcomb[(int64_t)p * HALF * tmax_steps + (t + 1) * HALF + d] * offset_mult;
                if (v[d] > w[d]) {
                    bmin[d] = w[d];
                    bmax[d] = v[d];
                } else {
                    bmin[d] = v[d];
                    bmax[d] = w[d];
                }
            }
            if (!po_is_motion_valid(v, w, bmin, bmax, n_obs, obstacles))
                hit = 1;
        }
        if (is_collision)
            is_collision[p] = hit;
        total += hit;
    }
    return total;
}

/* collisionProbability -- collisionProbability.cu:4-37: hits / nMC */
float po_collision_probability(const float *obstacles, int n_obs, const float *xcomb,
                               int tmax_steps, int n_mc, float offset_mult, const float *path,
                               int T)
{
    int hits = po_mccp_hits(obstacles, n_obs, xcomb, tmax_steps, n_mc, offset_mult, path, T, 0);
    return ((float)hits) / ((float)n_mc);
}

/* ======================================================================
 * PUMP search consumer -- PUMP.cu:3-405
 * ==================================================================== */

typedef struct ivec {
    int *v;
    int n, cap;
} ivec;

static void ivec_push(ivec *a, int x)
{
    if (a->n == a->cap) {
        a->cap = a->cap ? a->cap * 2 : 8;
        a->v = (int *)realloc(a->v, sizeof(int) * (size_t)a->cap);
    }
    a->v[a->n++] = x;
}

typedef struct goal_pair {
    float cost;
    int idx;
} goal_pair;

static int cmp_goal(const void *pa, const void *pb)
{
    const goal_pair *a = (const goal_pair *)pa, *b = (const goal_pair *)pb;
    if (a->cost < b->cost)
        return -1;
    if (b->cost < a->cost)
        return 1;
    return (a->idx > b->idx) - (a->idx < b->idx);
}

/*
 * The multiobjective wavefront search (PUMP.cu:3-290) and the goal-path
 * selection (PUMP.cu:354-405).  Dense NUM x NUM tables of the reference
 * (P, nnIdxs) are held as per-node lists / CSR rows with identical order:
 * out-edge nn of node i is edge row_start[i] + nn (destinations ascending),
 * which is exactly nnGoEdges[i][nn] and nnIdxs[j][i].
 * Validity rows are held for the paths that exist (stored paths plus the
 * current wavefront) instead of maxPathCount rows up front: every row is
 * written (copied from its parent by propagate) before it is read, as in the
 * reference, so only the allocation differs.
 * Pinned by: hand-derived searches on small roadmaps (tests/
 * test_oracle_search.py: per-iteration wavefronts, fates, buckets, the
 * stored path table and the returned path, each rule of PUMP.cu:178-405
 * exercised), and invariants (CP monotone along paths, returned path is a
 * chain of free roadmap edges, pathCost = sum of edge costs).
 */
int po_search(const po_search_in *in, po_search_out *out)
{
    memset(out, 0, sizeof(*out));
    out->soln_path_idx = -1;
    const int num = in->num;
    const int n_p = in->n_p;
    const int64_t maxp = in->max_path_count;
    int nb = in->num_buckets;

    int *path_prev = (int *)malloc(sizeof(int) * (size_t)maxp);
    int *path_node = (int *)malloc(sizeof(int) * (size_t)maxp);
    float *path_cost = (float *)malloc(sizeof(float) * (size_t)maxp);
    float *path_time = (float *)malloc(sizeof(float) * (size_t)maxp);
    float *path_cp = (float *)malloc(sizeof(float) * (size_t)maxp);
    float *d_path_cp = (float *)malloc(sizeof(float) * (size_t)maxp);   /* device mirror */
    float *d_path_time = (float *)malloc(sizeof(float) * (size_t)maxp); /* device mirror */
    int64_t vcap = 0; /* validity rows allocated */
    uint8_t *valid = NULL;
    int *G = (int *)malloc(sizeof(int) * (size_t)(nb * maxp));
    int *sizeG = (int *)calloc((size_t)nb, sizeof(int));
    int *wf_prev = (int *)malloc(sizeof(int) * (size_t)maxp);
    int *wf_next = (int *)malloc(sizeof(int) * (size_t)maxp);
    int *wf_edge = (int *)malloc(sizeof(int) * (size_t)maxp);
    ivec *P = (ivec *)calloc((size_t)num, sizeof(ivec));
    int *pnew = (int *)calloc((size_t)num, sizeof(int));
    int *fate = (int *)malloc(sizeof(int) * (size_t)maxp);
    po_search_trace *tr = in->trace;
    int status = 0;

    /* setupValidParticles (PUMP.cu:766-777): every row true, CP 1, time 0;
     * rows beyond path 0 are allocated when a wavefront needs them */
    vcap = 1;
    valid = (uint8_t *)malloc((size_t)n_p);
    memset(valid, 1, (size_t)n_p);
    for (int64_t k = 0; k < maxp; ++k) {
        d_path_cp[k] = 1;
        d_path_time[k] = 0;
    }
    if (tr)
        tr->wf_len = 0;
    for (int64_t k = 0; k < nb * maxp; ++k)
        G[k] = -1;

    const float dr = in->lambda * in->rn;
    const float cp_min_soln = in->cp_target / in->cp_factor;
    const float cp_cutoff = in->cp_target * in->cp_factor;

    int64_t num_paths = 1;
    int gidx = 0;
    int goal_cond = 0, goal_next = 0, empty_open = 0;
    ivec_push(&P[in->init_idx], 0);
    path_prev[0] = -2;
    path_node[0] = in->init_idx;
    path_cost[0] = 0;
    path_cp[0] = 0;
    path_time[0] = 0;
    sizeG[gidx]++;
    G[0] = 0;
    float cost_threshold = in->h[in->init_idx];
    int itrs = 0;
    if (tr && tr->path_cap > 0 && tr->path_bucket)
        tr->path_bucket[0] = -1;

    while (itrs < in->max_itrs && !goal_cond && !empty_open) {
        ++itrs;
        if (goal_next)
            goal_cond = 1;
        int n_new = 0;
        int overflow = 0;
        for (int g = 0; g < sizeG[gidx] && !overflow; ++g) {
            int pprev = G[gidx * maxp + g];
            G[gidx * maxp + g] = -1;
            int nprev = path_node[pprev];
            for (int64_t e = in->row_start[nprev]; e < in->row_start[nprev + 1]; ++e) {
                int nnext = in->edge_dst[e];
                if (!in->is_free_edges[e] || !in->is_free_samples[nnext])
                    continue;
                wf_prev[n_new] = pprev;
                wf_next[n_new] = nnext;
                wf_edge[n_new] = (int)e;
                n_new++;
                if (n_new >= maxp) { /* reference: > maxPathCount, writes past the end first */
                    overflow = 1;
                    break;
                }
            }
        }
        if (overflow) {
            status = -2;
            goto done;
        }
        if (n_new == 0)
            break;
        if (num_paths + n_new >= maxp) {
            status = -3;
            goto done;
        }
        sizeG[gidx] = 0;
        out->num_considered += n_new;
        if (num_paths + n_new > vcap) {
            int64_t want = num_paths + n_new;
            want += want / 2;
            if (want > maxp)
                want = maxp;
            uint8_t *nv = (uint8_t *)realloc(valid, (size_t)(want * n_p));
            if (!nv) {
                status = -7;
                goto done;
            }
            valid = nv;
            vcap = want;
        }
        for (int i = 0; i < n_new; ++i)
            fate[i] = 0;
        for (int w = 0; w < n_new; ++w)
            out->particle_steps +=
                (int64_t)n_p * (in->wp_offsets[wf_edge[w] + 1] - in->wp_offsets[wf_edge[w]]);

        out->clamped_time_idx += po_propagate_valid_particles(
            n_new, in->wp_offsets, wf_edge, wf_prev, num_paths, in->as, in->bs, d_path_time,
            in->topts_edge, in->dt, in->max_hs, in->offsets_time, in->tmax_steps, in->offset_mult,
            in->n_mc, valid, n_p, in->nthreads);
        po_calculate_cp_hsmc(n_new, num_paths, d_path_cp, valid, n_p);
        memcpy(&path_cp[num_paths], &d_path_cp[num_paths], sizeof(float) * (size_t)n_new);

        for (int i = 0; i < n_new; ++i) {
            path_time[num_paths + i] = path_time[wf_prev[i]] + in->topts_edge[wf_edge[i]];
            path_cost[num_paths + i] = path_cost[wf_prev[i]] + in->copts_edge[wf_edge[i]];
        }

        /* dominance: stage all surviving new paths at their node */
        for (int i = 0; i < n_new; ++i) {
            int node = wf_next[i];
            int64_t pidx = num_paths + i;
            if (path_cp[pidx] > cp_cutoff) {
                wf_next[i] = -1;
                fate[i] = -1;
                continue;
            }
            ivec_push(&P[node], (int)pidx);
            pnew[node]++;
        }
        for (int i = 0; i < n_new; ++i) {
            int node = wf_next[i];
            if (node == -1 || node == in->goal_idx)
                continue;
            int pidx = (int)(num_paths + i);
            for (int j = 0; j < P[node].n; ++j) {
                int cmp = P[node].v[j];
                if (cmp == pidx)
                    continue;
                if (path_cp[cmp] < in->eps_cp + path_cp[pidx] &&
                    path_cost[cmp] < in->eps_cost + path_cost[pidx]) {
                    if (path_cp[cmp] + in->eps_cp >= path_cp[pidx] &&
                        path_cost[cmp] + in->eps_cost >= path_cost[pidx] && pidx < cmp)
                        continue;
                    wf_next[i] = -1;
                    fate[i] = -2;
                    break;
                }
            }
        }
        /* drop the staged entries again; survivors are re-added compacted */
        for (int node = 0; node < num; ++node) {
            if (pnew[node]) {
                P[node].n -= pnew[node];
                pnew[node] = 0;
            }
        }

        /* store good paths (compaction) */
        int n_valid = 0;
        for (int i = 0; i < n_new; ++i) {
            int node = wf_next[i];
            int64_t pidx = num_paths + i;
            if (node == -1 || node == 0) {
                if (node == 0)
                    fate[i] = -3;
                wf_next[i] = -1;
                continue;
            }
            int64_t st = num_paths + n_valid;
            fate[i] = (int)st;
            path_cp[st] = path_cp[pidx];
            path_time[st] = path_time[pidx];
            path_cost[st] = path_cost[pidx];
            memmove(&valid[st * n_p], &valid[pidx * n_p], (size_t)n_p);
            float dcost = (path_cost[st] + in->h[node] - cost_threshold);
            dcost = dcost > 0 ? dcost : 0;
            int bucket = (((int)(dcost / dr)) + gidx + 1) % nb;
            if (bucket < 0 || bucket >= nb) { /* negative rn: the reference indexes out of G */
                status = -6;
                goto done;
            }
            G[bucket * maxp + sizeG[bucket]] = (int)st;
            sizeG[bucket]++;
            if (tr && st < tr->path_cap && tr->path_bucket)
                tr->path_bucket[st] = bucket;
            path_prev[st] = wf_prev[i];
            path_node[st] = node;
            ivec_push(&P[node], (int)st);
            wf_next[i] = -1;
            n_valid++;
        }
        memcpy(&d_path_cp[num_paths], &path_cp[num_paths], sizeof(float) * (size_t)n_valid);
        memcpy(&d_path_time[num_paths], &path_time[num_paths], sizeof(float) * (size_t)n_valid);
        num_paths += n_valid;
        if (tr) {
            int it = itrs - 1;
            if (it < tr->itr_cap) {
                if (tr->itr_gidx)
                    tr->itr_gidx[it] = gidx;
                if (tr->itr_n_new)
                    tr->itr_n_new[it] = n_new;
                if (tr->itr_n_valid)
                    tr->itr_n_valid[it] = n_valid;
                if (tr->itr_threshold)
                    tr->itr_threshold[it] = cost_threshold;
                if (tr->itr_size_g)
                    for (int b = 0; b < nb; ++b)
                        tr->itr_size_g[it * nb + b] = sizeG[b];
            }
            for (int i = 0; i < n_new; ++i) {
                int64_t k = tr->wf_len + i;
                if (k >= tr->wf_cap)
                    break;
                if (tr->wf_prev)
                    tr->wf_prev[k] = wf_prev[i];
                if (tr->wf_edge)
                    tr->wf_edge[k] = wf_edge[i];
                if (tr->wf_fate)
                    tr->wf_fate[k] = fate[i];
            }
            tr->wf_len += n_new;
        }

        for (int i = 0; i < P[in->goal_idx].n; ++i) {
            int pidx = P[in->goal_idx].v[i];
            if (path_cp[pidx] <= cp_min_soln) {
                goal_next = 1;
                goal_cond = 1;
            }
        }
        if (goal_cond)
            break;
        empty_open = 1;
        for (int b = 0; b < nb; ++b)
            empty_open = empty_open && (sizeG[b] == 0);
        if (empty_open)
            break;
        gidx = (gidx + 1) % nb;
        cost_threshold += dr;
        if (num_paths >= maxp) {
            status = -4;
            goto done;
        }
    }

    /* goal-path selection (PUMP.cu:354-405) */
    {
        int ng = P[in->goal_idx].n;
        out->num_goal_paths = ng;
        goal_pair *gp = (goal_pair *)malloc(sizeof(goal_pair) * (size_t)(ng > 0 ? ng : 1));
        for (int i = 0; i < ng; ++i) {
            gp[i].idx = P[in->goal_idx].v[i];
            gp[i].cost = path_cost[gp[i].idx];
        }
        qsort(gp, (size_t)ng, sizeof(goal_pair), cmp_goal);
        float *xs = (float *)malloc(sizeof(float) * DIM * 256);
        float *split = (float *)malloc(sizeof(float) * DIM * (size_t)in->path_cap);
        status = 1;
        for (int k = 0; k < ng; ++k) {
            int nodes[256];
            int cnt = 0;
            int pidx = gp[k].idx;
            while (pidx != -2 && cnt < 256) {
                nodes[cnt++] = path_node[pidx];
                pidx = path_prev[pidx];
            }
            for (int q = 0; q < cnt; ++q) /* init first */
                memcpy(&xs[q * DIM], &in->samples[nodes[cnt - 1 - q] * DIM], sizeof(float) * DIM);
            int len = 0;
            po_find_optimal_path(in->dt, split, xs, cnt, &len, in->path_cap);
            if (len < 0) {
                status = -5;
                break;
            }
            float cp = po_collision_probability(in->obstacles, in->n_obs, in->xcomb,
                                                in->tmax_steps, in->n_mc, in->offset_mult, split,
                                                len);
            out->mccp_calls++;
            out->soln_path_idx = gp[k].idx;
            out->soln_cp = cp;
            out->soln_cost = path_cost[gp[k].idx];
            out->soln_time = path_time[gp[k].idx];
            out->soln_hsmc_cp = path_cp[gp[k].idx];
            out->soln_len = cnt;
            for (int q = 0; q < cnt; ++q)
                out->soln_nodes[q] = nodes[cnt - 1 - q];
            if (cp < in->cp_target) {
                status = 0;
                break;
            }
        }
        if (status == 1 && ng == 0)
            out->soln_path_idx = -1;
        free(xs);
        free(split);
        free(gp);
    }

done:
    out->status = status;
    out->itrs = itrs;
    out->num_paths = num_paths;
    if (tr && tr->path_cap > 0) {
        int64_t n = num_paths < tr->path_cap ? num_paths : tr->path_cap;
        for (int64_t k = 0; k < n; ++k) {
            if (tr->path_prev)
                tr->path_prev[k] = path_prev[k];
            if (tr->path_node)
                tr->path_node[k] = path_node[k];
            if (tr->path_cost)
                tr->path_cost[k] = path_cost[k];
            if (tr->path_time)
                tr->path_time[k] = path_time[k];
            if (tr->path_cp)
                tr->path_cp[k] = path_cp[k];
        }
    }
    for (int n = 0; n < num; ++n)
        free(P[n].v);
    free(P);
    free(pnew);
    free(path_prev);
    free(path_node);
    free(path_cost);
    free(path_time);
    free(path_cp);
    free(d_path_cp);
    free(d_path_time);
    free(valid);
    free(G);
    free(sizeG);
    free(wf_prev);
    free(wf_next);
    free(wf_edge);
    free(fate);
    return status;
}

/* ======================================================================
 * FP64 oracle (accuracy reference for the chosen precision forms)
 * ==================================================================== */

/* c(tau) in exact form: 12D/t^3 - 12S/t^2 + 4R/t + 2g(vz0 - vz1) + (1+g^2) t,
 * D = sum dp^2, S = sum dp (v0+v1), R = sum (v0^2 + v0 v1 + v1^2). */
static void po64_drs(const float *x0, const float *x1, double *D, double *S, double *R)
{
    *D = *S = *R = 0;
    for (int d = 0; d < HALF; ++d) {
        double dp = (double)x1[d] - (double)x0[d];
        double v0 = x0[d + 3], v1 = x1[d + 3];
        *D += dp * dp;
        *S += dp * (v0 + v1);
        *R += v0 * v0 + v0 * v1 + v1 * v1;
    }
}

double po64_cost(double tau, const float *x0, const float *x1)
{
    double D, S, R;
    const double g = (double)PO_G;
    po64_drs(x0, x1, &D, &S, &R);
    return 12 * D / (tau * tau * tau) - 12 * S / (tau * tau) + 4 * R / tau +
           2 * g * ((double)x0[5] - (double)x1[5]) + (1 + g * g) * tau;
}

/* q(t) = t^4 c'(t) = (1+g^2) t^4 - 4R t^2 + 24 S t - 36 D */
static double po64_q(double t, double D, double S, double R)
{
    const double g = (double)PO_G;
    return (((1 + g * g) * t * t - 4 * R) * t + 24 * S) * t - 36 * D;
}

double po64_topt(const float *x0, const float *x1, double tmax)
{
    double D, S, R;
    po64_drs(x0, x1, &D, &S, &R);
    if (po64_q(tmax, D, S, R) < 0)
        return tmax;
    double tl = 0.01;
    for (int k = 0; k < 2000 && po64_q(tl, D, S, R) > 0; ++k)
        tl /= 2;
    double tu = tmax;
    for (int k = 0; k < 200 && tu - tl > 1e-15 * tu; ++k) {
        double m = 0.5 * (tl + tu);
        if (po64_q(m, D, S, R) > 0)
            tu = m;
        else
            tl = m;
    }
    return 0.5 * (tl + tu);
}

/* The root of q nearest to t_guess (the FP64 stationary point at the same
 * local minimum a bisection landed on): widen a bracket around t_guess until
 * q changes sign from - to +, then bisect to the last bit. */
double po64_topt_near(const float *x0, const float *x1, double t_guess, double tmax)
{
    double D, S, R;
    po64_drs(x0, x1, &D, &S, &R);
    if (t_guess >= tmax && po64_q(tmax, D, S, R) < 0)
        return tmax;
    double lo = t_guess, hi = t_guess;
    for (int k = 0; k < 200; ++k) {
        double w = t_guess * 1e-6 * (double)(1ull << (k < 60 ? k : 60));
        lo = t_guess - w > 1e-9 ? t_guess - w : 1e-9;
        hi = t_guess + w < tmax ? t_guess + w : tmax;
        if (po64_q(lo, D, S, R) <= 0 && po64_q(hi, D, S, R) >= 0)
            break;
    }
    for (int k = 0; k < 200 && hi - lo > 1e-15 * hi; ++k) {
        double m = 0.5 * (lo + hi);
        if (po64_q(m, D, S, R) > 0)
            hi = m;
        else
            lo = m;
    }
    return 0.5 * (lo + hi);
}

/* Cost errors of sampled rows against the FP64 oracle.  For each pair (i, j)
 * of the rows, tau64 = po64_topt_near(x_i, x_j, topts32, tmax) is the FP64
 * stationary point at the local minimum the bisection landed on and c64 =
 * po64_cost(tau64).  stats[0..2]: max |c - c64| / max(|c64|, 1e-2) for
 * c = po64_cost(topts_pol) (FP64-polished arrival time), c = copts32 (the
 * FP32 cost table) and c = po64_cost(topts32); stats[3..4]: max |t - tau64| /
 * tau64 for topts_pol and topts32; stats[5]: pairs compared.  Tables are
 * [nrows][num-1] in the row layout of fillCoptsTopts. */
void po64_rows_cost_error(const float *samples, int num, const int *rows, int nrows, float tmax,
                          const float *topts32, const float *copts32, const double *topts_pol,
                          double *stats, int nthreads)
{
    double m[5] = {0, 0, 0, 0, 0};
    const int64_t n = (int64_t)nrows * (num - 1);
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#endif
    double m0 = 0, m1 = 0, m2 = 0, m3 = 0, m4 = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) reduction(max : m0, m1, m2, m3, m4) if (nthreads != 1)
#endif
    for (int64_t q = 0; q < n; ++q) {
        const int i = rows[q / (num - 1)];
        const int k = (int)(q % (num - 1));
        const int j = k + (k >= i);
        const float *a = &samples[(int64_t)i * DIM], *b = &samples[(int64_t)j * DIM];
        const double t64 = po64_topt_near(a, b, (double)topts32[q], (double)tmax);
        const double c64 = po64_cost(t64, a, b);
        const double den = fabs(c64) > 1e-2 ? fabs(c64) : 1e-2;
        const double e0 = fabs(po64_cost(topts_pol[q], a, b) - c64) / den;
        const double e1 = fabs((double)copts32[q] - c64) / den;
        const double e2 = fabs(po64_cost((double)topts32[q], a, b) - c64) / den;
        const double e3 = fabs(topts_pol[q] - t64) / t64;
        const double e4 = fabs((double)topts32[q] - t64) / t64;
        m0 = e0 > m0 ? e0 : m0;
        m1 = e1 > m1 ? e1 : m1;
        m2 = e2 > m2 ? e2 : m2;
        m3 = e3 > m3 ? e3 : m3;
        m4 = e4 > m4 ? e4 : m4;
    }
    m[0] = m0;
    m[1] = m1;
    m[2] = m2;
    m[3] = m3;
    m[4] = m4;
    for (int k = 0; k < 5; ++k)
        stats[k] = m[k];
    stats[5] = (double)n;
    (void)nthreads;
}

static int po64_face(const double *v, const double *w, double lambda, int j, const float *lo,
                     const float *hi)
{
    for (int d = 0; d < HALF; ++d) {
        double proj = v[d] + (w[d] - v[d]) * lambda;
        if (d != j && !((double)lo[d] <= proj && proj <= (double)hi[d]))
            return 0;
    }
    return 1;
}

static int po64_seg_valid(const double *v, const double *w, int n_obs, const float *obstacles)
{
    for (int o = 0; o < n_obs; ++o) {
        const float *lo = &obstacles[o * 2 * DIM], *hi = lo + DIM;
        int separated = 0;
        for (int d = 0; d < HALF; ++d) {
            double mn = v[d] < w[d] ? v[d] : w[d];
            double mx = v[d] < w[d] ? w[d] : v[d];
            if (mx <= lo[d] || hi[d] <= mn) {
                separated = 1;
                break;
            }
        }
        if (separated)
            continue;
        for (int d = 0; d < HALF; ++d) {
            double vw = w[d] - v[d];
            double lambda = (v[d] < lo[d]) ? (lo[d] - v[d]) / vw : (hi[d] - v[d]) / vw;
            if (po64_face(v, w, lambda, d, lo, hi))
                return 0;
        }
    }
    return 1;
}

int po64_mccp_hits(const float *obstacles, int n_obs, const float *xcomb, int tmax_steps,
                   int n_mc, double offset_mult, const double *path, int T)
{
    int total = 0;
    for (int p = 0; p < n_mc; ++p) {
        for (int t = 0; t < T - 2; ++t) {
            double v[HALF], w[HALF];
            for (int d = 0; d < HALF; ++d) {
                v[d] = path[t * DIM + d] +
                       xcomb[(int64_t)p * HALF * tmax_steps + t * HALF + d] * offset_mult;
                w[d] = path[(t + 1) * DIM + d] +
                       xcomb[(int64_t)p * HALF * tmax_steps + (t + 1) * HALF + d] * offset_mult;
            }
            if (!po64_seg_valid(v, w, n_obs, obstacles)) {
                total++;
                break;
            }
        }
    }
    return total;
}

/* cubic state at time t along the edge x0 -> x1 of duration tau (FP64,
 * same polynomial as pathPoint, no gravity) */
static void po64_path_point(double t, double tau, const float *x0, const float *x1, double *x)
{
    for (int d = 0; d < HALF; ++d) {
        double p0 = x0[d], v0 = x0[d + 3], p1 = x1[d], v1 = x1[d + 3];
        double c3 = (2 * p0 - 2 * p1 + (v0 + v1) * tau) / (tau * tau * tau);
        double c2 = (3 * p0 - 3 * p1 + (2 * v0 + v1) * tau) / (tau * tau);
        x[d] = p0 + v0 * t - c2 * t * t + c3 * t * t * t;
        x[d + 3] = v0 - 2 * c2 * t + 3 * c3 * t * t;
    }
}

int po64_find_optimal_path(double dt, double *split, const float *xs, int n_wp, int cap)
{
    double toff = 0, t = 0;
    int rows = 0;
    for (int i = 0; i < n_wp - 1; ++i) {
        const float *x0 = &xs[i * DIM], *x1 = &xs[(i + 1) * DIM];
        double topt = po64_topt(x0, x1, 5.0);
        int T = (int)((topt - toff) / dt + 1);
        if (rows + (T > 0 ? T : 0) > cap)
            return -1;
        for (int k = 0; k < T; ++k) {
            t = dt * k + toff;
            po64_path_point(t, topt, x0, x1, &split[(rows + k) * DIM]);
        }
        rows += T;
        toff = (t - topt) + dt;
    }
    return rows;
}

/* ======================================================================
 * Smoothing and controller discretisation
 * ==================================================================== */

/* Smoothing -- PUMP.cu:406-540: the solution's node states xs are blended
 * towards the states that the obstacle-free optimal path from xs[0] to
 * xs[n-1] (findOptimalPath at dt/10) passes at the same time fractions,
 * xsSmooth = (1-alpha)*xs + alpha*xsOpt, with alpha bisected on the full-MC
 * CP (alpha = 1 first; CP < target raises the lower bound, else the upper
 * one), at most 30 iterations, stopping once target - CP lies in (0, 2e-4).
 * The kept path is the one whose CP is below the target and closest to it
 * (lower cost on a tie), starting from the solution itself.
 * Controller discretisation -- PUMP.cu:543-568: each edge of the last
 * iterate xsSmooth (the reference discretises the last iterate, not the
 * kept one) gives (int)(tau/0.05 + 1) rows of findDiscretizedPath, with tau
 * from a bisection with tmax = 2; the row count is accumulated as an int
 * with float additions (discretizedPathLength += tau/controlDt + 1).
 * Outputs: the kept path (path.txt) and its cost and CP; the last iterate;
 * the discretised rows (discretizedPath.txt); a trace of alpha, CP and cost
 * per iteration.  Returns 0, or -1 when a buffer capacity is exceeded (the
 * reference writes past its fixed arrays there).
 * Pinned by: an obstacle-free problem (every CP is 0, alpha stays 1, and the
 * kept cost is min(solution cost, the alpha = 1 cost)); the kept path equals
 * findOptimalPath of the kept iterate and the costs equal the sum of the
 * edge costs recomputed from the primitives; the discretised rows equal the
 * per-edge findDiscretizedPath rows. */
int po_smooth(const float *xs, int n, float soln_cost, float soln_cp, const float *obstacles,
              int n_obs, const float *xcomb, int tmax_steps, int n_mc, float offset_mult, float dt,
              float cp_target, int path_cap, float *best_path, int *best_rows, float *best_cost_cp,
              float *xs_smooth, int disc_cap, float *disc, int *disc_rows, float *trace, int *itrs)
{
    const int dt_mult = 10;
    const int opt_cap = path_cap * dt_mult;
    float *opt = (float *)malloc(sizeof(float) * (size_t)opt_cap * DIM);
    float *path = (float *)malloc(sizeof(float) * (size_t)path_cap * DIM);
    float *xs_opt = (float *)malloc(sizeof(float) * (size_t)n * DIM);
    float *cum = (float *)malloc(sizeof(float) * (size_t)n);
    int rc = 0;
    int opt_len = 0, len = 0;
    /* the kept path starts as the solution path */
    po_find_optimal_path(dt, best_path, xs, n, &len, path_cap);
    if (len < 0) {
        rc = -1;
        goto done;
    }
    *best_rows = len;
    float best_cost = soln_cost, best_cp = soln_cp;
    /* the obstacle-free optimal path from the first to the last state */
    {
        float ends[2 * DIM];
        for (int d = 0; d < DIM; ++d) {
            ends[d] = xs[d];
            ends[DIM + d] = xs[(n - 1) * DIM + d];
        }
        po_find_optimal_path(dt / dt_mult, opt, ends, 2, &opt_len, opt_cap);
        if (opt_len < 0) {
            rc = -1;
            goto done;
        }
    }
    cum[0] = 0;
    for (int i = 0; i < n - 1; ++i)
        cum[i + 1] = cum[i] + po_topt_bisection(&xs[i * DIM], &xs[(i + 1) * DIM], 2);
    for (int i = 0; i < n; ++i) {
        int idx = (int)((float)(opt_len - 1) * cum[i] / cum[n - 1]);
        for (int d = 0; d < DIM; ++d)
            xs_opt[i * DIM + d] = opt[idx * DIM + d];
    }
    {
        float alpha = 1.0f, alpha_u = 1.0f, alpha_l = 0.0f;
        int k = 0;
        for (k = 0; k < 30;) {
            for (int i = 0; i < n * DIM; ++i)
                xs_smooth[i] = (1 - alpha) * xs[i] + alpha * xs_opt[i];
            po_find_optimal_path(dt, path, xs_smooth, n, &len, path_cap);
            if (len < 0 || len > tmax_steps) {
                rc = -1;
                goto done;
            }
            float cost = 0;
            for (int i = 0; i < n - 1; ++i) {
                float tau = po_topt_bisection(&xs_smooth[i * DIM], &xs_smooth[(i + 1) * DIM], 2);
                cost += po_cost(tau, &xs_smooth[i * DIM], &xs_smooth[(i + 1) * DIM]);
            }
            int hits = po_mccp_hits(obstacles, n_obs, xcomb, tmax_steps, n_mc, offset_mult, path,
                                    len, 0);
            float cp = ((float)hits) / ((float)n_mc);
            trace[k * 3 + 0] = alpha;
            trace[k * 3 + 1] = cp;
            trace[k * 3 + 2] = cost;
            if (cp < cp_target)
                alpha_l = alpha;
            if (cp >= cp_target)
                alpha_u = alpha;
            alpha = (alpha_l + alpha_u) / 2;
            int satisfies = cp < cp_target;
            int closer = cp > best_cp;
            int equal_lower = (cp == best_cp && cost < best_cost);
            if (satisfies && (closer || equal_lower)) {
                memcpy(best_path, path, sizeof(float) * (size_t)len * DIM);
                *best_rows = len;
                best_cost = cost;
                best_cp = cp;
            }
            if (fabsf(cp_target - cp) < 0.0002f && cp_target - cp > 0) {
                ++k;
                break;
            }
            ++k;
        }
        *itrs = k;
    }
    best_cost_cp[0] = best_cost;
    best_cost_cp[1] = best_cp;
    /* controller discretisation of the last iterate */
    {
        const float control_dt = 0.05f;
        int rows = 1;
        for (int i = 0; i < n - 1; ++i) {
            float tau = po_topt_bisection(&xs_smooth[i * DIM], &xs_smooth[(i + 1) * DIM], 2);
            rows += tau / control_dt + 1;
        }
        if (rows > disc_cap) {
            rc = -1;
            goto done;
        }
        memset(disc, 0, sizeof(float) * (size_t)rows * DIM);
        int cur = 0;
        for (int i = 0; i < n - 1; ++i) {
            float tau = po_topt_bisection(&xs_smooth[i * DIM], &xs_smooth[(i + 1) * DIM], 2);
            int nw = tau / control_dt + 1;
            if (cur + nw + 1 > rows) {
                rc = -1;
                goto done;
            }
            po_find_discretized_path(&disc[(size_t)cur * DIM], &xs_smooth[i * DIM],
                                     &xs_smooth[(i + 1) * DIM], nw);
            cur += nw;
        }
        *disc_rows = rows;
    }
done:
    free(opt);
    free(path);
    free(xs_opt);
    free(cum);
    return rc;
}

/* createSamplesHalton -- sampler.cu:20-57: dimension d of sample n is the
 * radical inverse of n in base primes[d + skip], in the reference's float /
 * int steps, scaled to [lo, hi]; sample 0 = init, sample num-1 = goal.
 * Pinned by: base 2 is exact (the van der Corput values are dyadic), every
 * other base within a few ulps of the exact radical inverse. */
int po_halton_samples(int num, int skip, const float *init, const float *goal, const float *lo,
                      const float *hi, float *samples)
{
    static const int primes[25] = {2,  3,  5,  7,  11, 13, 17, 19, 23, 29, 31, 37, 41,
                                   43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97};
    if (skip < 0 || skip + DIM > 25)
        return -1;
    for (int d = 0; d < DIM; ++d) {
        const int b = primes[d + skip];
        for (int n = 0; n < num; ++n) {
            float hn = 0, f = 1 / ((float)b);
            int n0 = n;
            while (n0 > 0) {
                float n1 = n0 / b;
                int r = n0 - n1 * b;
                hn += f * r;
                f = f / b;
                n0 = n1;
            }
            samples[n * DIM + d] = hn * (hi[d] - lo[d]) + lo[d];
        }
    }
    for (int d = 0; d < DIM; ++d) {
        samples[d] = init[d];
        samples[(num - 1) * DIM + d] = goal[d];
    }
    return 0;
}

/* ======================================================================
 * Group Marching Tree and Euclidean neighbours (not in the reference's
 * build; listed API surface)
 * ==================================================================== */

/* calculateNumberNN + setupPrecompNN + precompNN -- precomp.cu:3-76: the
 * samples within squared distance r2 (all 6 dimensions, sum in dimension
 * order), in sample order, with their distance sqrtf(d2); -1 / 0 past the
 * last one.  nn_sizes receives every node's full count; at most nn_size are
 * kept.  Pinned by a float32 brute force in the same summation order. */
void po_precomp_nn(const float *samples, int num, float r2, int nn_size, int *nn_sizes,
                   float *distances, int *nn_edges)
{
    for (int node = 0; node < num; ++node) {
        int c = 0;
        for (int k = 0; k < nn_size; ++k) {
            nn_edges[(int64_t)node * nn_size + k] = -1;
            distances[(int64_t)node * nn_size + k] = 0;
        }
        for (int s = 0; s < num; ++s) {
            if (s == node)
                continue;
            float d2 = 0;
            for (int d = 0; d < DIM; ++d) {
                float diff = samples[node * DIM + d] - samples[s * DIM + d];
                d2 += diff * diff;
            }
            if (d2 < r2) {
                if (c < nn_size) {
                    nn_edges[(int64_t)node * nn_size + c] = s;
                    distances[(int64_t)node * nn_size + c] = sqrtf(d2);
                }
                ++c;
            }
        }
        nn_sizes[node] = c;
    }
}

/* GMTwavefront -- GMT.cu:3-261 with lambda = 1, one step at a time:
 * setupArrays; then per iteration findWavefront, the new wavefront,
 * findOptimalConnection (first minimum in come-edge order; a cost of 0 is
 * "unset"), verifyExpansion + removeInvalidExpansions (the connecting edge's
 * segments, found in the CSR rows), updateWavefront; until the goal has a
 * non-zero cost, the new wavefront is empty, or 100 iterations.
 * Pinned by: every reached node's cost is its parent's plus the come-edge
 * distance and its parent was reached earlier; the goal cost is at least the
 * shortest-path cost; on a chain the tree is the chain. */
int po_gmt_wavefront(const float *obstacles, int n_obs, const float *distances_come,
                     const int *nn_go, const int *nn_come, int nn_size, const float *disc,
                     const int64_t *wp_offsets, const int64_t *row_start, const int *edge_dst,
                     const uint8_t *is_free_samples, int num, float *costs, int *edges, int goal,
                     float *cost_goal, int *itrs_out)
{
    uint8_t *wf = (uint8_t *)calloc((size_t)num, 1), *wf_new = (uint8_t *)calloc((size_t)num, 1);
    uint8_t *was = (uint8_t *)calloc((size_t)num, 1), *unv = (uint8_t *)calloc((size_t)num, 1);
    int rc = 0;
    for (int n = 0; n < num; ++n) {
        unv[n] = is_free_samples[n];
        was[n] = !is_free_samples[n];
        wf[n] = n == 0;
        costs[n] = is_free_samples[n] ? 0.0f : -11.0f;
        edges[n] = is_free_samples[n] ? -1 : -2;
    }
    float cg = 0;
    int itrs = 0;
    while (itrs < 100 && cg == 0) {
        ++itrs;
        for (int node = 0; node < num; ++node) {
            if (!wf[node])
                continue;
            was[node] = 1;
            for (int i = 0; i < nn_size; ++i) {
                int nn = nn_go[(int64_t)node * nn_size + i];
                if (nn == -1)
                    break;
                if (unv[nn] && !wf[nn])
                    wf_new[nn] = 1;
            }
        }
        int n_new = 0;
        for (int node = 0; node < num; ++node)
            n_new += wf_new[node];
        if (n_new == 0)
            break;
        for (int node = 0; node < num; ++node) {
            if (!wf_new[node])
                continue;
            for (int i = 0; i < nn_size; ++i) {
                int nn = nn_come[(int64_t)node * nn_size + i];
                if (nn == -1)
                    break;
                if (wf[nn]) {
                    float c = costs[nn] + distances_come[(int64_t)node * nn_size + i];
                    if (costs[node] == 0 || c < costs[node]) {
                        costs[node] = c;
                        edges[node] = nn;
                    }
                }
            }
            int from = edges[node];
            int64_t e = -1;
            for (int64_t q = from >= 0 ? row_start[from] : 0; from >= 0 && q < row_start[from + 1];
                 ++q)
                if (edge_dst[q] == node) {
                    e = q;
                    break;
                }
            if (e < 0) {
                rc = -1;
                goto done;
            }
            int collides = 0;
            for (int64_t g = wp_offsets[e]; g + 1 < wp_offsets[e + 1] && !collides; ++g) {
                float v[HALF], w[HALF], bmin[HALF], bmax[HALF];
                for (int d = 0; d < HALF; ++d) {
                    v[d] = disc[g * DIM + d];
                    w[d] = disc[(g + 1) * DIM + d];
                    bmin[d] = v[d] > w[d] ? w[d] : v[d];
                    bmax[d] = v[d] > w[d] ? v[d] : w[d];
                }
                collides = !po_is_motion_valid(v, w, bmin, bmax, n_obs, obstacles);
            }
            if (collides) {
                costs[node] = 0;
                edges[node] = -1;
                wf_new[node] = 0;
            }
        }
        for (int node = 0; node < num; ++node) {
            if (!unv[node])
                continue;
            if (was[node]) {
                wf[node] = 0;
                unv[node] = 0;
            } else if (wf_new[node]) {
                wf[node] = 1;
                wf_new[node] = 0;
                if (node == goal)
                    cg = costs[node];
            }
        }
    }
    *cost_goal = cg;
    *itrs_out = itrs;
done:
    free(wf);
    free(wf_new);
    free(was);
    free(unv);
    return rc;
}
