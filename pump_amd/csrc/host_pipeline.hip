// host_pipeline.hip -- the host side above the C ABI (host C++; no kernels
// here).  It restates
//   * the roadmap precompute of mainPUMP.cu:248-504,621-681 (samples ->
//     fillCoptsTopts -> rn -> edges -> heuristic -> sampleFree -> freeEdges ->
//     setupAsBs/cacheHalfspaces), every stage an entry point of pump_abi.h
//     and every O(NUM^2) host loop of the reference moved to the device;
//   * the PUMP search consumer of PUMP.cu:3-405 (bucketed wavefront,
//     HSMC propagate + CP, epsilon-dominance, compaction, goal test, goal
//     path selection by full Monte Carlo), with the reference's operation
//     order on the host so that the returned path is the reference's.
// Dense NUM x NUM tables of the reference (nnIdxs, P) are replaced by the
// CSR rows of the edge numbering and per-node lists with the same order:
// out-edge nn of node i is edge row_start[i] + nn, destinations ascending.
#include <algorithm>
#include <cstdio>
#include <chrono>
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

#include "../../include/pump_host.h"
#include "pump_device.hpp"
#include "pump_internal.hpp"
#include "roadmap_internal.hpp"

#pragma clang fp contract(off)

using namespace pump;

namespace pump {
namespace rmb {

int rm_upload(pump_roadmap *rm, const float *samples_h, const float *obstacles_h,
              const float *offsets_time_h, const float *xcomb_h)
{
    const int num = rm->num;
    hipStream_t s = rm->stream;
    TRY(dalloc(&rm->samples, (size_t)num * 6));
    TRY(dalloc(&rm->obstacles, (size_t)rm->n_obs * 12));
    TRYH(hipMemcpyAsync(rm->samples, samples_h, sizeof(float) * num * 6, hipMemcpyHostToDevice, s));
    if (rm->n_obs)
        TRYH(hipMemcpyAsync(rm->obstacles, obstacles_h, sizeof(float) * rm->n_obs * 12,
                            hipMemcpyHostToDevice, s));
    if (rm->owns_noise && offsets_time_h && xcomb_h && rm->n_mc > 0) {
        size_t nt = (size_t)rm->tmax_steps * rm->n_mc * 3;
        TRY(dalloc(&rm->offsets_time, nt));
        TRY(dalloc(&rm->xcomb, nt));
        TRYH(hipMemcpyAsync(rm->offsets_time, offsets_time_h, sizeof(float) * nt,
                            hipMemcpyHostToDevice, s));
        TRYH(hipMemcpyAsync(rm->xcomb, xcomb_h, sizeof(float) * nt, hipMemcpyHostToDevice, s));
        TRY(dalloc(&rm->offsets_split,
                   pump_split_offsets_bytes(rm->tmax_steps, rm->n_mc) / sizeof(uint16_t)));
        TRY(pump_split_offsets(rm->offsets_time, rm->tmax_steps, rm->n_mc, rm->offsets_split, s));
    }
    rm->n_pairs = (int64_t)num * (num - 1);
    TRY(dalloc(&rm->copts, (size_t)rm->n_pairs));
    TRY(dalloc(&rm->topts, (size_t)rm->n_pairs));
    return PUMP_OK;
}

// (a) all ordered pairs, and the optional FP64 polish
int rm_pairs(pump_roadmap *rm)
{
    const pump_params &P = rm->prm;
    EventTimer tm(rm->stream);
    tm.start();
    TRY(fill_copts_topts(rm->samples, rm->num, P.tmax, rm->copts, rm->topts, rm->stream,
                         &rm->info.pair_steps));
    rm->info.ms_pairs = tm.stop();
    return PUMP_OK;
}

// rn, edges, neighbour tables, heuristic, sampleFree, waypoint offsets
int rm_tables(pump_roadmap *rm)
{
    const pump_params &P = rm->prm;
    const int num = rm->num;
    hipStream_t s = rm->stream;
    EventTimer tm(s);
    if (P.polish) {
        TRY(dalloc(&rm->topts64, (size_t)rm->n_pairs));
        TRY(pump_polish_topts(rm->samples, num, P.tmax, rm->topts, rm->topts64, s));
    }

    // rn (mainPUMP.cu:261-264,303-304)
    rm->rn_idx = pump_rn_index(num, P.rn_perc);
    if (rm->rn_idx < 0 || rm->rn_idx >= rm->n_pairs)
        return PUMP_E_RANGE;
    TRY(ensure_ws(rm, std::max(pump_select_rn_workspace(rm->n_pairs), pump_scan_workspace(num + 1))));
    tm.start();
    TRY(pump_select_rn(rm->copts, rm->n_pairs, rm->rn_idx, &rm->rn, rm->ws, rm->ws_bytes, s));
    rm->info.ms_select = tm.stop();

    // edges (mainPUMP.cu:313-348)
    TRY(dalloc(&rm->row_start, (size_t)num + 1));
    tm.start();
    TRY(pump_count_edges(rm->copts, num, rm->rn, rm->row_start, &rm->n_edges, rm->ws, rm->ws_bytes,
                         s));
    const size_t E = (size_t)rm->n_edges;
    TRY(dalloc(&rm->edge_src, E));
    TRY(dalloc(&rm->edge_dst, E));
    TRY(dalloc(&rm->copts_edge, E));
    TRY(dalloc(&rm->topts_edge, E));
    TRY(pump_fill_edges(rm->copts, rm->topts, num, rm->rn, rm->row_start, rm->edge_src,
                        rm->edge_dst, rm->copts_edge, rm->topts_edge, s));
    rm->info.ms_edges_build = tm.stop();

    // maxNNSize and heuristic (mainPUMP.cu:350-369)
    tm.start();
    TRY(dalloc(&rm->in_degree, (size_t)num));
    TRY(pump_max_nn_size(rm->row_start, rm->edge_dst, rm->n_edges, num, rm->in_degree, &rm->max_nn,
                         s));
    TRY(dalloc(&rm->h, (size_t)num));
    TRY(pump_heuristic(rm->samples, num, rm->goal, P.tmax, rm->h, s));
    rm->info.ms_nn = tm.stop();
    if (!P.keep_pairs) {
        TRYH(hipStreamSynchronize(s));
        dfree(rm->copts);
        dfree(rm->topts);
    }

    // sampleFree (mainPUMP.cu:429-439)
    TRY(dalloc(&rm->is_free_samples, (size_t)num));
    tm.start();
    TRY(pump_sample_free(rm->obstacles, rm->n_obs, rm->samples, num, rm->is_free_samples, s));
    rm->info.ms_sample_free = tm.stop();

    // waypoint offsets (mainPUMP.cu:441-463)
    TRY(dalloc(&rm->wp_offsets, E + 1));
    TRY(ensure_ws(rm, pump_scan_workspace((int64_t)E + 1)));
    TRY(pump_waypoint_offsets(rm->topts_edge, rm->n_edges, P.w_fixed, P.wp_dt, P.w_cap,
                              rm->wp_offsets, &rm->n_wp, rm->ws, rm->ws_bytes, s));
    TRY(dalloc(&rm->disc, (size_t)rm->n_wp * 6));
    TRY(dalloc(&rm->is_free_edges, E));
    return PUMP_OK;
}

// the fused (or, on request, split) edge pass
int rm_edge_pass(pump_roadmap *rm)
{
    const pump_params &P = rm->prm;
    hipStream_t s = rm->stream;
    EventTimer tm(s);
    tm.start();
    if (P.fused_edges) {
        TRY(pump_edges_discretize_check(rm->samples, rm->edge_src, rm->edge_dst, rm->topts_edge,
                                        rm->n_edges, rm->wp_offsets, 0, rm->obstacles, rm->n_obs,
                                        rm->disc, rm->is_free_edges, s));
    } else {
        TRY(pump_discretize_edges(rm->samples, rm->edge_src, rm->edge_dst, rm->topts_edge,
                                  rm->n_edges, rm->wp_offsets, 0, rm->disc, s));
        TRY(pump_free_edges_var(rm->obstacles, rm->n_obs, rm->wp_offsets, rm->disc,
                                rm->is_free_edges, rm->n_edges, s));
    }
    rm->info.ms_edge_pass = tm.stop();
    return PUMP_OK;
}

// half-space buffers (mainPUMP.cu:465-476)
int rm_alloc_halfspaces(pump_roadmap *rm)
{
    const pump_params &P = rm->prm;
    rm->max_hs = rm->n_obs > P.max_hs ? P.max_hs : rm->n_obs;
    rm->maxd2 = pump_maxd2(P.offset_mult);
    const size_t hs = (size_t)rm->n_wp * (size_t)(rm->max_hs > 0 ? rm->max_hs : 1);
    TRY(dalloc(&rm->as, hs * 3));
    TRY(dalloc(&rm->bs, hs));
    TRY(dalloc(&rm->count, (size_t)rm->n_wp));
    TRY(dalloc(&rm->edge_hs_mask, (size_t)rm->n_edges));
    if (rm->max_hs == 0)
        TRY(pump_setup_as_bs((int64_t)hs, rm->bs, rm->as, rm->stream));
    return PUMP_OK;
}

// half-spaces (mainPUMP.cu:477-504)
int rm_halfspaces(pump_roadmap *rm)
{
    hipStream_t s = rm->stream;
    EventTimer tm(s);
    tm.start();
    TRY(pump_cache_halfspaces(rm->n_edges, rm->disc, rm->is_free_edges, rm->n_obs, rm->max_hs,
                              rm->obstacles, 0, rm->wp_offsets, rm->as, rm->bs, rm->count,
                              rm->maxd2, 1, rm->edge_hs_mask, s));
    rm->info.ms_hs = tm.stop();
    return PUMP_OK;
}

int rm_finish(pump_roadmap *rm, std::chrono::steady_clock::time_point t_begin)
{
    TRYH(hipStreamSynchronize(rm->stream));
    auto t_end = std::chrono::steady_clock::now();
    pump_roadmap_info &I = rm->info;
    I.num = rm->num;
    I.n_obs = rm->n_obs;
    I.max_nn = rm->max_nn;
    I.max_hs = rm->max_hs;
    I.n_mc = rm->n_mc;
    I.tmax_steps = rm->tmax_steps;
    I.n_pairs = rm->n_pairs;
    I.rn_idx = rm->rn_idx;
    I.n_edges = rm->n_edges;
    I.n_waypoints = rm->n_wp;
    I.rn = rm->rn;
    I.maxd2 = rm->maxd2;
    I.ms_total = std::chrono::duration<float, std::milli>(t_end - t_begin).count();
    const size_t E = (size_t)rm->n_edges;
    std::vector<uint8_t> f(E);
    if (E)
        TRYH(hipMemcpy(f.data(), rm->is_free_edges, E, hipMemcpyDeviceToHost));
    int64_t c = 0;
    for (uint8_t v : f)
        c += v;
    I.n_free_edges = c;
    return PUMP_OK;
}

} // namespace rmb
} // namespace pump

using namespace pump::rmb;

namespace {

int build(pump_roadmap *rm, const float *samples_h, const float *obstacles_h,
          const float *offsets_time_h, const float *xcomb_h)
{
    const auto t_begin = std::chrono::steady_clock::now();
    TRY(rm_upload(rm, samples_h, obstacles_h, offsets_time_h, xcomb_h));
    TRY(rm_pairs(rm));
    TRY(rm_tables(rm));
    TRY(rm_edge_pass(rm));
    TRY(rm_alloc_halfspaces(rm));
    TRY(rm_halfspaces(rm));
    return rm_finish(rm, t_begin);
}

} // namespace

namespace pump {
static thread_local char g_last_error[256] = "";
void note_hip_error(hipError_t e, const char *where, int line)
{
    const char *base = where;
    for (const char *q = where; *q; ++q)
        if (*q == '/')
            base = q + 1;
    snprintf(g_last_error, sizeof(g_last_error), "%s (%d) at %s:%d", hipGetErrorString(e), (int)e,
             base, line);
}
} // namespace pump

extern "C" const char *pump_last_error(void) { return pump::g_last_error; }

extern "C" const char *pump_status_string(int status)
{
    switch (status) {
    case PUMP_OK: return "ok";
    case PUMP_E_INVALID: return "invalid argument";
    case PUMP_E_HIP: return "HIP runtime error";
    case PUMP_E_CAPACITY: return "buffer too small";
    case PUMP_E_NOMEM: return "device allocation failed";
    case PUMP_E_RANGE: return "argument outside the supported range";
    default: return "unknown status";
    }
}

extern "C" int pump_abi_version(void) { return 1; }

extern "C" void pump_params_defaults(pump_params *p)
{
    std::memset(p, 0, sizeof(*p));
    p->tmax = 5.0f;        // mainPUMP.cu:266
    p->rn_perc = 0.05f;    // mainPUMP.cu:261
    p->w_fixed = 9;        // numDisc = 8 (mainPUMP.cu:57)
    p->wp_dt = 0.05f;
    p->w_cap = 64;
    p->offset_mult = 0.5f; // mainPUMP.cu:62
    p->max_hs = 5;         // preMaxHSCount (mainPUMP.cu:69)
    p->dt = 0.05f;         // mainPUMP.cu:60
    p->n_p = 128;          // numHSMCParticles (mainPUMP.cu:68)
    p->lambda = 0.5f;      // mainPUMP.cu:66
    p->num_buckets = 2;    // mainPUMP.cu:67
    p->eps_cost = 0.0002f; // mainPUMP.cu:71
    p->eps_cp = 0.0002f;   // mainPUMP.cu:72
    p->cp_target = 0.01f;
    p->cp_factor = 2.0f;
    p->max_itrs = 20;      // PUMP.cu:34
    p->max_path_count = 0;
    p->validity_bits = 0;
    p->precision = 0;
    p->fused_edges = 1;
    p->keep_pairs = 0;
    p->polish = 0;
    p->path_cap = 4096;
    p->mccp_batch = 16;
    p->search_device = 1;
}

extern "C" int pump_roadmap_build(const float *samples_h, int num, const float *obstacles_h,
                                  int n_obs, const float *goal_h, const float *offsets_time_h,
                                  const float *xcomb_h, int n_mc, int tmax_steps,
                                  const pump_params *prm, pump_roadmap **out)
{
    if (!samples_h || num < 2 || n_obs < 0 || (n_obs > 0 && !obstacles_h) || !goal_h || !prm ||
        !out)
        return PUMP_E_INVALID;
    pump_roadmap *rm = new pump_roadmap();
    rm->num = num;
    rm->n_obs = n_obs;
    rm->n_mc = n_mc;
    rm->tmax_steps = tmax_steps;
    rm->prm = *prm;
    std::memcpy(rm->goal, goal_h, sizeof(rm->goal));
    rm->samples_h.assign(samples_h, samples_h + (size_t)num * 6);
    hipError_t ce = hipStreamCreate(&rm->stream);
    if (ce != hipSuccess) {
        note_hip_error(ce, __FILE__, __LINE__);
        delete rm;
        return PUMP_E_HIP;
    }
    int r = build(rm, samples_h, obstacles_h, offsets_time_h, xcomb_h);
    if (r != PUMP_OK) {
        pump_roadmap_free(rm);
        *out = nullptr;
        return r;
    }
    *out = rm;
    return PUMP_OK;
}

namespace {

template <typename T> int upload_n(T **dst, const T *src, size_t n, hipStream_t s)
{
    TRY(dalloc(dst, n));
    if (n)
        TRYH(hipMemcpyAsync(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice, s));
    return PUMP_OK;
}

int from_tables(pump_roadmap *rm, const pump_roadmap_tables &t, const float *offsets_time_h,
                const float *xcomb_h)
{
    const auto t_begin = std::chrono::steady_clock::now();
    hipStream_t s = rm->stream;
    const int num = t.num;
    const size_t E = (size_t)t.n_edges;
    // the host tables are checked before anything is uploaded
    if (t.row_start[0] != 0 || t.row_start[num] != t.n_edges || t.wp_offsets[0] != 0)
        return PUMP_E_INVALID;
    std::vector<int> src(E);
    for (int i = 0; i < num; ++i) {
        if (t.row_start[i + 1] < t.row_start[i])
            return PUMP_E_INVALID;
        for (int64_t e = t.row_start[i]; e < t.row_start[i + 1]; ++e) {
            const int d = t.edge_dst[e];
            if (d < 0 || d >= num || d == i || (e > t.row_start[i] && t.edge_dst[e - 1] >= d))
                return PUMP_E_INVALID;
            src[(size_t)e] = i;
        }
    }
    for (size_t e = 0; e < E; ++e)
        if (t.wp_offsets[e + 1] - t.wp_offsets[e] < 2)
            return PUMP_E_INVALID;
    const int64_t n_wp = t.wp_offsets[E];
    const int mh = t.max_hs;
    // the per-edge waypoint summary the batched HSMC pass reads (bit min(k, 63):
    // waypoint k stores a half-space whose b is not >= 1000) and the stored counts
    std::vector<uint64_t> mask(E, 0);
    std::vector<int> count((size_t)n_wp, 0);
    for (size_t e = 0; e < E; ++e)
        for (int64_t g = t.wp_offsets[e]; g < t.wp_offsets[e + 1]; ++g) {
            int live = 0;
            for (int h = 0; h < mh; ++h)
                live += !(t.bs[g * mh + h] >= 1000.0f);
            count[(size_t)g] = live;
            if (live) {
                const int64_t k = g - t.wp_offsets[e];
                mask[e] |= 1ull << (k < 63 ? k : 63);
            }
        }
    TRY(rm_upload(rm, t.samples, t.obstacles, offsets_time_h, xcomb_h));
    dfree(rm->copts); // no pair tables: the roadmap starts from its edges
    dfree(rm->topts);
    rm->n_pairs = 0;
    rm->rn = t.rn;
    rm->n_edges = t.n_edges;
    rm->n_wp = n_wp;
    rm->max_hs = mh;
    rm->maxd2 = pump_maxd2(rm->prm.offset_mult);
    TRY(upload_n(&rm->row_start, t.row_start, (size_t)num + 1, s));
    TRY(upload_n(&rm->edge_src, (const int *)src.data(), E, s));
    TRY(upload_n(&rm->edge_dst, t.edge_dst, E, s));
    TRY(upload_n(&rm->copts_edge, t.copts_edge, E, s));
    TRY(upload_n(&rm->topts_edge, t.topts_edge, E, s));
    TRY(upload_n(&rm->h, t.h, (size_t)num, s));
    TRY(upload_n(&rm->is_free_edges, t.is_free_edges, E, s));
    TRY(upload_n(&rm->is_free_samples, t.is_free_samples, (size_t)num, s));
    TRY(upload_n(&rm->wp_offsets, t.wp_offsets, E + 1, s));
    TRY(upload_n(&rm->as, t.as, (size_t)n_wp * mh * 3, s));
    TRY(upload_n(&rm->bs, t.bs, (size_t)n_wp * mh, s));
    TRY(upload_n(&rm->count, (const int *)count.data(), (size_t)n_wp, s));
    TRY(upload_n(&rm->edge_hs_mask, (const uint64_t *)mask.data(), E, s));
    TRY(dalloc(&rm->in_degree, (size_t)num));
    TRY(pump_max_nn_size(rm->row_start, rm->edge_dst, rm->n_edges, num, rm->in_degree, &rm->max_nn,
                         s));
    return rm_finish(rm, t_begin);
}

} // namespace

extern "C" int pump_roadmap_from_tables(const pump_roadmap_tables *t, const float *offsets_time_h,
                                        const float *xcomb_h, int n_mc, int tmax_steps,
                                        const pump_params *prm, pump_roadmap **out)
{
    if (!t || !out || !prm || t->num < 2 || !t->samples || t->n_obs < 0 ||
        (t->n_obs > 0 && !t->obstacles) || t->n_edges < 0 || !t->row_start || !t->wp_offsets ||
        !t->h || !t->is_free_samples || t->max_hs < 1 ||
        (t->n_edges > 0 && (!t->edge_dst || !t->copts_edge || !t->topts_edge ||
                            !t->is_free_edges || !t->as || !t->bs)) ||
        !offsets_time_h || !xcomb_h || n_mc < 1 || tmax_steps < 1)
        return PUMP_E_INVALID;
    if (t->n_edges > 0x7fffffff)
        return PUMP_E_RANGE;
    *out = nullptr;
    pump_roadmap *rm = new pump_roadmap();
    rm->num = t->num;
    rm->n_obs = t->n_obs;
    rm->n_mc = n_mc;
    rm->tmax_steps = tmax_steps;
    rm->prm = *prm;
    std::memcpy(rm->goal, t->samples + (size_t)(t->num - 1) * 6, sizeof(rm->goal));
    rm->samples_h.assign(t->samples, t->samples + (size_t)t->num * 6);
    hipError_t ce = hipStreamCreate(&rm->stream);
    if (ce != hipSuccess) {
        note_hip_error(ce, __FILE__, __LINE__);
        delete rm;
        return PUMP_E_HIP;
    }
    int r = from_tables(rm, *t, offsets_time_h, xcomb_h);
    if (r != PUMP_OK) {
        pump_roadmap_free(rm);
        return r;
    }
    *out = rm;
    return PUMP_OK;
}

extern "C" int pump_roadmap_info_get(const pump_roadmap *rm, pump_roadmap_info *info)
{
    if (!rm || !info)
        return PUMP_E_INVALID;
    *info = rm->info;
    return PUMP_OK;
}

void pump::rmb::rm_release(pump_roadmap *rm)
{
    if (!rm)
        return;
    if (rm->stream)
        hipStreamSynchronize(rm->stream);
    dfree(rm->samples);
    dfree(rm->obstacles);
    dfree(rm->copts);
    dfree(rm->topts);
    dfree(rm->topts64);
    dfree(rm->row_start);
    dfree(rm->edge_src);
    dfree(rm->edge_dst);
    dfree(rm->copts_edge);
    dfree(rm->topts_edge);
    dfree(rm->in_degree);
    dfree(rm->h);
    dfree(rm->wp_offsets);
    dfree(rm->disc);
    dfree(rm->is_free_edges);
    dfree(rm->is_free_samples);
    dfree(rm->as);
    dfree(rm->bs);
    dfree(rm->count);
    dfree(rm->edge_hs_mask);
    if (rm->owns_noise) {
        dfree(rm->offsets_time);
        dfree(rm->offsets_split);
        dfree(rm->xcomb);
    }
    dfree(rm->ws);
    delete rm->search_ws;
    rm->search_ws = nullptr;
    if (rm->stream && rm->owns_stream)
        hipStreamDestroy(rm->stream);
    delete rm;
}

extern "C" void pump_roadmap_free(pump_roadmap *rm)
{
    if (!rm || rm->in_batch)
        return;
    rm_release(rm);
}

static int d2h(void *dst, const void *src, size_t bytes)
{
    if (!dst || bytes == 0)
        return PUMP_OK;
    if (!src)
        return PUMP_E_INVALID;
    PUMP_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return PUMP_OK;
}

extern "C" int pump_roadmap_copy_pairs(const pump_roadmap *rm, float *copts, float *topts)
{
    if (!rm)
        return PUMP_E_INVALID;
    TRY(d2h(copts, rm->copts, sizeof(float) * rm->n_pairs));
    TRY(d2h(topts, rm->topts, sizeof(float) * rm->n_pairs));
    return PUMP_OK;
}

extern "C" int pump_roadmap_copy_polished(const pump_roadmap *rm, double *topts64)
{
    if (!rm)
        return PUMP_E_INVALID;
    return d2h(topts64, rm->topts64, sizeof(double) * rm->n_pairs);
}

extern "C" int pump_roadmap_copy_edges(const pump_roadmap *rm, int64_t *row_start, int *edge_src,
                                       int *edge_dst, float *copts_edge, float *topts_edge,
                                       float *h)
{
    if (!rm)
        return PUMP_E_INVALID;
    const size_t E = (size_t)rm->n_edges;
    TRY(d2h(row_start, rm->row_start, sizeof(int64_t) * ((size_t)rm->num + 1)));
    TRY(d2h(edge_src, rm->edge_src, sizeof(int) * E));
    TRY(d2h(edge_dst, rm->edge_dst, sizeof(int) * E));
    TRY(d2h(copts_edge, rm->copts_edge, sizeof(float) * E));
    TRY(d2h(topts_edge, rm->topts_edge, sizeof(float) * E));
    TRY(d2h(h, rm->h, sizeof(float) * rm->num));
    return PUMP_OK;
}

extern "C" int pump_roadmap_copy_nn(const pump_roadmap *rm, int *nn_go_edges, int *nn_come_edges,
                                    float *distances_come, int *nn_idxs)
{
    if (!rm)
        return PUMP_E_INVALID;
    const size_t tab = (size_t)rm->num * rm->max_nn;
    int *go = nullptr, *come = nullptr, *idx = nullptr;
    float *dist = nullptr;
    int r = PUMP_OK;
    if (nn_go_edges)
        r = r ? r : dalloc(&go, tab);
    if (nn_come_edges && distances_come) {
        r = r ? r : dalloc(&come, tab);
        r = r ? r : dalloc(&dist, tab);
    }
    if (nn_idxs)
        r = r ? r : dalloc(&idx, (size_t)rm->num * rm->num);
    if (r == PUMP_OK)
        r = pump_nn_tables(rm->row_start, rm->edge_src, rm->edge_dst, rm->copts_edge, rm->n_edges,
                           rm->num, rm->max_nn, go, come, dist, idx, rm->stream);
    if (r == PUMP_OK && hipStreamSynchronize(rm->stream) != hipSuccess)
        r = PUMP_E_HIP;
    if (r == PUMP_OK && go)
        r = d2h(nn_go_edges, go, sizeof(int) * tab);
    if (r == PUMP_OK && come)
        r = d2h(nn_come_edges, come, sizeof(int) * tab);
    if (r == PUMP_OK && dist)
        r = d2h(distances_come, dist, sizeof(float) * tab);
    if (r == PUMP_OK && idx)
        r = d2h(nn_idxs, idx, sizeof(int) * (size_t)rm->num * rm->num);
    dfree(go);
    dfree(come);
    dfree(dist);
    dfree(idx);
    return r;
}

extern "C" int pump_roadmap_copy_flags(const pump_roadmap *rm, int64_t *wp_offsets,
                                       uint8_t *is_free_edges, uint8_t *is_free_samples)
{
    if (!rm)
        return PUMP_E_INVALID;
    TRY(d2h(wp_offsets, rm->wp_offsets, sizeof(int64_t) * ((size_t)rm->n_edges + 1)));
    TRY(d2h(is_free_edges, rm->is_free_edges, (size_t)rm->n_edges));
    TRY(d2h(is_free_samples, rm->is_free_samples, (size_t)rm->num));
    return PUMP_OK;
}

extern "C" int pump_roadmap_copy_waypoints(const pump_roadmap *rm, float *disc)
{
    if (!rm)
        return PUMP_E_INVALID;
    return d2h(disc, rm->disc, sizeof(float) * (size_t)rm->n_wp * 6);
}

extern "C" int pump_roadmap_copy_halfspaces(const pump_roadmap *rm, float *as, float *bs,
                                            int *count_stored)
{
    if (!rm)
        return PUMP_E_INVALID;
    const size_t hs = (size_t)rm->n_wp * (size_t)(rm->max_hs > 0 ? rm->max_hs : 1);
    TRY(d2h(as, rm->as, sizeof(float) * hs * 3));
    TRY(d2h(bs, rm->bs, sizeof(float) * hs));
    TRY(d2h(count_stored, rm->count, sizeof(int) * (size_t)rm->n_wp));
    return PUMP_OK;
}

extern "C" int pump_roadmap_device(const pump_roadmap *rm, pump_roadmap_dev *d)
{
    if (!rm || !d)
        return PUMP_E_INVALID;
    d->samples = rm->samples;
    d->obstacles = rm->obstacles;
    d->copts = rm->copts;
    d->topts = rm->topts;
    d->copts_edge = rm->copts_edge;
    d->topts_edge = rm->topts_edge;
    d->h = rm->h;
    d->disc = rm->disc;
    d->as = rm->as;
    d->bs = rm->bs;
    d->offsets_time = rm->offsets_time;
    d->offsets_split = rm->offsets_split;
    d->xcomb = rm->xcomb;
    d->row_start = rm->row_start;
    d->wp_offsets = rm->wp_offsets;
    d->edge_src = rm->edge_src;
    d->edge_dst = rm->edge_dst;
    d->count_stored = rm->count;
    d->is_free_edges = rm->is_free_edges;
    d->is_free_samples = rm->is_free_samples;
    d->edge_hs_mask = rm->edge_hs_mask;
    return PUMP_OK;
}

// ---------------------------------------------------------------------------
// findOptimalPath (2pBVP.cu:5-36) with the library's arithmetic
// ---------------------------------------------------------------------------
extern "C" float pump_host_topt(const float *x0, const float *x1, float tmax)
{
    CostTerms c;
    cost_terms(c, x0, x1);
    return topt_bisection(c, tmax);
}

extern "C" float pump_host_cost(float tau, const float *x0, const float *x1)
{
    CostTerms c;
    cost_terms(c, x0, x1);
    return cost_eval(c, tau);
}

extern "C" void pump_host_path_point(float t, float tau, const float *x0, const float *x1,
                                     float *x)
{
    path_point(t, tau_powers(tau), x0, x1, x);
}

extern "C" void pump_host_div_rn(const double *a, const double *b, double *q, int64_t n)
{
    for (int64_t i = 0; i < n; ++i)
        q[i] = div_rn(a[i], b[i], 1.0 / b[i]);
}

extern "C" void pump_host_edge_bbox(const float *x0, const float *x1, float tau, float *lo,
                                    float *hi)
{
    edge_bbox(x0, x1, tau, lo, hi);
}

extern "C" void pump_host_div_magic(const uint32_t *n, int64_t count, uint32_t d, uint32_t *q)
{
    const DivMagic dm = div_magic(d);
    for (int64_t i = 0; i < count; ++i)
        q[i] = div_magic_apply(n[i], dm);
}

extern "C" void pump_host_div_sq(const double *a, const float *tau, double *q, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) {
        const double b = (double)tau[i] * (double)tau[i];
        q[i] = div_sq(a[i], b, 1.0 / b);
    }
}

extern "C" int pump_find_optimal_path(float dt, const float *xs, int n_wp, float *split, int cap,
                                      float *total_time)
{
    float tmax = 5.0f, toff = 0.0f, topt = 0.0f, t = 0.0f, ttotal = 0.0f;
    int rows = 0;
    for (int i = 0; i < n_wp - 1; ++i) {
        const float *x0 = &xs[i * 6], *x1 = &xs[(i + 1) * 6];
        topt = pump_host_topt(x0, x1, tmax);
        int T = (int)((topt - toff) / dt + 1.0f);
        if (rows + (T > 0 ? T : 0) > cap)
            return -1;
        TauPowers P = tau_powers(topt);
        for (int k = 0; k < T; ++k) {
            t = dt * (float)k + toff;
            path_point(t, P, x0, x1, &split[(size_t)(rows + k) * 6]);
        }
        rows += T;
        toff = (t - topt) + dt;
        ttotal += topt;
    }
    if (total_time)
        *total_time = ttotal;
    return rows;
}

// ---------------------------------------------------------------------------
// The PUMP search consumer (PUMP.cu:3-405)
// ---------------------------------------------------------------------------
// Goal-path selection (PUMP.cu:354-405): the goal paths sorted by (cost,
// index), then the first whose full-MC CP is below the target.  status is 0
// on success, 1 when no goal path qualifies, negative on a capacity error.
int pump::rmb::select_goal_path(pump_roadmap *rm, const pump_params &P,
                                const std::vector<int> &goal_list, const PathTable &T,
                                pump_search_result *res, int &status)
{
    hipStream_t s = rm->stream;
    // goal-path selection (PUMP.cu:354-405): cheapest first, first whose
    // full-MC CP is below the target
    std::vector<std::pair<float, int>> goal_paths;
    for (int pidx : goal_list)
        goal_paths.emplace_back(T.cost[pidx], pidx);
    std::sort(goal_paths.begin(), goal_paths.end());
    res->num_goal_paths = (int)goal_paths.size();
    status = 1;
    const int cap = P.path_cap;
    const int batch = P.mccp_batch > 0 ? P.mccp_batch : 1;
    SearchWs *W = search_ws(rm);
    if (!W)
        return PUMP_E_NOMEM;
    DevBuf &d_paths = W->sel_paths, &d_T = W->sel_T, &d_hits = W->sel_hits;
    std::vector<float> hpaths, xs;
    std::vector<int> hT, hits;
    std::vector<std::vector<int>> nodes_of;
    StreamDrain drain{s};
    for (size_t k0 = 0; k0 < goal_paths.size() && status == 1; k0 += batch) {
        const int nk = (int)std::min<size_t>(batch, goal_paths.size() - k0);
        hpaths.assign((size_t)nk * cap * 6, 0.0f);
        hT.assign(nk, 0);
        hits.assign(nk, 0);
        nodes_of.assign(nk, {});
        for (int q = 0; q < nk; ++q) {
            int pidx = goal_paths[k0 + q].second;
            std::vector<int> &nodes = nodes_of[q];
            while (pidx != -2) {
                nodes.push_back(T.node[pidx]);
                pidx = T.prev[pidx];
            }
            std::reverse(nodes.begin(), nodes.end());
            xs.resize(nodes.size() * 6);
            for (size_t r = 0; r < nodes.size(); ++r)
                std::memcpy(&xs[r * 6], &rm->samples_h[(size_t)nodes[r] * 6],
                            sizeof(float) * 6);
            float tt = 0;
            int rows = pump_find_optimal_path(P.dt, xs.data(), (int)nodes.size(),
                                              &hpaths[(size_t)q * cap * 6], cap, &tt);
            if (rows < 0 || rows > rm->tmax_steps + 1) {
                status = -5; // candidate path longer than the path / noise buffers
                break;
            }
            hT[q] = rows;
        }
        if (status != 1)
            break;
        TRY(d_paths.ensure(sizeof(float) * hpaths.size()));
        TRY(d_T.ensure(sizeof(int) * nk));
        TRY(d_hits.ensure(sizeof(int) * nk));
        TRYH(hipMemcpyAsync(d_paths.p, hpaths.data(), sizeof(float) * hpaths.size(),
                            hipMemcpyHostToDevice, s));
        TRYH(hipMemcpyAsync(d_T.p, hT.data(), sizeof(int) * nk, hipMemcpyHostToDevice, s));
        TRY(pump_collision_probability_many(rm->obstacles, rm->n_obs, rm->xcomb,
                                            rm->tmax_steps, rm->n_mc, P.offset_mult,
                                            (const float *)d_paths.p, cap, (const int *)d_T.p,
                                            nk, (int *)d_hits.p, 1, s));
        TRYH(hipMemcpyAsync(hits.data(), d_hits.p, sizeof(int) * nk, hipMemcpyDeviceToHost,
                            s));
        TRYH(hipStreamSynchronize(s));
        for (int q = 0; q < nk; ++q) {
            const int pidx = goal_paths[k0 + q].second;
            const float cp = ((float)hits[q]) / ((float)rm->n_mc);
            res->mccp_calls++;
            res->soln_path_idx = pidx;
            res->soln_cp = cp;
            res->soln_cost = T.cost[pidx];
            res->soln_time = T.time[pidx];
            res->soln_hsmc_cp = T.cp[pidx];
            res->soln_len = (int)std::min<size_t>(nodes_of[q].size(), 256);
            for (int r = 0; r < res->soln_len; ++r)
                res->soln_nodes[r] = nodes_of[q][r];
            if (cp < P.cp_target) {
                status = 0;
                break;
            }
        }
    }
    return PUMP_OK;
}

namespace {


struct SearchState {
    std::vector<int64_t> row_start;
    std::vector<int> edge_dst;
    std::vector<float> topts_edge, copts_edge, h;
    std::vector<uint8_t> free_edges, free_samples;
    std::vector<int64_t> wp_offsets;
};

int search(pump_roadmap *rm, const pump_params &P, pump_search_result *res)
{
    auto t0 = std::chrono::steady_clock::now();
    hipStream_t s = rm->stream;
    const int num = rm->num;
    const int n_p = P.n_p;
    if (n_p <= 0 || !rm->offsets_time || !rm->xcomb || rm->n_mc < n_p)
        return PUMP_E_INVALID;
    const int64_t maxp = P.max_path_count > 0
                             ? P.max_path_count
                             : (int64_t)((double)(2 * num) * std::sqrt((double)num));
    const int nb = P.num_buckets;
    const int init_idx = 0, goal_idx = num - 1;
    const bool bits = P.validity_bits != 0;
    const int64_t row_bytes = bits ? (int64_t)((n_p + 63) / 64) * 8 : (int64_t)n_p;

    SearchState S;
    const size_t E = (size_t)rm->n_edges;
    S.row_start.resize((size_t)num + 1);
    S.edge_dst.resize(E);
    S.topts_edge.resize(E);
    S.copts_edge.resize(E);
    S.h.resize(num);
    S.free_edges.resize(E);
    S.free_samples.resize(num);
    TRY(pump_roadmap_copy_edges(rm, S.row_start.data(), nullptr, S.edge_dst.data(),
                                S.copts_edge.data(), S.topts_edge.data(), S.h.data()));
    TRY(pump_roadmap_copy_flags(rm, nullptr, S.free_edges.data(), S.free_samples.data()));
    S.wp_offsets.resize(E + 1);
    TRY(d2h(S.wp_offsets.data(), rm->wp_offsets, sizeof(int64_t) * (E + 1)));

    // device search state: validity rows of stored paths, their times, and
    // staging for the current wavefront
    SearchWs *W = search_ws(rm);
    if (!W)
        return PUMP_E_NOMEM;
    DevBuf &valid = W->valid, &stage = W->stage, &d_time = W->time, &d_cp = W->cp,
           &d_prev = W->wf_prev, &d_edge = W->wf_edge, &d_rows = W->src_rows, &d_clamp = W->clamp,
           &d_ws = W->prop;
    TRY(valid.ensure((size_t)maxp * row_bytes));
    TRY(d_time.ensure(sizeof(float) * (size_t)maxp));
    TRY(d_clamp.ensure(2 * sizeof(unsigned long long)));
    TRYH(hipMemsetAsync(d_clamp.p, 0, 2 * sizeof(unsigned long long), s));
    if (bits) {
        TRY(pump_setup_valid_bits((uint64_t *)valid.p, n_p, 1, s));
    } else {
        TRYH(hipMemsetAsync(valid.p, 1, (size_t)row_bytes, s));
    }
    TRYH(hipMemsetAsync(d_time.p, 0, sizeof(float), s));

    std::vector<int> path_prev((size_t)maxp), path_node((size_t)maxp);
    std::vector<float> path_cost((size_t)maxp), path_time((size_t)maxp), path_cp((size_t)maxp);
    std::vector<int> G((size_t)nb * maxp, -1);
    std::vector<int> sizeG(nb, 0);
    std::vector<int> wf_prev, wf_next, wf_edge, src_rows;
    std::vector<std::vector<int>> Pn((size_t)num);
    std::vector<int> pnew((size_t)num, 0);
    std::vector<int> touched;
    StreamDrain drain{s};

    const float dr = P.lambda * rm->rn;
    const float cp_min_soln = P.cp_target / P.cp_factor;
    const float cp_cutoff = P.cp_target * P.cp_factor;

    int64_t num_paths = 1;
    int gidx = 0;
    bool goal_cond = false, goal_next = false, empty_open = false;
    Pn[init_idx].push_back(0);
    path_prev[0] = -2;
    path_node[0] = init_idx;
    path_cost[0] = 0;
    path_cp[0] = 0;
    path_time[0] = 0;
    sizeG[gidx]++;
    G[0] = 0;
    float cost_threshold = S.h[init_idx];
    int itrs = 0;
    int status = 0;
    float ms_prop = 0;
    EventTimer tm(s);

    while (itrs < P.max_itrs && !goal_cond && !empty_open) {
        ++itrs;
        if (goal_next)
            goal_cond = true;
        wf_prev.clear();
        wf_next.clear();
        wf_edge.clear();
        bool overflow = false;
        for (int g = 0; g < sizeG[gidx] && !overflow; ++g) {
            int pprev = G[(size_t)gidx * maxp + g];
            G[(size_t)gidx * maxp + g] = -1;
            int nprev = path_node[pprev];
            for (int64_t e = S.row_start[nprev]; e < S.row_start[nprev + 1]; ++e) {
                int nnext = S.edge_dst[e];
                if (!S.free_edges[e] || !S.free_samples[nnext])
                    continue;
                wf_prev.push_back(pprev);
                wf_next.push_back(nnext);
                wf_edge.push_back((int)e);
                if ((int64_t)wf_prev.size() >= maxp) {
                    overflow = true;
                    break;
                }
            }
        }
        if (overflow) {
            status = -2;
            break;
        }
        const int n_new = (int)wf_prev.size();
        if (n_new == 0)
            break;
        if (num_paths + n_new >= maxp) {
            status = -3;
            break;
        }
        sizeG[gidx] = 0;
        res->num_considered += n_new;
        int64_t wf_waypoints = 0;
        for (int w = 0; w < n_new; ++w)
            wf_waypoints += S.wp_offsets[wf_edge[w] + 1] - S.wp_offsets[wf_edge[w]];
        res->particle_steps += (int64_t)n_p * wf_waypoints;

        // HSMC step on the device (PUMP.cu:80-160); the workspace holds every
        // (path, waypoint) pair, so the binned form runs however dense the
        // stored half-spaces are
        TRY(stage.ensure((size_t)n_new * row_bytes));
        TRY(d_ws.ensure(pump_propagate_workspace_pairs(n_new, rm->tmax_steps, wf_waypoints)));
        TRY(d_cp.ensure(sizeof(float) * n_new));
        TRY(d_prev.ensure(sizeof(int) * n_new));
        TRY(d_edge.ensure(sizeof(int) * n_new));
        TRYH(hipMemcpyAsync(d_prev.p, wf_prev.data(), sizeof(int) * n_new, hipMemcpyHostToDevice,
                            s));
        TRYH(hipMemcpyAsync(d_edge.p, wf_edge.data(), sizeof(int) * n_new, hipMemcpyHostToDevice,
                            s));
        unsigned long long *clamp = (unsigned long long *)d_clamp.p;
        tm.start();
        if (bits) {
            TRY(pump_propagate_cp_bits(
                n_new, 0, rm->wp_offsets, (const int *)d_edge.p, (const int *)d_prev.p, rm->as,
                rm->bs, rm->edge_hs_mask, (const float *)d_time.p, rm->topts_edge, P.dt,
                rm->max_hs,
                rm->offsets_time, rm->tmax_steps, P.offset_mult, rm->n_mc,
                (const uint64_t *)valid.p, (uint64_t *)stage.p, 0, n_p, (float *)d_cp.p, 0,
                P.precision, rm->offsets_split, clamp + 1, clamp, d_ws.p, d_ws.bytes, s));
        } else {
            TRY(pump_propagate_cp(n_new, 0, rm->wp_offsets, (const int *)d_edge.p,
                                  (const int *)d_prev.p, rm->as, rm->bs, rm->edge_hs_mask,
                                  (const float *)d_time.p,
                                  rm->topts_edge, P.dt, rm->max_hs, rm->offsets_time,
                                  rm->tmax_steps, P.offset_mult, rm->n_mc,
                                  (const uint8_t *)valid.p, (uint8_t *)stage.p, 0, n_p,
                                  (float *)d_cp.p, 0, clamp, d_ws.p, d_ws.bytes, s));
        }
        ms_prop += tm.stop();
        TRYH(hipMemcpyAsync(&path_cp[num_paths], d_cp.p, sizeof(float) * n_new,
                            hipMemcpyDeviceToHost, s));
        TRYH(hipStreamSynchronize(s));

        for (int i = 0; i < n_new; ++i) {
            path_time[num_paths + i] = path_time[wf_prev[i]] + S.topts_edge[wf_edge[i]];
            path_cost[num_paths + i] = path_cost[wf_prev[i]] + S.copts_edge[wf_edge[i]];
        }

        // dominance (PUMP.cu:169-218)
        touched.clear();
        for (int i = 0; i < n_new; ++i) {
            int node = wf_next[i];
            int64_t pidx = num_paths + i;
            if (path_cp[pidx] > cp_cutoff) {
                wf_next[i] = -1;
                continue;
            }
            Pn[node].push_back((int)pidx);
            if (pnew[node]++ == 0)
                touched.push_back(node);
        }
        for (int i = 0; i < n_new; ++i) {
            int node = wf_next[i];
            if (node == -1 || node == goal_idx)
                continue;
            const int pidx = (int)(num_paths + i);
            const std::vector<int> &L = Pn[node];
            for (size_t j = 0; j < L.size(); ++j) {
                int cmp = L[j];
                if (cmp == pidx)
                    continue;
                if (path_cp[cmp] < P.eps_cp + path_cp[pidx] &&
                    path_cost[cmp] < P.eps_cost + path_cost[pidx]) {
                    if (path_cp[cmp] + P.eps_cp >= path_cp[pidx] &&
                        path_cost[cmp] + P.eps_cost >= path_cost[pidx] && pidx < cmp)
                        continue;
                    wf_next[i] = -1;
                    break;
                }
            }
        }
        for (int node : touched) {
            Pn[node].resize(Pn[node].size() - pnew[node]);
            pnew[node] = 0;
        }

        // store good paths (PUMP.cu:220-256)
        src_rows.clear();
        int n_valid = 0;
        for (int i = 0; i < n_new; ++i) {
            int node = wf_next[i];
            int64_t pidx = num_paths + i;
            if (node == -1 || node == 0) {
                wf_next[i] = -1;
                continue;
            }
            int64_t st = num_paths + n_valid;
            path_cp[st] = path_cp[pidx];
            path_time[st] = path_time[pidx];
            path_cost[st] = path_cost[pidx];
            src_rows.push_back(i);
            float dcost = (path_cost[st] + S.h[node] - cost_threshold);
            dcost = std::max(0.0f, dcost);
            int bucket = (((int)(dcost / dr)) + gidx + 1) % nb;
            if (bucket < 0 || bucket >= nb) {
                status = -6; // bucket index outside [0, nb): negative rn
                break;
            }
            G[(size_t)bucket * maxp + sizeG[bucket]] = (int)st;
            sizeG[bucket]++;
            path_prev[st] = wf_prev[i];
            path_node[st] = node;
            Pn[node].push_back((int)st);
            wf_next[i] = -1;
            n_valid++;
        }
        if (status == -6)
            break; // as the device search: stop before this iteration's paths count
        if (n_valid > 0) {
            TRY(d_rows.ensure(sizeof(int) * n_valid));
            TRYH(hipMemcpyAsync(d_rows.p, src_rows.data(), sizeof(int) * n_valid,
                                hipMemcpyHostToDevice, s));
            TRY(pump_gather_rows((const uint8_t *)stage.p, (const int *)d_rows.p, n_valid,
                                 (uint8_t *)valid.p, num_paths, row_bytes, s));
            TRYH(hipMemcpyAsync((float *)d_time.p + num_paths, &path_time[num_paths],
                                sizeof(float) * n_valid, hipMemcpyHostToDevice, s));
        }
        num_paths += n_valid;

        for (int pidx : Pn[goal_idx]) {
            if (path_cp[pidx] <= cp_min_soln) {
                goal_next = true;
                goal_cond = true;
            }
        }
        if (goal_cond)
            break;
        empty_open = true;
        for (int b = 0; b < nb; ++b)
            empty_open = empty_open && (sizeG[b] == 0);
        if (empty_open)
            break;
        gidx = (gidx + 1) % nb;
        cost_threshold += dr;
        if (num_paths >= maxp) {
            status = -4;
            break;
        }
    }
    TRYH(hipStreamSynchronize(s));
    {
        unsigned long long cl[2] = {0, 0};
        TRYH(hipMemcpy(cl, d_clamp.p, sizeof(cl), hipMemcpyDeviceToHost));
        res->clamped_time_idx = (int64_t)cl[0];
        res->band_particles = (int64_t)cl[1];
    }
    res->itrs = itrs;
    res->num_paths = num_paths;
    res->ms_propagate = ms_prop;
    res->soln_path_idx = -1;
    {
        // the stored path table, where pump_search_copy_paths reads it
        const size_t np = (size_t)num_paths;
        TRY(W->prev.ensure(sizeof(int) * np));
        TRY(W->node.ensure(sizeof(int) * np));
        TRY(W->cost.ensure(sizeof(float) * np));
        TRY(W->cpv.ensure(sizeof(float) * np));
        TRYH(hipMemcpyAsync(W->prev.p, path_prev.data(), sizeof(int) * np, hipMemcpyHostToDevice,
                            s));
        TRYH(hipMemcpyAsync(W->node.p, path_node.data(), sizeof(int) * np, hipMemcpyHostToDevice,
                            s));
        TRYH(hipMemcpyAsync(W->cost.p, path_cost.data(), sizeof(float) * np,
                            hipMemcpyHostToDevice, s));
        TRYH(hipMemcpyAsync(W->time.p, path_time.data(), sizeof(float) * np,
                            hipMemcpyHostToDevice, s));
        TRYH(hipMemcpyAsync(W->cpv.p, path_cp.data(), sizeof(float) * np, hipMemcpyHostToDevice,
                            s));
        TRYH(hipStreamSynchronize(s));
        W->last_num_paths = num_paths;
    }

    if (status == 0) {
        PathTable T{path_prev.data(), path_node.data(), path_cost.data(), path_time.data(),
                    path_cp.data()};
        TRY(select_goal_path(rm, P, Pn[goal_idx], T, res, status));
    }
    res->status = status;
    res->ms_total =
        std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return PUMP_OK;
}

} // namespace

extern "C" int pump_search_copy_paths(const pump_roadmap *rm, int64_t n, int *prev, int *node,
                                      float *cost, float *time, float *cp)
{
    if (!rm || n < 0)
        return PUMP_E_INVALID;
    const pump::rmb::SearchWs *W = rm->search_ws;
    if (n == 0)
        return PUMP_OK;
    if (!W || n > W->last_num_paths)
        return PUMP_E_RANGE;
    const size_t k = (size_t)n;
    TRY(d2h(prev, W->prev.p, sizeof(int) * k));
    TRY(d2h(node, W->node.p, sizeof(int) * k));
    TRY(d2h(cost, W->cost.p, sizeof(float) * k));
    TRY(d2h(time, W->time.p, sizeof(float) * k));
    TRY(d2h(cp, W->cpv.p, sizeof(float) * k));
    return PUMP_OK;
}

extern "C" int pump_search_run(pump_roadmap *rm, const pump_params *prm, pump_search_result *res)
{
    if (!rm || !prm || !res)
        return PUMP_E_INVALID;
    std::memset(res, 0, sizeof(*res));
    if (prm->search_device)
        return pump::rmb::search_device(rm, *prm, res);
    return search(rm, *prm, res);
}
