// roadmap_internal.hpp -- the roadmap object of the host pipeline and the
// phases of its build, shared by the single-instance pipeline
// (host_pipeline.hip) and the batched multi-instance build (batch.hip).
#pragma once

#include <chrono>
#include <new>
#include <vector>

#include "../../include/pump_host.h"
#include "pump_internal.hpp"

namespace pump {
namespace rmb {
struct SearchWs;
}
} // namespace pump

struct pump_roadmap {
    int num = 0, n_obs = 0, n_mc = 0, tmax_steps = 0;
    pump_params prm{};

    float goal[6] = {0, 0, 0, 0, 0, 0};
    std::vector<float> samples_h;
    hipStream_t stream = nullptr;
    float *samples = nullptr, *obstacles = nullptr, *copts = nullptr, *topts = nullptr;
    double *topts64 = nullptr;
    int64_t n_pairs = 0, rn_idx = 0;
    float rn = 0;
    int64_t *row_start = nullptr;
    int *edge_src = nullptr, *edge_dst = nullptr;
    float *copts_edge = nullptr, *topts_edge = nullptr;
    int64_t n_edges = 0;
    int *in_degree = nullptr;
    int max_nn = 0;
    float *h = nullptr;
    int64_t *wp_offsets = nullptr;
    int64_t n_wp = 0;
    float *disc = nullptr;
    uint8_t *is_free_edges = nullptr, *is_free_samples = nullptr;
    float *as = nullptr, *bs = nullptr;
    int *count = nullptr;
    uint64_t *edge_hs_mask = nullptr;
    int max_hs = 0;
    float maxd2 = 0;
    float *offsets_time = nullptr, *xcomb = nullptr;
    uint16_t *offsets_split = nullptr; // particle vectors of the split-BF16 test
    void *ws = nullptr;
    size_t ws_bytes = 0;
    // device buffers of the search, kept between searches (allocated on first use)
    pump::rmb::SearchWs *search_ws = nullptr;
    pump_roadmap_info info{};
    // a roadmap of a batch borrows the batch's stream and noise tables
    bool owns_stream = true, owns_noise = true;
    bool in_batch = false; // released by the batch, not by pump_roadmap_free
};

namespace pump {
namespace rmb {

template <typename T> inline int dalloc(T **p, size_t n)
{
    *p = nullptr;
    if (n == 0)
        n = 1;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * n);
    if (e != hipSuccess) {
        *p = nullptr;
        return PUMP_E_NOMEM;
    }
    return PUMP_OK;
}

template <typename T> inline void dfree(T *&p)
{
    if (p)
        hipFree((void *)p);
    p = nullptr;
}

inline int ensure_ws(pump_roadmap *rm, size_t bytes)
{
    if (rm->ws_bytes >= bytes)
        return PUMP_OK;
    dfree(rm->ws);
    char *w = nullptr;
    int r = dalloc(&w, bytes);
    rm->ws = w;
    rm->ws_bytes = r == PUMP_OK ? bytes : 0;
    return r;
}

// A device buffer that grows on demand: to twice the request, and to at least
// 4 MiB, so that a search whose wavefront grows from a few paths to hundreds
// of thousands reallocates each buffer a few times, not dozens (each hipFree
// also waits for the device).
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t b)
    {
        if (b <= bytes)
            return PUMP_OK;
        if (p)
            hipFree(p);
        p = nullptr;
        bytes = 0;
        size_t want = b * 2 > ((size_t)4 << 20) ? b * 2 : ((size_t)4 << 20);
        if (hipMalloc(&p, want) != hipSuccess) {
            // the failed attempt leaves its error as the thread's last error:
            // clear it, so that a later launch check does not report it
            (void)hipGetLastError();
            if (hipMalloc(&p, b) != hipSuccess) {
                (void)hipGetLastError();
                p = nullptr;
                return PUMP_E_NOMEM;
            }
            want = b;
        }
        bytes = want;
        return PUMP_OK;
    }
    ~DevBuf()
    {
        if (p)
            hipFree(p);
    }
};

// The search's device buffers. The roadmap keeps them from one search to the
// next: a repeated query on the same roadmap then makes no hipMalloc/hipFree
// once the buffers have grown to its sizes (each hipFree of a grown buffer
// also waits for the device). Every search initialises what it reads, so
// the contents a previous search left behind are never used. The host- and
// device-bookkeeping searches share the buffers whose meaning is the same;
// searches on one roadmap run one at a time (its one stream).
struct SearchWs {
    // free out-edges to free nodes per node (offsets, targets), scan workspace
    DevBuf voff, vedge, scan;
    // path table, per-node path lists, bucket lists, validity rows
    DevBuf prev, node, cost, time, cpv, next, head, open, valid, flags, steps, clamp;
    // per-iteration wavefront state
    DevBuf cnt, wf_prev, wf_edge, wf_next, stage, cp, new_time, new_cost, keep, node_cnt,
        node_cur, grp, surv, src_rows, bkt, bflag, prop, row_src;
    // goal-path selection: candidate paths, their lengths and MC hit counts
    DevBuf sel_paths, sel_T, sel_hits;
    // rows of the path table (prev, node, cost, time, cpv) the last search stored
    int64_t last_num_paths = 0;
};

// the roadmap's search buffers, created on first use (nullptr when out of host memory)
inline SearchWs *search_ws(pump_roadmap *rm)
{
    if (!rm->search_ws)
        rm->search_ws = new (std::nothrow) SearchWs();
    return rm->search_ws;
}

// Waits for the stream when it leaves scope. Declared after the host buffers
// that queued copies read or write, it keeps them alive until those copies
// are done on every return path, the error returns included.
struct StreamDrain {
    hipStream_t s;
    ~StreamDrain() { (void)hipStreamSynchronize(s); }
};

struct EventTimer {
    hipEvent_t a{}, b{};
    hipStream_t s;
    explicit EventTimer(hipStream_t st) : s(st)
    {
        hipEventCreate(&a);
        hipEventCreate(&b);
    }
    ~EventTimer()
    {
        hipEventDestroy(a);
        hipEventDestroy(b);
    }
    void start() { hipEventRecord(a, s); }
    float stop()
    {
        hipEventRecord(b, s);
        hipEventSynchronize(b);
        float ms = 0;
        hipEventElapsedTime(&ms, a, b);
        return ms;
    }
};

#define TRY(expr)                                                                                  \
    do {                                                                                           \
        int r_ = (expr);                                                                           \
        if (r_ != PUMP_OK)                                                                         \
            return r_;                                                                             \
    } while (0)

#define TRYH(expr)                                                                                 \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            ::pump::note_hip_error(e_, __FILE__, __LINE__);                                        \
            return PUMP_E_HIP;                                                                     \
        }                                                                                          \
    } while (0)

// Phases of the roadmap build (mainPUMP.cu:248-504): uploads, all ordered
// pairs, the tables from rn to the waypoint offsets, the fused edge pass
// and the half-spaces; rm_finish fills the info record.
int rm_upload(pump_roadmap *rm, const float *samples_h, const float *obstacles_h,
              const float *offsets_time_h, const float *xcomb_h);
int rm_pairs(pump_roadmap *rm);
int rm_tables(pump_roadmap *rm);
int rm_alloc_halfspaces(pump_roadmap *rm);
int rm_edge_pass(pump_roadmap *rm);
int rm_halfspaces(pump_roadmap *rm);
int rm_finish(pump_roadmap *rm, std::chrono::steady_clock::time_point t_begin);
// frees the roadmap and what it owns
void rm_release(pump_roadmap *rm);

// the search's path table: parent path, node, and the float sums of cost,
// time and the HSMC CP, indexed by path
struct PathTable {
    const int *prev, *node;
    const float *cost, *time, *cp;
};
int select_goal_path(pump_roadmap *rm, const pump_params &P, const std::vector<int> &goal_list,
                     const PathTable &T, pump_search_result *res, int &status);
// the search with its bookkeeping on the device (search_device.hip)
int search_device(pump_roadmap *rm, const pump_params &P, pump_search_result *res);

} // namespace rmb
} // namespace pump
