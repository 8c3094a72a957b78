// search_device.hip -- the PUMP search (PUMP.cu:3-290) with its bookkeeping
// on the device.  Same decisions as the host restatement in
// host_pipeline.hip, step for step:
//   * wavefront expansion (PUMP.cu:47-66): the open paths of the current
//     bucket in bucket order, each followed by its node's free out-edges to
//     free nodes in edge order -- a precomputed per-node list of those edges,
//     a scan of the open paths' degrees, and one wave per open path writing
//     its run;
//   * the HSMC step (PUMP.cu:80-160): pump_propagate_cp(_bits), unchanged;
//   * cost/time sums (PUMP.cu:164-167) as float adds in the same order;
//   * CP cutoff and epsilon-dominance (PUMP.cu:169-218): a new path is
//     dominated by any path of its node -- stored before this iteration, or
//     kept in this one -- that beats it by eps in CP and cost, except that of
//     two co-dominant new paths the lower index survives; the outcome of one
//     path does not depend on the others' outcomes, so every path is decided
//     by one lane against the node's stored-path list and its group of new
//     paths (a counting sort by node);
//   * compaction (PUMP.cu:220-256): survivors keep wavefront order (a scan),
//     their bucket ((int)(dcost/dr) + gidx + 1) % nb is computed as the host
//     does, and each bucket list is appended in wavefront order (a scan per
//     bucket); validity rows move by one gather;
//   * the goal test (PUMP.cu:259-265) on the paths stored at the goal.
// The per-node path lists are linked lists (head per node, next per path),
// since only their contents matter.  The host keeps the loop control and
// the goal-path selection (PUMP.cu:354-405), which reads the path table back
// once at the end.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/pump_host.h"
#include "pump_device.hpp"
#include "pump_internal.hpp"
#include "roadmap_internal.hpp"

#pragma clang fp contract(off)

namespace pump {

int scan_exclusive_i64(int64_t *d_a, int64_t n, void *d_ws, size_t ws_bytes, hipStream_t s);

namespace {

constexpr int kBlock = 256;

inline unsigned blocks_for(int64_t n) { return grid_for(n, kBlock, 1ll << 30); }

// free out-edges to free nodes, per node, in edge order
// One wave per node, lanes over its out-edges: the count of the edges whose
// flag and destination sample are free, then (with the counts scanned) those
// edges in edge order.
__global__ __launch_bounds__(kBlock) void k_valid_count(const int64_t *__restrict__ row_start,
                                                        const int *__restrict__ edge_dst,
                                                        const uint8_t *__restrict__ free_edges,
                                                        const uint8_t *__restrict__ free_samples,
                                                        int num, int64_t *__restrict__ vcnt)
{
    const int n = (int)((blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if (n >= num) // uniform within the wave
        return;
    const int64_t end = row_start[n + 1];
    int64_t c = 0;
    for (int64_t e0 = row_start[n]; e0 < end; e0 += 64) {
        const int64_t e = e0 + lane;
        const bool v = e < end && free_edges[e] && free_samples[edge_dst[e]];
        c += __popcll(__ballot(v));
    }
    if (lane == 0)
        vcnt[n] = c;
}

__global__ __launch_bounds__(kBlock) void k_valid_fill(const int64_t *__restrict__ row_start,
                                                       const int *__restrict__ edge_dst,
                                                       const uint8_t *__restrict__ free_edges,
                                                       const uint8_t *__restrict__ free_samples,
                                                       int num, const int64_t *__restrict__ voff,
                                                       int *__restrict__ vedge)
{
    const int n = (int)((blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if (n >= num)
        return;
    const int64_t end = row_start[n + 1];
    int64_t o = voff[n];
    for (int64_t e0 = row_start[n]; e0 < end; e0 += 64) {
        const int64_t e = e0 + lane;
        const bool v = e < end && free_edges[e] && free_samples[edge_dst[e]];
        const uint64_t m = __ballot(v);
        if (v)
            vedge[o + __popcll(m & ((1ull << lane) - 1ull))] = (int)e;
        o += __popcll(m);
    }
}

__global__ void k_init_search(int *node_head, int num, int init, int *path_prev, int *path_node,
                              float *path_cost, float *path_time, float *path_cp, int *path_next,
                              int *open0)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < num)
        node_head[n] = n == init ? 0 : -1;
    if (n == 0) {
        path_prev[0] = -2;
        path_node[0] = init;
        path_cost[0] = 0.0f;
        path_time[0] = 0.0f;
        path_cp[0] = 0.0f;
        path_next[0] = -1;
        open0[0] = 0;
    }
}

// expansion: degree of each open path's node
__global__ __launch_bounds__(kBlock) void k_open_count(const int *__restrict__ open, int n_open,
                                                       const int *__restrict__ path_node,
                                                       const int64_t *__restrict__ voff,
                                                       int64_t *__restrict__ cnt)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_open)
        return;
    const int n = path_node[open[g]];
    cnt[g] = voff[n + 1] - voff[n];
}

// expansion: one wave per open path writes its run of new paths
__global__ __launch_bounds__(kBlock) void k_open_fill(
    const int *__restrict__ open, int n_open, const int *__restrict__ path_node,
    const int64_t *__restrict__ voff, const int *__restrict__ vedge,
    const int *__restrict__ edge_dst, const int64_t *__restrict__ run, const int64_t *wp_offsets,
    int w_fixed, int n_p, int *__restrict__ wf_prev, int *__restrict__ wf_edge,
    int *__restrict__ wf_next, unsigned long long *__restrict__ steps)
{
    const int g = blockIdx.x * (kBlock / 64) + (int)(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (g >= n_open)
        return; // whole wave
    const int p = open[g];
    const int n = path_node[p];
    const int64_t v0 = voff[n];
    const int deg = (int)(voff[n + 1] - v0);
    const int64_t o = run[g];
    unsigned long long w = 0;
    for (int k = lane; k < deg; k += 64) {
        const int e = vedge[v0 + k];
        wf_prev[o + k] = p;
        wf_edge[o + k] = e;
        wf_next[o + k] = edge_dst[e];
        w += wp_offsets ? (unsigned long long)(wp_offsets[e + 1] - wp_offsets[e])
                        : (unsigned long long)w_fixed;
    }
    for (int m = 32; m > 0; m >>= 1)
        w += __shfl_xor(w, m, 64);
    if (lane == 0 && w)
        atomicAdd(steps, w * (unsigned long long)n_p);
}

// cost/time sums of the new paths, CP cutoff, per-node counts of kept paths
__global__ __launch_bounds__(kBlock) void k_new_paths(
    int n_new, const int *__restrict__ wf_prev, const int *__restrict__ wf_edge,
    const int *__restrict__ wf_next, const float *__restrict__ cp,
    const float *__restrict__ path_time, const float *__restrict__ path_cost,
    const float *__restrict__ topts_edge, const float *__restrict__ copts_edge, float cp_cutoff,
    float *__restrict__ new_time, float *__restrict__ new_cost, uint8_t *__restrict__ keep,
    int64_t *__restrict__ node_cnt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_new)
        return;
    const int prev = wf_prev[i], e = wf_edge[i];
    new_time[i] = path_time[prev] + topts_edge[e];
    new_cost[i] = path_cost[prev] + copts_edge[e];
    const bool k = !(cp[i] > cp_cutoff);
    keep[i] = k ? 1 : 0;
    if (k)
        atomicAdd(reinterpret_cast<unsigned long long *>(&node_cnt[wf_next[i]]), 1ull);
}

// the kept new paths grouped by node (order within a node is irrelevant)
__global__ __launch_bounds__(kBlock) void k_group(int n_new, const int *__restrict__ wf_next,
                                                  const uint8_t *__restrict__ keep,
                                                  const int64_t *__restrict__ node_off,
                                                  int *__restrict__ node_cur,
                                                  int *__restrict__ grp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_new || !keep[i])
        return;
    const int n = wf_next[i];
    grp[node_off[n] + atomicAdd(&node_cur[n], 1)] = i;
}

// epsilon-dominance (PUMP.cu:188-218); survivors exclude the init node
__global__ __launch_bounds__(kBlock) void k_dominance(
    int n_new, const int *__restrict__ wf_next, const uint8_t *__restrict__ keep,
    const float *__restrict__ cp, const float *__restrict__ new_cost,
    const int *__restrict__ node_head, const int *__restrict__ path_next,
    const float *__restrict__ path_cp, const float *__restrict__ path_cost,
    const int64_t *__restrict__ node_off, const int *__restrict__ grp, float eps_cp,
    float eps_cost, int init, int goal, int64_t *__restrict__ surv)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_new)
        return;
    bool s = keep[i] != 0;
    const int n = wf_next[i];
    if (s && n != goal) {
        const float ci = cp[i], ki = new_cost[i];
        bool dom = false;
        // paths stored before: their index is below every new path's, so the
        // co-dominance tie-break (pidx < cmp) never spares the new one
        for (int c = node_head[n]; c >= 0 && !dom; c = path_next[c])
            dom = path_cp[c] < eps_cp + ci && path_cost[c] < eps_cost + ki;
        for (int64_t q = node_off[n]; q < node_off[n + 1] && !dom; ++q) {
            const int j = grp[q];
            if (j == i)
                continue;
            const float cj = cp[j], kj = new_cost[j];
            if (cj < eps_cp + ci && kj < eps_cost + ki) {
                if (cj + eps_cp >= ci && kj + eps_cost >= ki && i < j)
                    continue;
                dom = true;
            }
        }
        s = !dom;
    }
    surv[i] = (s && n != init) ? 1 : 0;
}

// store the survivors at num_paths + rank, in wavefront order (PUMP.cu:220-256)
__global__ __launch_bounds__(kBlock) void k_store(
    int n_new, int64_t num_paths, const int64_t *__restrict__ pos, const int *__restrict__ wf_prev,
    const int *__restrict__ wf_next, const float *__restrict__ cp,
    const float *__restrict__ new_cost, const float *__restrict__ new_time,
    const float *__restrict__ h, float cost_threshold, float dr, int gidx, int nb, int goal,
    float cp_min_soln, int *__restrict__ path_prev, int *__restrict__ path_node,
    float *__restrict__ path_cost, float *__restrict__ path_time, float *__restrict__ path_cp,
    int *__restrict__ node_head, int *__restrict__ path_next, int *__restrict__ src_rows,
    int *__restrict__ bkt, int *__restrict__ flags)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_new || pos[i + 1] == pos[i])
        return;
    const int64_t k = pos[i];
    const int64_t st = num_paths + k;
    const int n = wf_next[i];
    const float c = new_cost[i];
    path_prev[st] = wf_prev[i];
    path_node[st] = n;
    path_cost[st] = c;
    path_time[st] = new_time[i];
    path_cp[st] = cp[i];
    src_rows[k] = i;
    float dcost = (c + h[n] - cost_threshold);
    dcost = (0.0f < dcost) ? dcost : 0.0f;
    const int b = (((int)(dcost / dr)) + gidx + 1) % nb;
    bkt[k] = b;
    if (b < 0 || b >= nb)
        atomicOr(&flags[1], 1); // a bucket outside [0, nb): negative rn
    path_next[st] = atomicExch(&node_head[n], (int)st);
    if (n == goal && cp[i] <= cp_min_soln)
        atomicOr(&flags[0], 1);
}

__global__ __launch_bounds__(kBlock) void k_bucket_flag(int64_t n, const int *__restrict__ bkt,
                                                        int b, int64_t *__restrict__ f)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n)
        f[k] = bkt[k] == b ? 1 : 0;
}

__global__ __launch_bounds__(kBlock) void k_bucket_place(int64_t n, int64_t num_paths,
                                                         const int64_t *__restrict__ f,
                                                         int *__restrict__ list, int64_t base)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n && f[k + 1] != f[k])
        list[base + f[k]] = (int)(num_paths + k);
}

__global__ void k_zero_i64(int64_t *a, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        a[i] = 0;
}

__global__ void k_zero_int(int *a, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        a[i] = 0;
}

template <typename T> T *as(rmb::DevBuf &b) { return reinterpret_cast<T *>(b.p); }


} // namespace

using namespace rmb;

int rmb::search_device(pump_roadmap *rm, const pump_params &P, pump_search_result *res)
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
    if (maxp > 0x7fffffff)
        return PUMP_E_RANGE;
    const int nb = P.num_buckets;
    if (nb < 1)
        return PUMP_E_INVALID;
    const int init_idx = 0, goal_idx = num - 1;
    const bool bits = P.validity_bits != 0;
    const int64_t row_bytes = bits ? (int64_t)((n_p + 63) / 64) * 8 : (int64_t)n_p;
    const float dr = P.lambda * rm->rn;
    const float cp_min_soln = P.cp_target / P.cp_factor;
    const float cp_cutoff = P.cp_target * P.cp_factor;

    SearchWs *W = search_ws(rm);
    if (!W)
        return PUMP_E_NOMEM;
    // free out-edges to free nodes per node
    DevBuf &voff = W->voff, &vedge = W->vedge, &ws = W->scan;
    TRY(voff.ensure(sizeof(int64_t) * ((size_t)num + 1)));
    TRY(ws.ensure(pump_scan_workspace(std::max<int64_t>(num, maxp) + 1)));
    hipLaunchKernelGGL(k_valid_count, dim3(blocks_for((int64_t)num * 64)), dim3(kBlock), 0, s,
                       rm->row_start, rm->edge_dst, rm->is_free_edges, rm->is_free_samples, num,
                       as<int64_t>(voff));
    PUMP_LAUNCHED();
    TRY(scan_exclusive_i64(as<int64_t>(voff), num, ws.p, ws.bytes, s));
    int64_t n_valid_edges = 0;
    TRYH(hipMemcpyAsync(&n_valid_edges, as<int64_t>(voff) + num, sizeof(int64_t),
                        hipMemcpyDeviceToHost, s));
    TRYH(hipStreamSynchronize(s));
    TRY(vedge.ensure(sizeof(int) * (size_t)std::max<int64_t>(n_valid_edges, 1)));
    hipLaunchKernelGGL(k_valid_fill, dim3(blocks_for((int64_t)num * 64)), dim3(kBlock), 0, s,
                       rm->row_start, rm->edge_dst, rm->is_free_edges, rm->is_free_samples, num,
                       as<int64_t>(voff), as<int>(vedge));
    PUMP_LAUNCHED();

    // the path table, per-node path lists, bucket lists, validity rows
    DevBuf &prev = W->prev, &node = W->node, &cost = W->cost, &time = W->time, &cpv = W->cpv,
           &next = W->next, &head = W->head, &open = W->open, &valid = W->valid,
           &flags = W->flags, &steps = W->steps, &clamp = W->clamp;
    TRY(prev.ensure(sizeof(int) * (size_t)maxp));
    TRY(node.ensure(sizeof(int) * (size_t)maxp));
    TRY(cost.ensure(sizeof(float) * (size_t)maxp));
    TRY(time.ensure(sizeof(float) * (size_t)maxp));
    TRY(cpv.ensure(sizeof(float) * (size_t)maxp));
    TRY(next.ensure(sizeof(int) * (size_t)maxp));
    TRY(head.ensure(sizeof(int) * (size_t)num));
    TRY(open.ensure(sizeof(int) * (size_t)nb * maxp));
    TRY(valid.ensure((size_t)maxp * row_bytes));
    TRY(flags.ensure(2 * sizeof(int)));
    TRY(steps.ensure(sizeof(unsigned long long)));
    TRY(clamp.ensure(2 * sizeof(unsigned long long)));
    hipLaunchKernelGGL(k_init_search, dim3(blocks_for(num)), dim3(kBlock), 0, s, as<int>(head),
                       num, init_idx, as<int>(prev), as<int>(node), as<float>(cost),
                       as<float>(time), as<float>(cpv), as<int>(next), as<int>(open));
    PUMP_LAUNCHED();
    hipLaunchKernelGGL(k_zero_i64, dim3(1), dim3(64), 0, s, as<int64_t>(steps), 1);
    PUMP_LAUNCHED();
    hipLaunchKernelGGL(k_zero_i64, dim3(1), dim3(64), 0, s, as<int64_t>(clamp), 2);
    PUMP_LAUNCHED();
    if (bits) {
        TRY(pump_setup_valid_bits(as<uint64_t>(valid), n_p, 1, s));
    } else {
        TRYH(hipMemsetAsync(valid.p, 1, (size_t)row_bytes, s));
    }
    float cost_threshold = 0.0f;
    TRYH(hipMemcpyAsync(&cost_threshold, rm->h + init_idx, sizeof(float), hipMemcpyDeviceToHost,
                        s));
    TRYH(hipStreamSynchronize(s));

    std::vector<int64_t> sizeG(nb, 0);
    sizeG[0] = 1;
    DevBuf &cnt = W->cnt, &wf_prev = W->wf_prev, &wf_edge = W->wf_edge, &wf_next = W->wf_next,
           &stage = W->stage, &d_cp = W->cp, &new_time = W->new_time, &new_cost = W->new_cost,
           &keep = W->keep, &node_cnt = W->node_cnt, &node_cur = W->node_cur, &grp = W->grp,
           &surv = W->surv, &src_rows = W->src_rows, &bkt = W->bkt, &bflag = W->bflag,
           &d_ws = W->prop, &row_src = W->row_src;
    TRY(node_cnt.ensure(sizeof(int64_t) * ((size_t)num + 1)));
    TRY(node_cur.ensure(sizeof(int) * (size_t)num));

    int64_t num_paths = 1;
    int gidx = 0;
    bool goal_cond = false, goal_next = false, empty_open = false;
    int itrs = 0, status = 0;
    float ms_prop = 0;
    EventTimer tm(s);

    while (itrs < P.max_itrs && !goal_cond && !empty_open) {
        ++itrs;
        if (goal_next)
            goal_cond = true;
        const int n_open = (int)sizeG[gidx];
        int *open_g = as<int>(open) + (size_t)gidx * maxp;
        // expansion (PUMP.cu:47-66)
        TRY(cnt.ensure(sizeof(int64_t) * ((size_t)n_open + 1)));
        hipLaunchKernelGGL(k_open_count, dim3(blocks_for(n_open)), dim3(kBlock), 0, s, open_g,
                           n_open, as<int>(node), as<int64_t>(voff), as<int64_t>(cnt));
        PUMP_LAUNCHED();
        TRY(scan_exclusive_i64(as<int64_t>(cnt), n_open, ws.p, ws.bytes, s));
        int64_t n_new64 = 0;
        TRYH(hipMemcpyAsync(&n_new64, as<int64_t>(cnt) + n_open, sizeof(int64_t),
                            hipMemcpyDeviceToHost, s));
        TRYH(hipStreamSynchronize(s));
        if (n_new64 >= maxp) {
            status = -2;
            break;
        }
        const int n_new = (int)n_new64;
        if (n_new == 0)
            break;
        if (num_paths + n_new >= maxp) {
            status = -3;
            break;
        }
        TRY(wf_prev.ensure(sizeof(int) * n_new));
        TRY(wf_edge.ensure(sizeof(int) * n_new));
        TRY(wf_next.ensure(sizeof(int) * n_new));
        hipLaunchKernelGGL(k_open_fill, dim3((unsigned)((n_open + 3) / 4)), dim3(kBlock), 0, s,
                           open_g, n_open, as<int>(node), as<int64_t>(voff), as<int>(vedge),
                           rm->edge_dst, as<int64_t>(cnt), rm->wp_offsets, P.w_fixed, n_p,
                           as<int>(wf_prev), as<int>(wf_edge), as<int>(wf_next),
                           as<unsigned long long>(steps));
        PUMP_LAUNCHED();
        sizeG[gidx] = 0;
        res->num_considered += n_new;

        // HSMC step (PUMP.cu:80-160)
        TRY(stage.ensure((size_t)n_new * row_bytes));
        // room for n_new paths of the longest edge's waypoints, so the
        // binned form runs however dense the stored half-spaces are
        const int64_t w_bound = rm->prm.w_fixed > 0 ? rm->prm.w_fixed : rm->prm.w_cap;
        TRY(d_ws.ensure(pump_propagate_workspace_pairs(n_new, rm->tmax_steps,
                                                       (int64_t)n_new * w_bound)));
        TRY(d_cp.ensure(sizeof(float) * n_new));
        TRY(row_src.ensure(sizeof(int) * n_new));
        unsigned long long *cl = as<unsigned long long>(clamp);
        tm.start();
        // copy-on-write: the children that meet no half-space keep their
        // parent's row and CP; the survivors' rows are materialised below
        TRY(pump_propagate_cp_cow(
            bits ? 1 : 0, n_new, 0, rm->wp_offsets, as<int>(wf_edge), as<int>(wf_prev), rm->as,
            rm->bs, rm->edge_hs_mask, as<float>(time), rm->topts_edge, P.dt, rm->max_hs,
            rm->offsets_time, rm->tmax_steps, P.offset_mult, rm->n_mc, valid.p, as<float>(cpv),
            stage.p, 0, as<int>(row_src), n_p, as<float>(d_cp), 0, bits ? P.precision : 0,
            rm->offsets_split, cl + 1, cl, d_ws.p, d_ws.bytes, s));
        ms_prop += tm.stop();

        // sums, cutoff, dominance (PUMP.cu:164-218)
        TRY(new_time.ensure(sizeof(float) * n_new));
        TRY(new_cost.ensure(sizeof(float) * n_new));
        TRY(keep.ensure((size_t)n_new));
        TRY(grp.ensure(sizeof(int) * n_new));
        TRY(surv.ensure(sizeof(int64_t) * ((size_t)n_new + 1)));
        hipLaunchKernelGGL(k_zero_i64, dim3(blocks_for(num + 1)), dim3(kBlock), 0, s,
                           as<int64_t>(node_cnt), (int64_t)num + 1);
        PUMP_LAUNCHED();
        hipLaunchKernelGGL(k_zero_int, dim3(blocks_for(num)), dim3(kBlock), 0, s,
                           as<int>(node_cur), (int64_t)num);
        PUMP_LAUNCHED();
        hipLaunchKernelGGL(k_new_paths, dim3(blocks_for(n_new)), dim3(kBlock), 0, s, n_new,
                           as<int>(wf_prev), as<int>(wf_edge), as<int>(wf_next), as<float>(d_cp),
                           as<float>(time), as<float>(cost), rm->topts_edge, rm->copts_edge,
                           cp_cutoff, as<float>(new_time), as<float>(new_cost), as<uint8_t>(keep),
                           as<int64_t>(node_cnt));
        PUMP_LAUNCHED();
        TRY(scan_exclusive_i64(as<int64_t>(node_cnt), num, ws.p, ws.bytes, s));
        hipLaunchKernelGGL(k_group, dim3(blocks_for(n_new)), dim3(kBlock), 0, s, n_new,
                           as<int>(wf_next), as<uint8_t>(keep), as<int64_t>(node_cnt),
                           as<int>(node_cur), as<int>(grp));
        PUMP_LAUNCHED();
        hipLaunchKernelGGL(k_dominance, dim3(blocks_for(n_new)), dim3(kBlock), 0, s, n_new,
                           as<int>(wf_next), as<uint8_t>(keep), as<float>(d_cp),
                           as<float>(new_cost), as<int>(head), as<int>(next), as<float>(cpv),
                           as<float>(cost), as<int64_t>(node_cnt), as<int>(grp), P.eps_cp,
                           P.eps_cost, init_idx, goal_idx, as<int64_t>(surv));
        PUMP_LAUNCHED();
        TRY(scan_exclusive_i64(as<int64_t>(surv), n_new, ws.p, ws.bytes, s));

        // store good paths (PUMP.cu:220-256)
        TRY(src_rows.ensure(sizeof(int) * n_new));
        TRY(bkt.ensure(sizeof(int) * n_new));
        TRY(bflag.ensure(sizeof(int64_t) * ((size_t)n_new + 1)));
        hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s, as<int>(flags), (int64_t)2);
        PUMP_LAUNCHED();
        hipLaunchKernelGGL(k_store, dim3(blocks_for(n_new)), dim3(kBlock), 0, s, n_new, num_paths,
                           as<int64_t>(surv), as<int>(wf_prev), as<int>(wf_next), as<float>(d_cp),
                           as<float>(new_cost), as<float>(new_time), rm->h, cost_threshold, dr,
                           gidx, nb, goal_idx, cp_min_soln, as<int>(prev), as<int>(node),
                           as<float>(cost), as<float>(time), as<float>(cpv), as<int>(head),
                           as<int>(next), as<int>(src_rows), as<int>(bkt), as<int>(flags));
        PUMP_LAUNCHED();
        int64_t n_valid = 0;
        int hflags[2] = {0, 0};
        TRYH(hipMemcpyAsync(&n_valid, as<int64_t>(surv) + n_new, sizeof(int64_t),
                            hipMemcpyDeviceToHost, s));
        TRYH(hipMemcpyAsync(hflags, flags.p, sizeof(hflags), hipMemcpyDeviceToHost, s));
        TRYH(hipStreamSynchronize(s));
        if (hflags[1]) {
            status = -6; // bucket index outside [0, nb)
            break;
        }
        if (n_valid > 0) {
            std::vector<int64_t> add(nb, 0);
            auto place = [&]() -> int {
                for (int b = 0; b < nb; ++b) {
                    hipLaunchKernelGGL(k_bucket_flag, dim3(blocks_for(n_valid)), dim3(kBlock), 0,
                                       s, n_valid, as<int>(bkt), b, as<int64_t>(bflag));
                    PUMP_LAUNCHED();
                    TRY(scan_exclusive_i64(as<int64_t>(bflag), n_valid, ws.p, ws.bytes, s));
                    hipLaunchKernelGGL(k_bucket_place, dim3(blocks_for(n_valid)), dim3(kBlock), 0,
                                       s, n_valid, num_paths, as<int64_t>(bflag),
                                       as<int>(open) + (size_t)b * maxp, sizeG[b]);
                    PUMP_LAUNCHED();
                    // the next bucket's flags reuse bflag: the stream orders
                    // this copy before them
                    TRYH(hipMemcpyAsync(&add[b], as<int64_t>(bflag) + n_valid, sizeof(int64_t),
                                        hipMemcpyDeviceToHost, s));
                }
                return PUMP_OK;
            };
            const int placed = place();
            // one wait for every bucket's count; also on an error, since
            // copies into add may still be queued
            TRYH(hipStreamSynchronize(s));
            TRY(placed);
            for (int b = 0; b < nb; ++b)
                sizeG[b] += add[b];
            TRY(pump_gather_rows_cow(stage.p, valid.p, as<int>(row_src), as<int>(src_rows),
                                     (int)n_valid, valid.p, num_paths, row_bytes, bits ? 1 : 0,
                                     s));
        }
        num_paths += n_valid;

        if (hflags[0]) {
            goal_next = true;
            goal_cond = true;
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
        unsigned long long st = 0;
        TRYH(hipMemcpy(cl, clamp.p, sizeof(cl), hipMemcpyDeviceToHost));
        TRYH(hipMemcpy(&st, steps.p, sizeof(st), hipMemcpyDeviceToHost));
        res->clamped_time_idx = (int64_t)cl[0];
        res->band_particles = (int64_t)cl[1];
        res->particle_steps = (int64_t)st;
    }
    res->itrs = itrs;
    res->num_paths = num_paths;
    res->ms_propagate = ms_prop;
    res->soln_path_idx = -1;
    W->last_num_paths = num_paths;

    if (status == 0) {
        const size_t np = (size_t)num_paths;
        std::vector<int> h_prev(np), h_node(np);
        std::vector<float> h_cost(np), h_time(np), h_cp(np);
        TRYH(hipMemcpy(h_prev.data(), prev.p, sizeof(int) * np, hipMemcpyDeviceToHost));
        TRYH(hipMemcpy(h_node.data(), node.p, sizeof(int) * np, hipMemcpyDeviceToHost));
        TRYH(hipMemcpy(h_cost.data(), cost.p, sizeof(float) * np, hipMemcpyDeviceToHost));
        TRYH(hipMemcpy(h_time.data(), time.p, sizeof(float) * np, hipMemcpyDeviceToHost));
        TRYH(hipMemcpy(h_cp.data(), cpv.p, sizeof(float) * np, hipMemcpyDeviceToHost));
        std::vector<int> goal_list;
        for (size_t q = 0; q < np; ++q)
            if (h_node[q] == goal_idx)
                goal_list.push_back((int)q);
        PathTable T{h_prev.data(), h_node.data(), h_cost.data(), h_time.data(), h_cp.data()};
        TRY(select_goal_path(rm, P, goal_list, T, res, status));
    }
    res->status = status;
    res->ms_total =
        std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return PUMP_OK;
}

} // namespace pump
