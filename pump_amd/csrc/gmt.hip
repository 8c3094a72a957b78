// gmt.hip -- the reference's single-objective Group Marching Tree over the
// roadmap (GMT.cu:3-261, lambda = 1: the whole wavefront expands at once)
// and its Euclidean-ball neighbour precompute (precomp.cu:3-76).  Neither is
// in the reference's build; both are listed API surface.
//
// GMT, per iteration, as GMT.cu:42-113:
//   1. the wavefront's nodes (in node order) mark every unvisited node
//      outside the wavefront that they reach by a go-edge as new, and
//      themselves as expanded (findWavefront);
//   2. each new node takes the cheapest connection from a wavefront node
//      among its come-edges, the first in come-edge order on a tie; a cost
//      of exactly 0 counts as unset, as in the reference
//      (findOptimalConnection);
//   3. a new node whose connecting edge collides in any segment is dropped
//      (verifyExpansion + removeInvalidExpansions);
//   4. expanded nodes leave the wavefront and the unvisited set, new nodes
//      join the wavefront, and the goal's cost is recorded when it joins
//      (updateWavefront).
// It stops when the goal has a non-zero cost, the new wavefront is empty, or
// after 100 iterations.  Node 0 starts the wavefront (setupArrays), as in
// the reference, whatever init_idx says.  Every step reads only state the
// previous step finished, so the result does not depend on thread order.
#include <cmath>
#include <cstring>

#include "pump_device.hpp"
#include "pump_internal.hpp"

#pragma clang fp contract(off)

namespace pump {

int scan_exclusive_i64(int64_t *d_a, int64_t n, void *d_ws, size_t ws_bytes, hipStream_t s);

namespace {

constexpr int kBlock = 256;
constexpr int kGmtMaxItrs = 100; // maxItrs (GMT.cu:8)

inline unsigned blocks_for(int64_t n) { return grid_for(n, kBlock, 1ll << 30); }

// ---- precomp: Euclidean neighbours within r (precomp.cu:3-76) -------------
// one lane per node, the samples streamed through LDS in tiles
constexpr int kNnTile = 256;

template <bool kFill>
__global__ __launch_bounds__(kBlock) void k_precomp_nn(const float *__restrict__ samples, int num,
                                                       float r2, int nn_size,
                                                       int *__restrict__ nn_sizes,
                                                       float *__restrict__ distances,
                                                       int *__restrict__ nn_edges)
{
    __shared__ float tile[kNnTile][kDim];
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    float x[kDim];
    for (int d = 0; d < kDim; ++d)
        x[d] = node < num ? samples[(int64_t)node * kDim + d] : 0.0f;
    int count = 0;
    for (int base = 0; base < num; base += kNnTile) {
        __syncthreads();
        for (int q = threadIdx.x; q < kNnTile * kDim; q += blockDim.x) {
            const int s = base + q / kDim;
            tile[q / kDim][q % kDim] = s < num ? samples[(int64_t)s * kDim + q % kDim] : 0.0f;
        }
        __syncthreads();
        const int m = num - base < kNnTile ? num - base : kNnTile;
        if (node < num)
            for (int j = 0; j < m; ++j) {
                const int sample = base + j;
                if (sample == node)
                    continue;
                float distance2 = 0;
                for (int d = 0; d < kDim; ++d) {
                    const float difference = x[d] - tile[j][d];
                    distance2 += difference * difference;
                }
                if (distance2 < r2) {
                    if (kFill && count < nn_size) {
                        nn_edges[(int64_t)node * nn_size + count] = sample;
                        distances[(int64_t)node * nn_size + count] = sqrtf(distance2);
                    }
                    ++count;
                }
            }
    }
    if (node >= num)
        return;
    if (!kFill) {
        nn_sizes[node] = count;
    } else {
        for (int k = count; k < nn_size; ++k) { // setupPrecompNN's defaults
            nn_edges[(int64_t)node * nn_size + k] = -1;
            distances[(int64_t)node * nn_size + k] = 0.0f;
        }
    }
}

// ---- GMT ------------------------------------------------------------------
struct GmtState {
    uint8_t *wavefront, *wavefront_new, *wavefront_was, *unvisited;
    int *active;
    int64_t *scan;
    float *cost_goal;
};

__global__ __launch_bounds__(kBlock) void k_gmt_setup(GmtState S, const uint8_t *is_free_samples,
                                                      float *costs, int *edges, int num)
{
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= num)
        return;
    const bool free = is_free_samples[node] != 0;
    S.unvisited[node] = free;
    S.wavefront_new[node] = 0;
    S.wavefront_was[node] = !free;
    S.wavefront[node] = node == 0;
    if (node == 0)
        *S.cost_goal = 0.0f;
    costs[node] = free ? 0.0f : -11.0f;
    edges[node] = free ? -1 : -2;
}

__global__ __launch_bounds__(kBlock) void k_flag_i64(const uint8_t *__restrict__ f, int num,
                                                     int64_t *__restrict__ out)
{
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node < num)
        out[node] = f[node] ? 1 : 0;
}

// fillWavefront (GMT.cu:169-179): the flagged nodes in node order
__global__ __launch_bounds__(kBlock) void k_fill_active(const uint8_t *__restrict__ f, int num,
                                                        const int64_t *__restrict__ scan,
                                                        int *__restrict__ active)
{
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node < num && f[node])
        active[scan[node]] = node;
}

// findWavefront (GMT.cu:146-167)
__global__ __launch_bounds__(kBlock) void k_find_wavefront(GmtState S, const int *__restrict__ nn_go,
                                                           int nn_size, int n_active)
{
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= n_active)
        return;
    const int node = S.active[tid];
    S.wavefront_was[node] = 1;
    for (int i = 0; i < nn_size; ++i) {
        const int nn = nn_go[(int64_t)node * nn_size + i];
        if (nn == -1)
            return;
        if (S.unvisited[nn] && !S.wavefront[nn])
            S.wavefront_new[nn] = 1; // every writer writes 1
    }
}

// the edge of the roadmap from node `from` to node `to`
__device__ __forceinline__ int64_t edge_of(int from, int to, int num, const int *nn_idxs,
                                           const int64_t *row_start, const int *edge_dst)
{
    if (nn_idxs)
        return nn_idxs[(int64_t)to * num + from]; // nnIdxs[to*NUM + from] (mainPUMP.cu:338)
    int64_t lo = row_start[from], hi = row_start[from + 1];
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (edge_dst[mid] < to)
            lo = mid + 1;
        else
            hi = mid;
    }
    return (lo < row_start[from + 1] && edge_dst[lo] == to) ? lo : -1;
}

// findOptimalConnection (GMT.cu:181-201), then the edge check of
// verifyExpansion / removeInvalidExpansions (GMT.cu:203-243,
// collisionCheck.cu:39-63): the connecting edge's segments in order
__global__ __launch_bounds__(kBlock) void k_connect(
    GmtState S, float *__restrict__ costs, int *__restrict__ edges,
    const float *__restrict__ distances_come, const int *__restrict__ nn_come, int nn_size,
    int n_active, int num, const int *nn_idxs, const int64_t *row_start, const int *edge_dst,
    const float *__restrict__ disc, const int64_t *wp_offsets, int num_disc,
    const float *__restrict__ obstacles, int n_obs, int *__restrict__ bad_edge)
{
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= n_active)
        return;
    const int node = S.active[tid];
    for (int i = 0; i < nn_size; ++i) {
        const int nn = nn_come[(int64_t)node * nn_size + i];
        if (nn == -1)
            break;
        if (S.wavefront[nn]) {
            const float cost_tmp = costs[nn] + distances_come[(int64_t)node * nn_size + i];
            if (costs[node] == 0 || cost_tmp < costs[node]) {
                costs[node] = cost_tmp;
                edges[node] = nn;
            }
        }
    }
    const int from = edges[node];
    const int64_t e = from >= 0 ? edge_of(from, node, num, nn_idxs, row_start, edge_dst) : -1;
    if (e < 0) {
        atomicOr(bad_edge, 1); // no roadmap edge for the connection
        return;
    }
    const int64_t g0 = wp_offsets ? wp_offsets[e] : e * (int64_t)(num_disc + 1);
    const int n_seg = wp_offsets ? (int)(wp_offsets[e + 1] - g0) - 1 : num_disc;
    bool collides = false;
    for (int k = 0; k < n_seg && !collides; ++k) {
        float v[3], w[3], bmin[3], bmax[3];
        for (int d = 0; d < 3; ++d) {
            v[d] = disc[(g0 + k) * kDim + d];
            w[d] = disc[(g0 + k + 1) * kDim + d];
        }
        seg_bbox(v, w, bmin, bmax);
        for (int b = 0; b < n_obs && !collides; ++b) {
            const float lo[3] = {obstacles[b * 12 + 0], obstacles[b * 12 + 1],
                                 obstacles[b * 12 + 2]};
            const float hi[3] = {obstacles[b * 12 + 6], obstacles[b * 12 + 7],
                                 obstacles[b * 12 + 8]};
            collides = segment_hits_box(v, w, bmin, bmax, lo, hi);
        }
    }
    if (collides) {
        costs[node] = 0.0f;
        edges[node] = -1;
        S.wavefront_new[node] = 0;
    }
}

// updateWavefront (GMT.cu:245-261)
__global__ __launch_bounds__(kBlock) void k_update_wavefront(GmtState S, const float *costs,
                                                             int num, int goal)
{
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= num || !S.unvisited[node])
        return;
    if (S.wavefront_was[node]) {
        S.wavefront[node] = 0;
        S.unvisited[node] = 0;
    } else if (S.wavefront_new[node]) {
        S.wavefront[node] = 1;
        S.wavefront_new[node] = 0;
        if (node == goal)
            *S.cost_goal = costs[node];
    }
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

} // namespace
} // namespace pump

using namespace pump;

extern "C" int pump_precomp_nn_count(const float *d_samples, int num, float r2, int *d_nn_sizes,
                                     pump_stream_t stream)
{
    PUMP_REQUIRE(d_samples && d_nn_sizes && num >= 1);
    hipLaunchKernelGGL(k_precomp_nn<false>, dim3(blocks_for(num)), dim3(kBlock), 0,
                       as_stream(stream), d_samples, num, r2, 0, d_nn_sizes, nullptr, nullptr);
    PUMP_LAUNCHED();
    return PUMP_OK;
}

extern "C" int pump_precomp_nn(const float *d_samples, int num, float r2, int nn_size,
                               float *d_distances, int *d_nn_edges, pump_stream_t stream)
{
    PUMP_REQUIRE(d_samples && d_distances && d_nn_edges && num >= 1 && nn_size >= 0);
    hipLaunchKernelGGL(k_precomp_nn<true>, dim3(blocks_for(num)), dim3(kBlock), 0,
                       as_stream(stream), d_samples, num, r2, nn_size, nullptr, d_distances,
                       d_nn_edges);
    PUMP_LAUNCHED();
    return PUMP_OK;
}

extern "C" size_t pump_gmt_workspace(int num)
{
    if (num < 1)
        return 0;
    return align256((size_t)num) * 4 + align256(sizeof(int) * (size_t)num) +
           align256(sizeof(int64_t) * ((size_t)num + 1)) + pump_scan_workspace(num) + 512;
}

extern "C" int pump_gmt_wavefront(const float *d_obstacles, int n_obs,
                                  const float *d_distances_come, const int *d_nn_go_edges,
                                  const int *d_nn_come_edges, int nn_size, const float *d_disc,
                                  const int64_t *d_wp_offsets, int num_disc, const int *d_nn_idxs,
                                  const int64_t *d_row_start, const int *d_edge_dst,
                                  const uint8_t *d_is_free_samples, int num, float *d_costs,
                                  int *d_edges, int goal_idx, void *d_ws, size_t ws_bytes,
                                  float *cost_goal_host, int *itrs_host, pump_stream_t stream)
{
    PUMP_REQUIRE(num >= 2 && goal_idx >= 0 && goal_idx < num && nn_size >= 0 && n_obs >= 0);
    PUMP_REQUIRE(d_distances_come && d_nn_go_edges && d_nn_come_edges && d_disc &&
                 d_is_free_samples && d_costs && d_edges && cost_goal_host && itrs_host);
    PUMP_REQUIRE(n_obs == 0 || d_obstacles);
    PUMP_REQUIRE(d_nn_idxs || (d_row_start && d_edge_dst));
    PUMP_REQUIRE(d_wp_offsets || num_disc >= 1);
    PUMP_REQUIRE(d_ws && ws_bytes >= pump_gmt_workspace(num));
    hipStream_t s = as_stream(stream);
    char *w = static_cast<char *>(d_ws);
    GmtState S;
    S.wavefront = reinterpret_cast<uint8_t *>(w);
    S.wavefront_new = S.wavefront + align256(num);
    S.wavefront_was = S.wavefront_new + align256(num);
    S.unvisited = S.wavefront_was + align256(num);
    S.active = reinterpret_cast<int *>(S.unvisited + align256(num));
    S.scan = reinterpret_cast<int64_t *>(reinterpret_cast<char *>(S.active) +
                                         align256(sizeof(int) * (size_t)num));
    char *tail = reinterpret_cast<char *>(S.scan) + align256(sizeof(int64_t) * ((size_t)num + 1));
    S.cost_goal = reinterpret_cast<float *>(tail);
    int *bad_edge = reinterpret_cast<int *>(tail + 256);
    void *scan_ws = tail + 512;
    const size_t scan_bytes = pump_scan_workspace(num);

    hipLaunchKernelGGL(k_gmt_setup, dim3(blocks_for(num)), dim3(kBlock), 0, s, S,
                       d_is_free_samples, d_costs, d_edges, num);
    PUMP_LAUNCHED();
    PUMP_HIP(hipMemsetAsync(bad_edge, 0, sizeof(int), s));
    // the flagged nodes of f, in node order, into S.active; returns their count
    auto gather = [&](const uint8_t *f, int64_t *count) -> int {
        hipLaunchKernelGGL(k_flag_i64, dim3(blocks_for(num)), dim3(kBlock), 0, s, f, num, S.scan);
        PUMP_LAUNCHED();
        int r = scan_exclusive_i64(S.scan, num, scan_ws, scan_bytes, s);
        if (r != PUMP_OK)
            return r;
        hipLaunchKernelGGL(k_fill_active, dim3(blocks_for(num)), dim3(kBlock), 0, s, f, num,
                           S.scan, S.active);
        PUMP_LAUNCHED();
        PUMP_HIP(hipMemcpyAsync(count, S.scan + num, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        PUMP_HIP(hipStreamSynchronize(s));
        return PUMP_OK;
    };
    float cost_goal = 0.0f;
    int itrs = 0;
    while (itrs < kGmtMaxItrs && cost_goal == 0) {
        ++itrs;
        int64_t n_active = 0;
        int r = gather(S.wavefront, &n_active);
        if (r != PUMP_OK)
            return r;
        if (n_active > 0) {
            hipLaunchKernelGGL(k_find_wavefront, dim3(blocks_for(n_active)), dim3(kBlock), 0, s,
                               S, d_nn_go_edges, nn_size, (int)n_active);
            PUMP_LAUNCHED();
        }
        r = gather(S.wavefront_new, &n_active);
        if (r != PUMP_OK)
            return r;
        if (n_active == 0)
            break; // the next wavefront is empty
        hipLaunchKernelGGL(k_connect, dim3(blocks_for(n_active)), dim3(kBlock), 0, s, S, d_costs,
                           d_edges, d_distances_come, d_nn_come_edges, nn_size, (int)n_active,
                           num, d_nn_idxs, d_row_start, d_edge_dst, d_disc, d_wp_offsets,
                           num_disc, d_obstacles, n_obs, bad_edge);
        PUMP_LAUNCHED();
        hipLaunchKernelGGL(k_update_wavefront, dim3(blocks_for(num)), dim3(kBlock), 0, s, S,
                           d_costs, num, goal_idx);
        PUMP_LAUNCHED();
        int bad = 0;
        PUMP_HIP(hipMemcpyAsync(&cost_goal, S.cost_goal, sizeof(float), hipMemcpyDeviceToHost, s));
        PUMP_HIP(hipMemcpyAsync(&bad, bad_edge, sizeof(int), hipMemcpyDeviceToHost, s));
        PUMP_HIP(hipStreamSynchronize(s));
        if (bad)
            return PUMP_E_INVALID; // a connection the roadmap has no edge for
    }
    *cost_goal_host = cost_goal;
    *itrs_host = itrs;
    return PUMP_OK;
}
