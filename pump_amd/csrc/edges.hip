// edges.hip -- kernel (b): edge waypoint generation fused with the
// segment/AABB test (reference: findDiscretizedPath 2pBVP.cu:53-69 driven by
// the host loop mainPUMP.cu:327-346, then freeEdges collisionCheck.cu:5-37),
// plus the reference-shaped freeEdges and sampleFree (sampler.cu:61-88).
//
// Layout of the fused kernel: a workgroup owns a run of consecutive edges
// (kEdgesPerBlock); its per-edge data and the obstacle boxes (structure of
// arrays, read as broadcasts) are staged once in LDS.  Each wave then walks
// tiles of 64 waypoints, one per lane, with no further workgroup barrier:
// the waypoints go to discMotions (consecutive lanes write consecutive
// 24-byte records), segment (k, k+1) is tested by the lane holding waypoint
// k, with the far end shuffled from the next lane, against the boxes that
// are not separated from the tile's bounding box.  An edge already known to
// be blocked skips its remaining segments, and a lane stops scanning boxes
// at its first hit.
#include "pump_device.hpp"
#include "pump_internal.hpp"

#pragma clang fp contract(off)

namespace pump {
namespace {

constexpr int kEdgesPerBlock = 128;
constexpr int kTile = 256;

constexpr int kMaskWords = kMaxObsLds / 64;

__device__ __forceinline__ float wave_min(float x)
{
    for (int m = 32; m > 0; m >>= 1)
        x = fminf(x, __shfl_xor(x, m, 64));
    return x;
}

__device__ __forceinline__ float wave_max(float x)
{
    for (int m = 32; m > 0; m >>= 1)
        x = fmaxf(x, __shfl_xor(x, m, 64));
    return x;
}

// Each wave walks its own tiles of 64 waypoints (consecutive tiles overlap
// by one waypoint, so a tile tests 63 segments); the segment's far end comes
// from the next lane by a shuffle.  The candidate boxes of a tile are those
// not separated from the bounding box of the tile's waypoints: every segment
// the tile tests lies inside it, so a box separated from it (with the
// reference's <= test, collisionCheck.cu:104-112) is rejected by the broad
// phase for each of those segments and skipping it changes no result.
// Waypoints are finite (tau >= the bisection's positive lower end).
template <bool kCheck, bool kWrite, int EPB>
__device__ __forceinline__ void edge_block(const EdgeJob &J, int64_t e0)
{
    const float *__restrict__ samples = J.samples;
    const int *__restrict__ edge_src = J.edge_src;
    const int *__restrict__ edge_dst = J.edge_dst;
    const float *__restrict__ topts_edge = J.topts_edge;
    const int64_t n_edges = J.n_edges;
    const int64_t *__restrict__ wp_offsets = J.wp_offsets;
    const int w_fixed = J.w_fixed;
    const float *__restrict__ obstacles = J.obstacles;
    const int n_obs = J.n_obs;
    float *__restrict__ disc = J.disc;
    uint8_t *__restrict__ is_free = J.is_free;
    constexpr int kWaves = kTile / 64;
    __shared__ ObsLds o;
    __shared__ float x0s[EPB][kDim], x1s[EPB][kDim];
    __shared__ TauPowers tps[EPB]; // per-edge powers, reciprocals
    __shared__ float dts[EPB];     // and waypoint spacing
    __shared__ int64_t off[EPB + 1];
    __shared__ int ok[EPB];

    const int ne = (int)((n_edges - e0) < EPB ? (n_edges - e0) : EPB);
    const int t = threadIdx.x, lane = t & 63;
    const int wid = __builtin_amdgcn_readfirstlane(t >> 6);
    // boxes past the LDS capacity are tested for every segment (no mask)
    const int n_masked = n_obs < kMaxObsLds ? n_obs : kMaxObsLds;
    const int n_words = (n_masked + 63) >> 6;
    if (kCheck)
        stage_obstacles(o, obstacles, n_obs);
    for (int q = t; q < ne * kDim; q += kTile) {
        int le = q / kDim, d = q - kDim * (q / kDim);
        x0s[le][d] = samples[(int64_t)edge_src[e0 + le] * kDim + d];
        x1s[le][d] = samples[(int64_t)edge_dst[e0 + le] * kDim + d];
    }
    for (int q = t; q <= ne; q += kTile)
        off[q] = wp_offsets ? wp_offsets[e0 + q] : (e0 + q) * (int64_t)w_fixed;
    __syncthreads();
    for (int q = t; q < ne; q += kTile) {
        const float tau = topts_edge[e0 + q];
        tps[q] = tau_powers(tau);
        dts[q] = tau / (float)(off[q + 1] - off[q] - 1);
        ok[q] = 1;
    }
    __syncthreads();

    const int64_t g_begin = off[0], g_end = off[ne];
    // tiles start every 63 waypoints; tile j belongs to wave j % kWaves
    for (int64_t base = g_begin + (int64_t)wid * 63; base < g_end; base += (int64_t)kWaves * 63) {
        const int64_t g = base + lane;
        const bool have = g < g_end;
        int le = 0, k = 0, W = 0;
        float x[kDim] = {0, 0, 0, 0, 0, 0};
        if (have) {
            le = find_edge(off, ne, g);
            k = (int)(g - off[le]);
            W = (int)(off[le + 1] - off[le]);
            const TauPowers P = tps[le];
            waypoint_dt(k, W, P, dts[le], x0s[le], x1s[le], x);
            // lane 63's waypoint is lane 0's of the next tile
            if (kWrite && (lane < 63 || g + 1 >= g_end))
                for (int d = 0; d < kDim; ++d)
                    disc[g * kDim + d] = x[d];
        }
        if (!kCheck)
            continue;
        // segment (k, k+1): far end from the next lane
        float w[3];
        for (int d = 0; d < 3; ++d)
            w[d] = __shfl_down(x[d], 1, 64);
        const bool test = have && lane < 63 && k < W - 1 && ok[le];
        // candidate boxes of the tile: not separated from its waypoints' box
        float tlo[3], thi[3];
        for (int d = 0; d < 3; ++d) {
            tlo[d] = wave_min(have ? x[d] : INFINITY);
            thi[d] = wave_max(have ? x[d] : -INFINITY);
        }
        if (__ballot(test) == 0ull)
            continue;
        uint64_t cand[kMaskWords];
        for (int wd = 0; wd < n_words; ++wd) {
            const int b = wd * 64 + lane;
            bool near = false;
            if (b < n_masked) {
                near = true;
                for (int d = 0; d < 3; ++d)
                    near = near && !(thi[d] <= o.lo[d][b] || o.hi[d][b] <= tlo[d]);
            }
            cand[wd] = __ballot(near);
        }
        if (test) {
            const float v[3] = {x[0], x[1], x[2]};
            float bmin[3], bmax[3];
            seg_bbox(v, w, bmin, bmax);
            bool hit = false;
            for (int wd = 0; wd < n_words && !hit; ++wd) {
                uint64_t m = cand[wd];
                while (m && !hit) {
                    const int b = wd * 64 + __builtin_ctzll(m);
                    m &= m - 1;
                    float lo[3] = {o.lo[0][b], o.lo[1][b], o.lo[2][b]};
                    float hi[3] = {o.hi[0][b], o.hi[1][b], o.hi[2][b]};
                    hit = segment_hits_box(v, w, bmin, bmax, lo, hi);
                }
            }
            for (int b = n_masked; b < n_obs && !hit; ++b) {
                float lo[3], hi[3];
                box_bounds(o, obstacles, b, lo, hi);
                hit = segment_hits_box(v, w, bmin, bmax, lo, hi);
            }
            if (hit)
                ok[le] = 0; // only zeros are written: concurrent writes agree
        }
    }
    if (kCheck) {
        __syncthreads();
        for (int q = t; q < ne; q += kTile)
            is_free[e0 + q] = (uint8_t)ok[q];
    }
}

template <bool kCheck, bool kWrite>
__global__ __launch_bounds__(kTile) void k_edges(
    const float *__restrict__ samples, const int *__restrict__ edge_src,
    const int *__restrict__ edge_dst, const float *__restrict__ topts_edge, int64_t n_edges,
    const int64_t *__restrict__ wp_offsets, int w_fixed, const float *__restrict__ obstacles,
    int n_obs, float *__restrict__ disc, uint8_t *__restrict__ is_free)
{
    const EdgeJob J{samples, edge_src, edge_dst, topts_edge, n_edges, wp_offsets,
                    w_fixed, obstacles, n_obs, disc, is_free};
    edge_block<kCheck, kWrite, kEdgesPerBlock>(J, (int64_t)blockIdx.x * kEdgesPerBlock);
}

// the fused edge pass of many instances in one launch
template <int EPB>
__global__ __launch_bounds__(kTile) void k_edges_many(const EdgeJob *__restrict__ jobs,
                                                      const int64_t *__restrict__ prefix,
                                                      int n_inst)
{
    const int k = instance_of(prefix, n_inst, blockIdx.x);
    const EdgeJob J = jobs[k];
    edge_block<true, true, EPB>(J, ((int64_t)blockIdx.x - prefix[k]) * EPB);
}

// freeEdges over precomputed discMotions: one lane per segment; the edge's
// flag lives in LDS for the workgroup that owns the edge.
__global__ __launch_bounds__(kTile) void k_free_edges(const float *__restrict__ obstacles,
                                                      int n_obs, int w_fixed,
                                                      const int64_t *__restrict__ wp_offsets,
                                                      const float *__restrict__ disc,
                                                      uint8_t *__restrict__ is_free,
                                                      int64_t n_edges)
{
    __shared__ ObsLds o;
    __shared__ int64_t off[kEdgesPerBlock + 1];
    __shared__ int ok[kEdgesPerBlock];
    const int64_t e0 = (int64_t)blockIdx.x * kEdgesPerBlock;
    const int ne = (int)((n_edges - e0) < kEdgesPerBlock ? (n_edges - e0) : kEdgesPerBlock);
    const int t = threadIdx.x;
    stage_obstacles(o, obstacles, n_obs);
    for (int q = t; q <= ne; q += kTile)
        off[q] = wp_offsets ? wp_offsets[e0 + q] : (e0 + q) * (int64_t)w_fixed;
    for (int q = t; q < ne; q += kTile)
        ok[q] = 1;
    __syncthreads();
    const int64_t g_begin = off[0], g_end = off[ne];
    for (int64_t g = g_begin + t; g < g_end; g += kTile) {
        int le = find_edge(off, ne, g);
        int k = (int)(g - off[le]);
        int W = (int)(off[le + 1] - off[le]);
        if (k < W - 1 && ok[le]) {
            float v[3], w[3];
            for (int d = 0; d < 3; ++d) {
                v[d] = disc[g * kDim + d];
                w[d] = disc[(g + 1) * kDim + d];
            }
            if (segment_blocked(v, w, o, obstacles, n_obs))
                ok[le] = 0;
        }
    }
    __syncthreads();
    for (int q = t; q < ne; q += kTile)
        is_free[e0 + q] = (uint8_t)ok[q];
}

__global__ __launch_bounds__(256) void k_sample_free(const float *__restrict__ obstacles, int n_obs,
                                                     const float *__restrict__ samples, int num,
                                                     uint8_t *__restrict__ is_free)
{
    __shared__ ObsLds o;
    stage_obstacles(o, obstacles, n_obs);
    __syncthreads();
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= num)
        return;
    float p[3] = {samples[(int64_t)n * kDim], samples[(int64_t)n * kDim + 1],
                  samples[(int64_t)n * kDim + 2]};
    bool free_ = true;
    for (int b = 0; b < n_obs && free_; ++b) {
        float lo[3], hi[3];
        box_bounds(o, obstacles, b, lo, hi);
        if (p[0] > lo[0] && p[0] < hi[0] && p[1] > lo[1] && p[1] < hi[1] && p[2] > lo[2] &&
            p[2] < hi[2])
            free_ = false;
    }
    is_free[n] = free_ ? 1 : 0;
}

} // namespace
} // namespace pump

using namespace pump;

static int edge_blocks(int64_t n_edges, unsigned &blocks)
{
    int64_t b = (n_edges + kEdgesPerBlock - 1) / kEdgesPerBlock;
    if (b > 0x7fffffff)
        return PUMP_E_RANGE;
    blocks = (unsigned)(b > 0 ? b : 1);
    return PUMP_OK;
}

extern "C" int pump_discretize_edges(const float *d_samples, const int *d_edge_src,
                                     const int *d_edge_dst, const float *d_topts_edge,
                                     int64_t n_edges, const int64_t *d_wp_offsets,
                                     int waypoints_count, float *d_disc_motions,
                                     pump_stream_t stream)
{
    PUMP_REQUIRE(n_edges >= 0);
    if (n_edges == 0)
        return PUMP_OK;
    PUMP_REQUIRE(d_samples && d_edge_src && d_edge_dst && d_topts_edge && d_disc_motions);
    PUMP_REQUIRE(d_wp_offsets || waypoints_count >= 2);
    unsigned blocks;
    int r = edge_blocks(n_edges, blocks);
    if (r)
        return r;
    hipLaunchKernelGGL((k_edges<false, true>), dim3(blocks), dim3(kTile), 0, as_stream(stream),
                       d_samples, d_edge_src, d_edge_dst, d_topts_edge, n_edges, d_wp_offsets,
                       waypoints_count, nullptr, 0, d_disc_motions, nullptr);
    PUMP_LAUNCHED();
    return PUMP_OK;
}

extern "C" int pump_edges_discretize_check(const float *d_samples, const int *d_edge_src,
                                           const int *d_edge_dst, const float *d_topts_edge,
                                           int64_t n_edges, const int64_t *d_wp_offsets,
                                           int waypoints_count, const float *d_obstacles,
                                           int n_obs, float *d_disc_motions,
                                           uint8_t *d_is_free_edges, pump_stream_t stream)
{
    PUMP_REQUIRE(n_edges >= 0 && n_obs >= 0);
    if (n_edges == 0)
        return PUMP_OK;
    PUMP_REQUIRE(d_samples && d_edge_src && d_edge_dst && d_topts_edge && d_is_free_edges);
    PUMP_REQUIRE(n_obs == 0 || d_obstacles);
    PUMP_REQUIRE(d_wp_offsets || waypoints_count >= 2);
    unsigned blocks;
    int r = edge_blocks(n_edges, blocks);
    if (r)
        return r;
    if (d_disc_motions)
        hipLaunchKernelGGL((k_edges<true, true>), dim3(blocks), dim3(kTile), 0, as_stream(stream),
                           d_samples, d_edge_src, d_edge_dst, d_topts_edge, n_edges, d_wp_offsets,
                           waypoints_count, d_obstacles, n_obs, d_disc_motions, d_is_free_edges);
    else
        hipLaunchKernelGGL((k_edges<true, false>), dim3(blocks), dim3(kTile), 0,
                           as_stream(stream), d_samples, d_edge_src, d_edge_dst, d_topts_edge,
                           n_edges, d_wp_offsets, waypoints_count, d_obstacles, n_obs, nullptr,
                           d_is_free_edges);
    PUMP_LAUNCHED();
    return PUMP_OK;
}

extern "C" int pump_free_edges(const float *d_obstacles, int n_obs, const float *d_samples,
                               const uint8_t *d_is_free_samples, int waypoints_count,
                               const int64_t *d_wp_offsets, const float *d_disc_motions,
                               uint8_t *d_is_free_edges, int64_t n_edges, pump_stream_t stream)
{
    (void)d_samples;          // passed but unused, as in the reference
    (void)d_is_free_samples;  // idem (collisionCheck.cu:5-37)
    PUMP_REQUIRE(n_edges >= 0 && n_obs >= 0);
    if (n_edges == 0)
        return PUMP_OK;
    PUMP_REQUIRE(d_disc_motions && d_is_free_edges && (n_obs == 0 || d_obstacles));
    PUMP_REQUIRE(d_wp_offsets || waypoints_count >= 1);
    unsigned blocks;
    int r = edge_blocks(n_edges, blocks);
    if (r)
        return r;
    hipLaunchKernelGGL(k_free_edges, dim3(blocks), dim3(kTile), 0, as_stream(stream), d_obstacles,
                       n_obs, waypoints_count, d_wp_offsets, d_disc_motions, d_is_free_edges,
                       n_edges);
    PUMP_LAUNCHED();
    return PUMP_OK;
}

extern "C" int pump_sample_free(const float *d_obstacles, int n_obs, const float *d_samples,
                                int num, uint8_t *d_is_free_samples, pump_stream_t stream)
{
    PUMP_REQUIRE(num >= 0 && n_obs >= 0);
    if (num == 0)
        return PUMP_OK;
    PUMP_REQUIRE(d_samples && d_is_free_samples && (n_obs == 0 || d_obstacles));
    hipLaunchKernelGGL(k_sample_free, dim3((num + 255) / 256), dim3(256), 0, as_stream(stream),
                       d_obstacles, n_obs, d_samples, num, d_is_free_samples);
    PUMP_LAUNCHED();
    return PUMP_OK;
}

int pump::edges_many(const EdgeJob *d_jobs, const int64_t *d_block_prefix, int n_inst,
                     int64_t n_blocks, int edges_per_block, hipStream_t s)
{
    PUMP_REQUIRE(d_jobs && d_block_prefix && n_inst > 0 && n_blocks >= 0 &&
                 n_blocks <= 0x7fffffff);
    if (n_blocks == 0)
        return PUMP_OK;
    if (edges_per_block == 64)
        hipLaunchKernelGGL(k_edges_many<64>, dim3((unsigned)n_blocks), dim3(kTile), 0, s, d_jobs,
                           d_block_prefix, n_inst);
    else if (edges_per_block == 128)
        hipLaunchKernelGGL(k_edges_many<128>, dim3((unsigned)n_blocks), dim3(kTile), 0, s, d_jobs,
                           d_block_prefix, n_inst);
    else if (edges_per_block == 256)
        hipLaunchKernelGGL(k_edges_many<256>, dim3((unsigned)n_blocks), dim3(kTile), 0, s, d_jobs,
                           d_block_prefix, n_inst);
    else
        return PUMP_E_INVALID;
    PUMP_LAUNCHED();
    return PUMP_OK;
}

