// mccp.hip -- full Monte Carlo collision probability (reference:
// collisionProbability + MCCP, collisionProbability.cu:4-92).
//
// The reference runs one thread per particle in 32 workgroups of 64 and walks
// the path's segments serially.  Here one wave owns one particle and its 64
// lanes test 64 consecutive segments at once (the particle's noise row is
// contiguous in t, so the loads coalesce); a wave ballot ends the particle at
// the first blocked segment.  Obstacles are staged in LDS, and each lane
// tests only the boxes that meet the union of the wave's 64 segment boxes.  Paths of a batch
// are the grid's y dimension, so many candidate paths share one launch.
#include "pump_device.hpp"
#include "pump_internal.hpp"

#pragma clang fp contract(off)

namespace pump {
namespace {

__global__ __launch_bounds__(256) void k_mccp(const float *__restrict__ obstacles, int n_obs,
                                              const float *__restrict__ xcomb, int tmax_steps,
                                              int n_mc, float om, const float *__restrict__ paths,
                                              int64_t stride_rows, const int *__restrict__ Ts,
                                              int *__restrict__ hits)
{
    __shared__ ObsLds o;
    stage_obstacles(o, obstacles, n_obs);
    __syncthreads();
    const int k = blockIdx.y;
    const int lane = threadIdx.x & 63;
    const int p = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (p >= n_mc)
        return;
    const float *path = paths + (int64_t)k * stride_rows * kDim;
    // segments t < T-2 (the reference skips the final segment); bounded by
    // the noise table and the path buffer
    int nseg = Ts[k] - 2;
    if (nseg > tmax_steps - 1)
        nseg = tmax_steps - 1;
    if (nseg > (int)stride_rows - 1)
        nseg = (int)stride_rows - 1;
    const float *xc = xcomb + (int64_t)p * 3 * tmax_steps;
    for (int t0 = 0; t0 < nseg; t0 += 64) {
        const int t = t0 + lane;
        const bool act = t < nseg;
        float v[3] = {0, 0, 0}, w[3] = {0, 0, 0};
        float bmin[3] = {INFINITY, INFINITY, INFINITY}, bmax[3] = {-INFINITY, -INFINITY, -INFINITY};
        if (act) {
            for (int d = 0; d < 3; ++d) {
                v[d] = path[t * kDim + d] + xc[t * 3 + d] * om;
                w[d] = path[(t + 1) * kDim + d] + xc[(t + 1) * 3 + d] * om;
            }
            seg_bbox(v, w, bmin, bmax);
        }
        // Only boxes that meet the union of the wave's segment boxes can be
        // hit (the broad phase of segment_hits_box uses the same <= test,
        // and a box separated from the union is separated from each
        // segment's box), so each lane tests just those: lane-parallel
        // over the boxes, 64 at a time, then the candidates one by one.
        float umin[3], umax[3];
        for (int d = 0; d < 3; ++d) {
            umin[d] = wave_min(bmin[d]);
            umax[d] = wave_max(bmax[d]);
        }
        bool hit = false;
        for (int b0 = 0; b0 < n_obs; b0 += 64) {
            bool meets = false;
            if (b0 + lane < n_obs) {
                float lo[3], hi[3];
                box_bounds(o, obstacles, b0 + lane, lo, hi);
                meets = !((umax[0] <= lo[0]) | (hi[0] <= umin[0]) | (umax[1] <= lo[1]) |
                          (hi[1] <= umin[1]) | (umax[2] <= lo[2]) | (hi[2] <= umin[2]));
            }
            unsigned long long cand = __ballot(meets);
            while (cand) {
                const int b = b0 + __builtin_ctzll(cand);
                cand &= cand - 1;
                float lo[3], hi[3];
                box_bounds(o, obstacles, b, lo, hi);
                if (act && !hit)
                    hit = segment_hits_box(v, w, bmin, bmax, lo, hi);
            }
        }
        if (__ballot(hit) != 0ull) {
            if (lane == 0)
                atomicAdd(&hits[k], 1);
            break;
        }
    }
}

__global__ void k_zero_hits(int *h, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        h[i] = 0;
}

} // namespace
} // namespace pump

using namespace pump;

extern "C" int pump_collision_probability_many(const float *d_obstacles, int n_obs,
                                               const float *d_xcomb, int tmax_steps, int n_mc,
                                               float offset_mult, const float *d_paths,
                                               int64_t path_stride_rows, const int *d_T,
                                               int n_paths, int *d_hits, int zero_hits,
                                               pump_stream_t stream)
{
    PUMP_REQUIRE(n_paths >= 0 && n_obs >= 0 && n_mc > 0 && tmax_steps > 0);
    if (n_paths == 0)
        return PUMP_OK;
    PUMP_REQUIRE(d_xcomb && d_paths && d_T && d_hits && path_stride_rows > 0);
    PUMP_REQUIRE(n_obs == 0 || d_obstacles);
    if (n_paths > 65535)
        return PUMP_E_RANGE;
    hipStream_t s = as_stream(stream);
    if (zero_hits) {
        hipLaunchKernelGGL(k_zero_hits, dim3(1), dim3(256), 0, s, d_hits, n_paths);
        PUMP_LAUNCHED();
    }
    dim3 grid((unsigned)((n_mc + 3) / 4), (unsigned)n_paths);
    hipLaunchKernelGGL(k_mccp, grid, dim3(256), 0, s, d_obstacles, n_obs, d_xcomb, tmax_steps, n_mc,
                       offset_mult, d_paths, path_stride_rows, d_T, d_hits);
    PUMP_LAUNCHED();
    return PUMP_OK;
}

extern "C" int pump_collision_probability(const float *d_obstacles, int n_obs,
                                          const float *d_xcomb, int tmax_steps, int n_mc,
                                          float offset_mult, const float *d_path, int T,
                                          float *cp_host, pump_stream_t stream)
{
    PUMP_REQUIRE(cp_host && d_path && T >= 0);
    if (T > tmax_steps + 1)
        return PUMP_E_RANGE;
    hipStream_t s = as_stream(stream);
    int *scratch = nullptr; // [0] = T, [1] = hits
    PUMP_HIP(hipMalloc((void **)&scratch, 2 * sizeof(int)));
    int hostT[2] = {T, 0};
    hipError_t ce = hipMemcpyAsync(scratch, hostT, 2 * sizeof(int), hipMemcpyHostToDevice, s);
    int r = PUMP_OK;
    if (ce == hipSuccess)
        r = pump_collision_probability_many(d_obstacles, n_obs, d_xcomb, tmax_steps, n_mc,
                                            offset_mult, d_path, T > 1 ? T : 1, scratch, 1,
                                            scratch + 1, 0, stream);
    int hits = 0;
    hipError_t de = hipMemcpyAsync(&hits, scratch + 1, sizeof(int), hipMemcpyDeviceToHost, s);
    hipError_t se = hipStreamSynchronize(s);
    hipFree(scratch);
    PUMP_HIP(ce);
    PUMP_HIP(de);
    PUMP_HIP(se);
    if (r != PUMP_OK)
        return r;
    *cp_host = ((float)hits) / ((float)n_mc);
    return PUMP_OK;
}
