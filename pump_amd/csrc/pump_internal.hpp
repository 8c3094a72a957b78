// pump_internal.hpp -- shared host-side helpers of the HIP entry points.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pump_abi.h"

namespace pump {

inline hipStream_t as_stream(pump_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// Records the failing HIP call for pump_last_error() (thread-local).
void note_hip_error(hipError_t e, const char *where, int line);

// Returns PUMP_E_HIP from the enclosing entry point on a HIP error.
#define PUMP_HIP(expr)                                                                             \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            ::pump::note_hip_error(e_, __FILE__, __LINE__);                                        \
            return PUMP_E_HIP;                                                                     \
        }                                                                                          \
    } while (0)

// Launch check: catches configuration errors of the preceding launch.
#define PUMP_LAUNCHED()                                                                            \
    do {                                                                                           \
        hipError_t e_ = hipGetLastError();                                                         \
        if (e_ != hipSuccess) {                                                                    \
            ::pump::note_hip_error(e_, __FILE__, __LINE__);                                        \
            return PUMP_E_HIP;                                                                     \
        }                                                                                          \
    } while (0)

#define PUMP_REQUIRE(cond)                                                                         \
    do {                                                                                           \
        if (!(cond))                                                                               \
            return PUMP_E_INVALID;                                                                 \
    } while (0)

// Waypoints per edge stay below this bound (the edge pass indexes the
// waypoints of a run of up to 256 edges with 32-bit integers).
constexpr int kMaxWaypointsPerEdge = 1 << 23;

// One instance of a batched (multi-instance) launch: pairs, the fused edge
// pass, and the half-space cache.  Instance k of a launch owns the blocks (or
// pair indices) [prefix[k], prefix[k+1]) of a prefix table in device memory.
// A pointer that a kernel reads from device memory (a job record of a
// batched launch) is a generic pointer to the compiler: its accesses become
// flat instructions with 64-bit vector addresses, where a kernel's own
// pointer arguments give global instructions with scalar bases (and fewer
// VGPRs: 109 against 77 in the binned half-space pass).  gptr reads such a
// field as a global pointer, which every job pointer is (device memory from
// hipMalloc).
template <class T>
__device__ __forceinline__ T *gptr(T *const &field)
{
    typedef __attribute__((address_space(1))) T *GlobalPtr;
    return (T *)*reinterpret_cast<const GlobalPtr *>(&field);
}

struct PairJob {
    const float *samples;
    float *copts, *topts;
    int num;
    float tmax;
    // magic number of num - 1 for the pair's row (DivMagic's fields; m = 0:
    // 64-bit division, for 2^32 pairs or more)
    uint32_t dm_m, dm_s1, dm_s2;
};

struct EdgeJob {
    const float *samples;
    const int *edge_src, *edge_dst;
    const float *topts_edge;
    int64_t n_edges;
    const int64_t *wp_offsets;
    int w_fixed;
    const float *obstacles;
    int n_obs;
    float *disc;
    uint8_t *is_free;
};

struct HsJob {
    int64_t n_edges;
    const float *disc;
    const int64_t *wp_offsets;
    int w_fixed;
    const uint8_t *is_free;
    int n_obs, max_hs;
    const float *obstacles;
    float *as, *bs;
    int *count;
    float maxd2;
    uint64_t *edge_mask;
};

__device__ __forceinline__ PairJob global_job(const PairJob &j)
{
    PairJob g = j;
    g.samples = gptr(j.samples);
    g.copts = gptr(j.copts);
    g.topts = gptr(j.topts);
    return g;
}
__device__ __forceinline__ EdgeJob global_job(const EdgeJob &j)
{
    EdgeJob g = j;
    g.samples = gptr(j.samples);
    g.edge_src = gptr(j.edge_src);
    g.edge_dst = gptr(j.edge_dst);
    g.topts_edge = gptr(j.topts_edge);
    g.wp_offsets = gptr(j.wp_offsets);
    g.obstacles = gptr(j.obstacles);
    g.disc = gptr(j.disc);
    g.is_free = gptr(j.is_free);
    return g;
}
__device__ __forceinline__ HsJob global_job(const HsJob &j)
{
    HsJob g = j;
    g.disc = gptr(j.disc);
    g.wp_offsets = gptr(j.wp_offsets);
    g.is_free = gptr(j.is_free);
    g.obstacles = gptr(j.obstacles);
    g.as = gptr(j.as);
    g.bs = gptr(j.bs);
    g.count = gptr(j.count);
    g.edge_mask = gptr(j.edge_mask);
    return g;
}

// all ordered pairs of one instance (pump_fill_copts_topts), with the
// bisection step count of the persistent launch in *steps (may be null)
int fill_copts_topts(const float *d_samples, int num, float tmax, float *d_copts, float *d_topts,
                     hipStream_t s, int64_t *steps);
int fill_pairs_many(const PairJob *d_jobs, const int64_t *d_prefix, int n_inst, int64_t n_pairs,
                    int blocks_per_cu, unsigned long long *d_counter, hipStream_t s);
// edges per workgroup of the batched edge pass: 64, 128 or 256
// (max_obs: the largest box count of the instances; obs_tile: boxes staged
// in LDS with candidate masks, the rest are tested from global memory)
int edges_many(const EdgeJob *d_jobs, const int64_t *d_block_prefix, int n_inst,
               int64_t n_blocks, int edges_per_block, int max_obs, int obs_tile,
               hipStream_t s);
constexpr int kHsEdgesPerBlock = 128;
// (max_obs, max_hs: the largest box count and slot count of the instances)
int cache_hs_many(const HsJob *d_jobs, const int64_t *d_block_prefix, int n_inst,
                  int64_t n_blocks, int max_obs, int max_hs, hipStream_t s);

// Grid size for a flat launch over n items, grid-stride beyond the cap.
inline unsigned grid_for(int64_t n, int block, int64_t cap = 1 << 20)
{
    int64_t g = (n + block - 1) / block;
    if (g < 1)
        g = 1;
    if (g > cap)
        g = cap;
    return (unsigned)g;
}

} // namespace pump
