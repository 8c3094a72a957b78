// hsmc.hip -- half-space cache and the half-space Monte Carlo step
// (reference: setupAsBs / cacheHalfspaces hsmc.cu:10-99,201-211,
// bvlsShortcut bvls.cu:259-273, setupValidParticles / propagateValidParticles /
// calculateCP_HSMC PUMP.cu:613-777).
//
// cacheHalfspaces: a workgroup owns a run of edges and walks their waypoints
// one per lane; obstacles are staged in LDS (any count: boxes past the LDS
// capacity are read from global memory).  It also records, per edge, which
// waypoints store a half-space (a 64-bit summary).
//
// propagate + CP comes in three forms, all with the reference's particle
// decisions (float order of PUMP.cu:645-651, no contraction; the AND over
// (waypoint, half-space) pairs is order-independent), hence identical rows:
//   * one wave per new path (the drop-in entry point and small wavefronts);
//   * one workgroup per new path with the path's (waypoint, half-space)
//     pairs staged in LDS (selected with PUMP_HSMC_KERNEL=block);
//   * the batched form for large wavefronts: classify, copy, and the
//     time-binned half-space pass of hsmc_bins.hip.
#include <algorithm>
#include <array>
#include <cstdlib>
#include <map>
#include <mutex>
#include <vector>

#include "hsmc_wave.hpp"
#include "pump_device.hpp"
#include "pump_internal.hpp"

#pragma clang fp contract(off)

namespace pump {
namespace {

using namespace wave;

constexpr int kEdgesPerBlock = 128;
constexpr int kHsBlock = 256;
constexpr int kListCap = 1024; // (waypoint, half-space) entries staged per pass

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_setup_as_bs(int64_t n, float *__restrict__ bs,
                                                     float *__restrict__ as)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        as[i * 3 + 0] = -1.0f;
        as[i * 3 + 1] = -1.0f;
        as[i * 3 + 2] = -1.0f;
        bs[i] = 1111.0f;
    }
}

constexpr int kHsObsLds = 512; // boxes staged in LDS by the half-space cache
constexpr int kStageHs = 8;    // largest max_hs whose slots are staged in LDS
constexpr int kHsMapCap = 3072; // block waypoints mapped to their edge in LDS
static_assert(kEdgesPerBlock <= 256, "edge index of the waypoint map is a byte");

// LDS of a half-space-cache workgroup, sized at launch: obs_cap staged boxes
// and, when stage_hs > 0, each wave's 64 x stage_hs slots of as/bs, written
// out as one contiguous run per wave (slots and sentinels together).
struct HsLds {
    int off, em, fr, obs, stage, wmap, bytes;
};

__host__ __device__ inline HsLds hs_lds(int obs_cap, int stage_hs)
{
    HsLds L{};
    int o = 0;
    auto take = [&o](int bytes, int align) {
        o = (o + align - 1) / align * align;
        const int at = o;
        o += bytes;
        return at;
    };
    L.off = take((kEdgesPerBlock + 1) * 8, 8);
    L.em = take(kEdgesPerBlock * 8, 8);
    L.fr = take(kEdgesPerBlock, 4);
    L.obs = take(obs_cap * 6 * 4, 4);
    L.stage = take((kHsBlock / 64) * 64 * stage_hs * 4 * 4, 16);
    L.wmap = take(kHsMapCap, 4);
    L.bytes = (o + 15) / 16 * 16;
    return L;
}

__host__ __device__ inline int hs_obs_cap(int n_obs) { return n_obs < kHsObsLds ? n_obs : kHsObsLds; }

// One wave writes a run of cnt floats from LDS to global memory: the elements
// before the first 16-byte boundary singly, the body as 16-byte stores (one
// KiB per wave instruction), the tail singly.
__device__ __forceinline__ void copy_run(float *__restrict__ g, const float *l, int cnt, int lane)
{
    const int mis = (int)((reinterpret_cast<uintptr_t>(g) >> 2) & 3);
    const int head = mis ? (4 - mis < cnt ? 4 - mis : cnt) : 0;
    if (lane < head)
        g[lane] = l[lane];
    const int n4 = (cnt - head) >> 2;
    float4 *g4 = reinterpret_cast<float4 *>(g + head);
    for (int i = lane; i < n4; i += 64) {
        const float *s = l + head + 4 * i;
        g4[i] = make_float4(s[0], s[1], s[2], s[3]);
    }
    const int done = head + 4 * n4;
    if (lane < cnt - done)
        g[done + lane] = l[done + lane];
}

// cacheHalfspaces (hsmc.cu:10-99) for the waypoints of kEdgesPerBlock
// consecutive edges, one lane per waypoint; with kSentinels the unfilled
// slots get setupAsBs's values (hsmc.cu:201-211) in the same pass.
template <bool kSentinels>
__device__ __forceinline__ void hs_block(const HsJob &J, int64_t e0, char *smem, int obs_cap,
                                         int stage_hs)
{
    const int64_t n_edges = J.n_edges;
    const float *__restrict__ disc = J.disc;
    const int64_t *__restrict__ wp_offsets = J.wp_offsets;
    const int w_fixed = J.w_fixed;
    const uint8_t *__restrict__ is_free = J.is_free;
    const int n_obs = J.n_obs, max_hs = J.max_hs;
    const float *__restrict__ obstacles = J.obstacles;
    float *__restrict__ as = J.as;
    float *__restrict__ bs = J.bs;
    int *__restrict__ count_stored = J.count;
    const float maxd2 = J.maxd2;
    uint64_t *__restrict__ edge_mask = J.edge_mask;
    const HsLds L = hs_lds(obs_cap, stage_hs);
    int64_t *off = reinterpret_cast<int64_t *>(smem + L.off);
    unsigned long long *em = reinterpret_cast<unsigned long long *>(smem + L.em);
    uint8_t *fr = reinterpret_cast<uint8_t *>(smem + L.fr);
    // The staged boxes and slots are typed as LDS pointers: where a branch
    // reads or writes either them or global memory, the compiler would
    // otherwise merge the two accesses into one flat access through a
    // selected pointer.
    typedef __attribute__((address_space(3))) float LdsFloat;
    LdsFloat *olo = (LdsFloat *)(smem + L.obs); // lo[d][b], then hi[d][b]
    LdsFloat *ohi = olo + 3 * obs_cap;
    uint8_t *wmap = reinterpret_cast<uint8_t *>(smem + L.wmap); // edge of each local waypoint
    const int ne = (int)((n_edges - e0) < kEdgesPerBlock ? (n_edges - e0) : kEdgesPerBlock);
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    const int n_staged = n_obs < obs_cap ? n_obs : obs_cap;
    // slots staged in LDS and written out per wave (sentinel pass only)
    const bool stage = kSentinels && max_hs > 0 && max_hs <= stage_hs;
    LdsFloat *sa = (LdsFloat *)(smem + L.stage) + wid * 64 * stage_hs * 4;
    LdsFloat *sb = sa + 64 * max_hs * 3;
    for (int q = t; q < n_staged * 3; q += kHsBlock) {
        const int b = q / 3, d = q - 3 * (q / 3);
        olo[d * obs_cap + b] = obstacles[b * 12 + d];
        ohi[d * obs_cap + b] = obstacles[b * 12 + 6 + d];
    }
    auto bounds = [&](int b, float *lo, float *hi) {
        if (b < n_staged) {
            for (int d = 0; d < 3; ++d) {
                lo[d] = olo[d * obs_cap + b];
                hi[d] = ohi[d * obs_cap + b];
            }
        } else {
            for (int d = 0; d < 3; ++d) {
                lo[d] = obstacles[b * 12 + d];
                hi[d] = obstacles[b * 12 + 6 + d];
            }
        }
    };
    for (int q = t; q <= ne; q += kHsBlock)
        off[q] = wp_offsets ? wp_offsets[e0 + q] : (e0 + q) * (int64_t)w_fixed;
    for (int q = t; q < ne; q += kHsBlock) {
        fr[q] = is_free[e0 + q];
        em[q] = 0ull;
    }
    __syncthreads();
    const int64_t g_begin = off[0], g_end = off[ne];
    const bool mapped = g_end - g_begin <= kHsMapCap;
    if (mapped) {
        for (int q = t; q < ne; q += kHsBlock)
            for (int64_t j = off[q]; j < off[q + 1]; ++j)
                wmap[j - g_begin] = (uint8_t)q;
        __syncthreads();
    }
    const int n_words = (n_obs + 63) >> 6;
    // A box whose distance from the tile's waypoint bounds exceeds sqrt(maxd2)
    // along one axis has d2 > maxd2 (hsmc.cu:69) at every waypoint of the
    // tile, so only the boxes near those bounds are scanned, in index order.
    // The reach exceeds sqrt(maxd2) by far more than the rounding of the
    // clamp distances; a NaN or infinite maxd2 scans every box.
    const float reach =
        (maxd2 >= 0.0f && maxd2 < INFINITY) ? sqrtf(maxd2) * 1.001f + 1e-6f : INFINITY;
    for (int64_t base = g_begin; base < g_end; base += kHsBlock) {
        const int64_t g = base + t;
        const bool have = g < g_end;
        const int le = !have ? 0 : mapped ? (int)wmap[g - g_begin] : find_edge(off, ne, g);
        const bool work = have && fr[le];
        float pos[3] = {0, 0, 0}, vel[3] = {0, 0, 0};
        if (work)
            for (int d = 0; d < 3; ++d) {
                pos[d] = disc[g * kDim + d];
                vel[d] = disc[g * kDim + 3 + d];
            }
        // slot h of this waypoint: as/bs[slot0 + h], in LDS when staged.
        // stage is uniform, and the two branches keep each store's address
        // space known (LDS or global stores, not flat ones).
        const int64_t slot0 = stage ? (int64_t)lane * max_hs : g * max_hs;
        auto put = [&](int64_t hs, float a0, float a1, float a2, float b) {
            if (stage) {
                sa[hs * 3 + 0] = a0;
                sa[hs * 3 + 1] = a1;
                sa[hs * 3 + 2] = a2;
                sb[hs] = b;
            } else {
                as[hs * 3 + 0] = a0;
                as[hs * 3 + 1] = a1;
                as[hs * 3 + 2] = a2;
                bs[hs] = b;
            }
        };
        int stored = 0;
        bool live = false; // some stored b passes the sentinel test of PUMP.cu:642
        if (__ballot(work) != 0ull) {
            float tlo[3], thi[3];
            for (int d = 0; d < 3; ++d) {
                tlo[d] = wave_min(work ? pos[d] : INFINITY) - reach;
                thi[d] = wave_max(work ? pos[d] : -INFINITY) + reach;
            }
            for (int wd = 0; wd < n_words; ++wd) {
                const int bb = wd * 64 + lane;
                bool near = bb < n_obs;
                if (near) {
                    float lo[3], hi[3];
                    bounds(bb, lo, hi);
                    for (int d = 0; d < 3; ++d)
                        near = near && !(hi[d] < tlo[d] || thi[d] < lo[d]);
                }
                unsigned long long m = __ballot(near);
                while (work && m && stored < max_hs) {
                    const int b = wd * 64 + __builtin_ctzll(m);
                    m &= m - 1;
                    float lo[3], hi[3], x[3];
                    bounds(b, lo, hi);
                    for (int d = 0; d < 3; ++d)
                        x[d] = (pos[d] <= lo[d]) ? lo[d] : (pos[d] >= hi[d]) ? hi[d] : pos[d];
                    for (int d = 0; d < 3; ++d)
                        x[d] = x[d] - pos[d];
                    float d2 = 0.0f;
                    for (int d = 0; d < 3; ++d)
                        d2 = d2 + x[d] * x[d];
                    if (d2 > maxd2)
                        continue;
                    float xdv = 0.0f, vdv = 0.0f;
                    for (int d = 0; d < 3; ++d) {
                        xdv = xdv + x[d] * vel[d];
                        vdv = vdv + vel[d] * vel[d];
                    }
                    if (vdv == 0.0f)
                        vdv = 1e-5f;
                    float ratio = xdv / vdv;
                    for (int d = 0; d < 3; ++d)
                        x[d] = x[d] - ratio * vel[d];
                    float xx = 0.0f;
                    for (int d = 0; d < 3; ++d)
                        xx = xx + x[d] * x[d];
                    put(slot0 + stored, x[0], x[1], x[2], xx);
                    live = live || !(xx >= 1000.0f);
                    ++stored;
                }
            }
        }
        if (have) {
            count_stored[g] = stored;
            if (live && edge_mask) {
                const int k = (int)(g - off[le]);
                atomicOr(&em[le], 1ull << (k < 63 ? k : 63));
            }
            if (kSentinels)
                for (int h = stored; h < max_hs; ++h)
                    put(slot0 + h, -1.0f, -1.0f, -1.0f, 1111.0f);
        }
        if (stage) {
            // the wave's waypoints are the prefix of its lanes with have set;
            // their slots are one contiguous run of as and of bs
            const int64_t w0 = base + (int64_t)wid * 64;
            const int64_t left = g_end - w0;
            const int nw = left <= 0 ? 0 : left < 64 ? (int)left : 64;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            copy_run(as + w0 * max_hs * 3, (const float *)sa, nw * max_hs * 3, lane);
            copy_run(bs + w0 * max_hs, (const float *)sb, nw * max_hs, lane);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
    if (edge_mask) {
        __syncthreads();
        for (int q = t; q < ne; q += kHsBlock)
            edge_mask[e0 + q] = em[q];
    }
}

template <bool kSentinels>
__global__ __launch_bounds__(kHsBlock) void k_cache_hs(
    int64_t n_edges, const float *__restrict__ disc, const int64_t *__restrict__ wp_offsets,
    int w_fixed, const uint8_t *__restrict__ is_free, int n_obs, int max_hs,
    const float *__restrict__ obstacles, float *__restrict__ as, float *__restrict__ bs,
    int *__restrict__ count_stored, float maxd2, uint64_t *__restrict__ edge_mask, int stage_hs)
{
    extern __shared__ __align__(16) char smem[];
    const HsJob J{n_edges, disc, wp_offsets, w_fixed, is_free, n_obs, max_hs,
                  obstacles, as, bs, count_stored, maxd2, edge_mask};
    hs_block<kSentinels>(J, (int64_t)blockIdx.x * kEdgesPerBlock, smem, hs_obs_cap(n_obs),
                         stage_hs);
}

// the half-space cache of many instances in one launch (with sentinel rows);
// LDS holds obs_cap boxes and stage_hs slots per waypoint (the largest of the
// instances, or 0)
__global__ __launch_bounds__(kHsBlock) void k_cache_hs_many(const HsJob *__restrict__ jobs,
                                                            const int64_t *__restrict__ prefix,
                                                            int n_inst, int obs_cap, int stage_hs)
{
    extern __shared__ __align__(16) char smem[];
    const int k = instance_of(prefix, n_inst, blockIdx.x);
    const HsJob J = global_job(jobs[k]);
    hs_block<true>(J, ((int64_t)blockIdx.x - prefix[k]) * kEdgesPerBlock, smem, obs_cap,
                   stage_hs);
}
static_assert(kEdgesPerBlock == kHsEdgesPerBlock, "edges per half-space cache block");

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_setup_valid(uint8_t *__restrict__ valid, int64_t n_bytes,
                                                     float *__restrict__ cp,
                                                     float *__restrict__ time, int64_t n_paths)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t words = n_bytes / 4;
    uint32_t *v32 = reinterpret_cast<uint32_t *>(valid);
    for (int64_t i = i0; i < words; i += stride)
        v32[i] = 0x01010101u;
    for (int64_t i = words * 4 + i0; i < n_bytes; i += stride)
        valid[i] = 1;
    for (int64_t i = i0; i < n_paths; i += stride) {
        cp[i] = 1.0f;
        time[i] = 0.0f;
    }
}

// one (waypoint, half-space) entry of a path's edge
struct HsEntry {
    float a0, a1, a2, b;
    int ti;
};

// Gather the entries of waypoints [i0, i1) of edge e into LDS; returns count.
// Also counts time indices clamped into [0, tmax_steps).
__device__ __forceinline__ int gather_entries(HsEntry *list, int *cnt, int i0, int i1, int64_t g0,
                                              int max_hs, const float *__restrict__ as,
                                              const float *__restrict__ bs, float tbase,
                                              float dtl, float dt, int tmax_steps,
                                              unsigned long long *clamped)
{
    if (threadIdx.x == 0)
        *cnt = 0;
    __syncthreads();
    for (int i = i0 + (int)threadIdx.x; i < i1; i += blockDim.x) {
        float time = tbase + (float)i * dtl;
        int ti = (int)(time / dt);
        bool any = false;
        const int64_t head = (g0 + i) * (int64_t)max_hs;
        for (int h = 0; h < max_hs; ++h) {
            float b = bs[head + h];
            if (b >= 1000.0f)
                continue;
            if (!any) {
                any = true;
                if (ti < 0 || ti >= tmax_steps) {
                    if (clamped)
                        atomicAdd(clamped, 1ull);
                    ti = ti < 0 ? 0 : tmax_steps - 1;
                }
            }
            int slot = atomicAdd(cnt, 1);
            HsEntry en;
            en.a0 = as[(head + h) * 3 + 0];
            en.a1 = as[(head + h) * 3 + 1];
            en.a2 = as[(head + h) * 3 + 2];
            en.b = b;
            en.ti = ti;
            list[slot] = en;
        }
    }
    __syncthreads();
    return *cnt;
}

__device__ __forceinline__ bool particle_passes(const HsEntry &en, const float *__restrict__ off,
                                                int n_mc, int p, float om)
{
    const float *o = off + ((int64_t)en.ti * n_mc + p) * 3;
    return hs_pass(o[0], o[1], o[2], en.a0, en.a1, en.a2, om, en.b);
}

// Byte-validity propagate (+ optional fused CP).  Reads parent rows of
// valid_src, writes child row dst_base+w of valid_dst.
__global__ __launch_bounds__(256) void k_propagate(
    int n_new, const int64_t *__restrict__ wp_offsets, int w_fixed,
    const int *__restrict__ wf_edge, const int *__restrict__ wf_prev,
    const float *__restrict__ as, const float *__restrict__ bs, const float *__restrict__ times,
    const float *__restrict__ topts, float dt, int max_hs, const float *__restrict__ offsets_time,
    int tmax_steps, float om, int n_mc, const uint8_t *__restrict__ valid_src,
    uint8_t *__restrict__ valid_dst, int64_t dst_base, int n_p, float *__restrict__ path_cp,
    int64_t cp_base, unsigned long long *__restrict__ clamped)
{
    __shared__ HsEntry list[kListCap];
    __shared__ int cnt;
    __shared__ int red[4];
    const int w = blockIdx.x;
    const int e = wf_edge[w];
    const int prev = wf_prev[w];
    const int64_t g0 = wp_offsets ? wp_offsets[e] : (int64_t)e * w_fixed;
    const int W = (int)((wp_offsets ? wp_offsets[e + 1] : g0 + w_fixed) - g0);
    const float dtl = topts[e] / (float)(W - 1);
    const float tbase = times[prev];
    const uint8_t *src = valid_src + (int64_t)prev * n_p;
    uint8_t *dst = valid_dst + (dst_base + w) * (int64_t)n_p;
    const int chunk = max_hs > 0 ? (kListCap / max_hs) : W;

    int invalid = 0;
    for (int i0 = 0; i0 < W; i0 += chunk) {
        const int i1 = i0 + chunk < W ? i0 + chunk : W;
        const int n = gather_entries(list, &cnt, i0, i1, g0, max_hs, as, bs, tbase, dtl, dt,
                                     tmax_steps, clamped);
        const bool last = i1 >= W;
        const uint8_t *in = (i0 == 0) ? src : dst;
        for (int p = threadIdx.x; p < n_p; p += blockDim.x) {
            bool v = in[p] != 0;
            for (int q = 0; q < n && v; ++q)
                v = particle_passes(list[q], offsets_time, n_mc, p, om);
            dst[p] = v ? 1 : 0;
            if (last)
                invalid += v ? 0 : 1;
        }
        __syncthreads();
    }
    if (path_cp) {
        for (int off = 32; off > 0; off >>= 1)
            invalid += __shfl_down(invalid, off, 64);
        if ((threadIdx.x & 63) == 0)
            red[threadIdx.x >> 6] = invalid;
        __syncthreads();
        if (threadIdx.x == 0)
            path_cp[cp_base + w] = ((float)(red[0] + red[1] + red[2] + red[3])) / (float)n_p;
    }
}

// calculateCP_HSMC: one workgroup per path
// calculateCP_HSMC: one wave per path (grid-stride), 16-byte loads
__global__ __launch_bounds__(256) void k_cp(int n_new, int64_t base, float *__restrict__ cp,
                                            const uint8_t *__restrict__ valid, int n_p)
{
    const int lane = threadIdx.x & 63;
    const int64_t n_waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t w = wave_index(); w < n_new; w += n_waves) {
        const int c = wave_sum(invalid_bytes(valid + (base + w) * (int64_t)n_p, n_p, lane));
        if (lane == 0)
            cp[base + w] = ((float)c) / (float)n_p;
    }
}

__global__ __launch_bounds__(256) void k_gather_rows(const uint8_t *__restrict__ src,
                                                     const int *__restrict__ rows, int n_rows,
                                                     uint8_t *__restrict__ dst, int64_t dst_base,
                                                     int64_t row_bytes)
{
    const int k = blockIdx.x;
    const uint8_t *s = src + (int64_t)rows[k] * row_bytes;
    uint8_t *d = dst + (dst_base + k) * row_bytes;
    const bool aligned = ((row_bytes & 15) == 0) && ((((uintptr_t)src) & 15) == 0) &&
                         ((((uintptr_t)dst) & 15) == 0);
    if (aligned) {
        const uint4 *s4 = reinterpret_cast<const uint4 *>(s);
        uint4 *d4 = reinterpret_cast<uint4 *>(d);
        for (int64_t q = threadIdx.x; q < row_bytes / 16; q += blockDim.x)
            d4[q] = s4[q];
    } else {
        for (int64_t q = threadIdx.x; q < row_bytes; q += blockDim.x)
            d[q] = s[q];
    }
}

// ---------------------------------------------------------------------------
// Bit-mask validity (chosen form): one wave handles 64 particles per step;
// the parent word is a uniform load and the child word is the ballot of the
// lanes' decisions.
__global__ __launch_bounds__(256) void k_propagate_bits(
    int n_new, const int64_t *__restrict__ wp_offsets, int w_fixed,
    const int *__restrict__ wf_edge, const int *__restrict__ wf_prev,
    const float *__restrict__ as, const float *__restrict__ bs, const float *__restrict__ times,
    const float *__restrict__ topts, float dt, int max_hs, const float *__restrict__ offsets_time,
    int tmax_steps, float om, int n_mc, const uint64_t *__restrict__ valid_src,
    uint64_t *__restrict__ valid_dst, int64_t dst_base, int n_p, float *__restrict__ path_cp,
    int64_t cp_base, unsigned long long *__restrict__ clamped)
{
    __shared__ HsEntry list[kListCap];
    __shared__ int cnt;
    __shared__ int red[4];
    const int w = blockIdx.x;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int e = wf_edge[w];
    const int prev = wf_prev[w];
    const int64_t g0 = wp_offsets ? wp_offsets[e] : (int64_t)e * w_fixed;
    const int W = (int)((wp_offsets ? wp_offsets[e + 1] : g0 + w_fixed) - g0);
    const float dtl = topts[e] / (float)(W - 1);
    const float tbase = times[prev];
    const int nwords = (n_p + 63) / 64;
    const uint64_t *src = valid_src + (int64_t)prev * nwords;
    uint64_t *dst = valid_dst + (dst_base + w) * (int64_t)nwords;
    const int chunk = max_hs > 0 ? (kListCap / max_hs) : W;

    int valid_count = 0;
    for (int i0 = 0; i0 < W; i0 += chunk) {
        const int i1 = i0 + chunk < W ? i0 + chunk : W;
        const int n = gather_entries(list, &cnt, i0, i1, g0, max_hs, as, bs, tbase, dtl, dt,
                                     tmax_steps, clamped);
        const bool last = i1 >= W;
        const uint64_t *in = (i0 == 0) ? src : dst;
        for (int wd = wid; wd < nwords; wd += 4) {
            const uint64_t word = in[wd];
            const int p = wd * 64 + lane;
            bool v = (p < n_p) && ((word >> lane) & 1ull);
            if (__ballot(v) != 0ull) {
                for (int q = 0; q < n; ++q) {
                    if (v)
                        v = particle_passes(list[q], offsets_time, n_mc, p, om);
                }
            }
            const uint64_t out = __ballot(v);
            if (lane == 0)
                dst[wd] = out;
            if (last)
                valid_count += (lane == 0) ? __popcll(out) : 0;
        }
        __syncthreads();
    }
    if (path_cp) {
        if (lane == 0)
            red[wid] = valid_count;
        __syncthreads();
        if (threadIdx.x == 0) {
            int valid_all = red[0] + red[1] + red[2] + red[3];
            path_cp[cp_base + w] = ((float)(n_p - valid_all)) / (float)n_p;
        }
    }
}

// ---------------------------------------------------------------------------
// Wave-level forms of propagate + CP.  A wave first marks, per group of 64
// waypoints, the waypoints that hold at least one stored half-space (b <
// 1000): from the edge's waypoint mask when one is given (bit k: waypoint k,
// bit 63: some waypoint >= 63), else by ballots over bs.  Particles are then
// streamed 64 at a time (lane = particle) and tested only against those
// waypoints' half-spaces, whose data are wave-uniform loads.  Decisions use
// the reference's float order.  No LDS, no workgroup barriers.
// One wave per new path (single launch; small wavefronts).
// One path's propagate + CP by one wave (the body of the wave-per-path
// forms): marked waypoints from the edge's summary or from bs, particles 64 at
// a time, the child row written in full, the CP from its final count.
template <bool kBits>
__device__ __forceinline__ void propagate_path(
    int64_t w, int e, int prev, const PathInfo &pi, const float *__restrict__ as,
    const float *__restrict__ bs, const uint64_t *__restrict__ edge_mask, float dt, int max_hs,
    const float *__restrict__ offsets_time, int tmax_steps, float om, int n_mc,
    const void *__restrict__ valid_src, void *__restrict__ valid_dst, int64_t dst_base, int n_p,
    float *__restrict__ path_cp, int64_t cp_base, unsigned long long *__restrict__ clamped)
{
    // eight particle groups per round: a path's particles in few rounds of
    // memory latency
    constexpr int NQ = 2 * kWalkGroups;
    const int lane = threadIdx.x & 63;
    const uint64_t emask = edge_mask ? edge_mask[e] : 0ull;
    int count = 0; // valid particles (bits) or invalid particles (bytes), per lane
    // waypoints in groups of 256; later groups continue from the child row
    for (int i0 = 0; i0 < pi.W; i0 += 256) {
        uint64_t masks[4];
        const uint64_t any = hs_masks(masks, pi, i0, edge_mask != nullptr, emask, max_hs, bs,
                                      dt, tmax_steps, clamped);
        const bool last = i0 + 256 >= pi.W;
        if (kBits) {
            const int nwords = (n_p + 63) >> 6;
            const uint64_t *src =
                i0 == 0 ? (const uint64_t *)valid_src + (int64_t)prev * nwords
                        : (const uint64_t *)valid_dst + (dst_base + w) * (int64_t)nwords;
            uint64_t *dst = (uint64_t *)valid_dst + (dst_base + w) * (int64_t)nwords;
            if (!any) {
                for (int wd = lane; wd < nwords; wd += 64) {
                    const uint64_t x = src[wd];
                    if (i0 == 0)
                        dst[wd] = x;
                    if (last)
                        count += __popcll(x);
                }
            } else {
                for (int wd0 = 0; wd0 < nwords; wd0 += NQ) {
                    bool v[NQ];
                    int p[NQ];
                    uint64_t alive = 0;
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        const int wd = wd0 + q;
                        p[q] = wd * 64 + lane;
                        const uint64_t word = wd < nwords ? src[wd] : 0ull;
                        v[q] = (p[q] < n_p) && ((word >> lane) & 1ull);
                        alive |= __ballot(v[q]);
                    }
                    if (alive)
                        particle_walk<NQ>(v, p, n_p, masks, i0, pi, max_hs, as, bs,
                                          offsets_time, n_mc, tmax_steps, dt, om);
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        const uint64_t out = __ballot(v[q]);
                        if (lane == q && wd0 + q < nwords) {
                            dst[wd0 + q] = out;
                            if (last)
                                count += __popcll(out);
                        }
                    }
                }
            }
        } else {
            const uint8_t *src = i0 == 0
                                     ? (const uint8_t *)valid_src + (int64_t)prev * n_p
                                     : (const uint8_t *)valid_dst + (dst_base + w) * (int64_t)n_p;
            uint8_t *dst = (uint8_t *)valid_dst + (dst_base + w) * (int64_t)n_p;
            const bool words = ((n_p & 3) == 0) &&
                               ((((uintptr_t)src) | ((uintptr_t)dst)) & 3) == 0;
            const bool wide = ((n_p & 15) == 0) &&
                              ((((uintptr_t)src) | ((uintptr_t)dst)) & 15) == 0;
            if (!any && wide) {
                const uint4 *s16 = (const uint4 *)src;
                uint4 *d16 = (uint4 *)dst;
                for (int q = lane; q < (n_p >> 4); q += 64) {
                    uint4 x = s16[q];
                    x.x = nonzero_bytes01(x.x);
                    x.y = nonzero_bytes01(x.y);
                    x.z = nonzero_bytes01(x.z);
                    x.w = nonzero_bytes01(x.w);
                    d16[q] = x;
                    if (last)
                        count += 16 - __popc(x.x) - __popc(x.y) - __popc(x.z) - __popc(x.w);
                }
            } else if (!any && words) {
                const uint32_t *s4 = (const uint32_t *)src;
                uint32_t *d4 = (uint32_t *)dst;
                const int nq = n_p >> 2;
                for (int q = lane; q < nq; q += 64) {
                    const uint32_t x = nonzero_bytes01(s4[q]);
                    d4[q] = x;
                    if (last)
                        count += 4 - __popc(x);
                }
            } else {
                for (int p0 = 0; p0 < n_p; p0 += 64 * NQ) {
                    bool v[NQ];
                    int p[NQ];
                    uint64_t alive = 0;
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        p[q] = p0 + q * 64 + lane;
                        v[q] = (p[q] < n_p) && src[p[q]] != 0;
                        alive |= __ballot(v[q]);
                    }
                    if (any && alive)
                        particle_walk<NQ>(v, p, n_p, masks, i0, pi, max_hs, as, bs,
                                          offsets_time, n_mc, tmax_steps, dt, om);
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        if (p[q] < n_p) {
                            dst[p[q]] = v[q] ? 1 : 0;
                            if (last)
                                count += v[q] ? 0 : 1;
                        }
                    }
                }
            }
        }
    }
    count = wave_sum(count);
    if (path_cp && lane == 0) {
        const int invalid = kBits ? n_p - count : count;
        path_cp[cp_base + w] = ((float)invalid) / (float)n_p;
    }
}

// One wave per new path (single launch; small wavefronts).
template <bool kBits>
__global__ __launch_bounds__(256) void k_propagate_wave(
    int n_new, const int64_t *__restrict__ wp_offsets, int w_fixed,
    const int *__restrict__ wf_edge, const int *__restrict__ wf_prev,
    const float *__restrict__ as, const float *__restrict__ bs,
    const uint64_t *__restrict__ edge_mask, const float *__restrict__ times,
    const float *__restrict__ topts, float dt, int max_hs, const float *__restrict__ offsets_time,
    int tmax_steps, float om, int n_mc, const void *__restrict__ valid_src,
    void *__restrict__ valid_dst, int64_t dst_base, int n_p, float *__restrict__ path_cp,
    int64_t cp_base, unsigned long long *__restrict__ clamped)
{
    const int64_t n_waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t w = wave_index(); w < n_new; w += n_waves) {
        const int e = wf_edge[w];
        const int prev = wf_prev[w];
        const PathInfo pi = path_info(e, prev, wp_offsets, w_fixed, topts, times);
        propagate_path<kBits>(w, e, prev, pi, as, bs, edge_mask, dt, max_hs, offsets_time,
                              tmax_steps, om, n_mc, valid_src, valid_dst, dst_base, n_p, path_cp,
                              cp_base, clamped);
    }
}

// ---------------------------------------------------------------------------
// Batched form for large wavefronts:
//  (1) k_classify: lane = path; a path needs half-space tests iff its edge
//      has a marked waypoint; those paths are appended to a list of path
//      records (one atomic per workgroup) and their marked waypoints counted
//      per time index;
//  (2) the time-binned pass (hsmc_bins.hip) scans that histogram;
//  (3) k_copy_rows: parent rows copied to the child rows by a flat, fully
//      parallel kernel that also counts the unlisted paths' particles;
//  (4) the time-binned pass tests the listed paths' (waypoint, time index)
//      pairs and ANDs the results into their rows, then finishes their CP.
// When the pairs overflow the workspace, (4) is replaced by the group form:
// one wave per (listed path, 512 particles).
constexpr int kClassifyBlock = 1024;

__device__ __forceinline__ void classify_body(
    int n_new, const int64_t *__restrict__ wp_offsets, int w_fixed,
    const int *__restrict__ wf_edge, const int *__restrict__ wf_prev,
    const float *__restrict__ bs, const uint64_t *__restrict__ edge_mask, int max_hs,
    const float *__restrict__ times, const float *__restrict__ topts, float dt, int tmax_steps,
    uint8_t *__restrict__ has_hs, HsPath *__restrict__ recs, int *__restrict__ ctl,
    int *__restrict__ hist, unsigned long long *__restrict__ acc,
    unsigned long long *__restrict__ clamped, int *__restrict__ row_src,
    const float *__restrict__ parent_cp, float *__restrict__ path_cp, int64_t cp_base,
    const VGrid &G)
{
    __shared__ int wcount[kClassifyBlock / 64];
    __shared__ int block_base;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int64_t stride = G.nb * kClassifyBlock;
    // workgroup-uniform trip count: every thread reaches the barriers below
    for (int64_t base = G.bid * kClassifyBlock; base < n_new; base += stride) {
        const int64_t w = base + threadIdx.x;
        bool has = false;
        int n_clamped = 0;
        HsPath r;
        if (w < n_new) {
            const int e = wf_edge[w];
            // an all-clear summary settles the path with one load
            const uint64_t em0 = edge_mask ? edge_mask[e] : ~0ull;
            if (em0 != 0ull) {
                r.prev = wf_prev[w];
                const PathInfo pi = path_info(e, r.prev, wp_offsets, w_fixed, topts, times);
                r.emask = edge_mask ? mask_bits(em0, pi.W) : 0ull;
                for_each_marked(pi, edge_mask != nullptr, r.emask, bs, max_hs, [&](int i) {
                    has = true;
                    const int ti = time_index(pi, i, dt);
                    const int tc = clamp_time(ti, tmax_steps);
                    n_clamped += ti != tc ? 1 : 0;
                    if (hist)
                        atomicAdd(&hist[tc * kSubBins + (int)(w & (kSubBins - 1))], 1);
                });
                r.g0 = pi.g0;
                r.W = pi.W;
                r.dtl = pi.dtl;
                r.tbase = pi.tbase;
                r.w = (int)w;
                r.pad = 0;
            }
            has_hs[w] = has ? 1 : 0;
            // copy-on-write: a path that meets no half-space keeps its
            // parent's row and CP (the same count, the same float)
            if (row_src) {
                if (has) {
                    row_src[w] = (int)w;
                } else {
                    const int prev = wf_prev[w];
                    row_src[w] = -1 - prev;
                    if (path_cp)
                        path_cp[cp_base + w] = parent_cp[prev];
                }
            }
        }
        if (n_clamped && clamped)
            atomicAdd(clamped, (unsigned long long)n_clamped);
        const uint64_t m = __ballot(has);
        if (lane == 0)
            wcount[wid] = __popcll(m);
        __syncthreads();
        if (threadIdx.x == 0) {
            int total = 0;
            for (int k = 0; k < kClassifyBlock / 64; ++k) {
                const int c = wcount[k];
                wcount[k] = total;
                total += c;
            }
            block_base = total ? atomicAdd(&ctl[kCtlListed], total) : 0;
        }
        __syncthreads();
        if (has) {
            const int slot = block_base + wcount[wid] + __popcll(m & ((1ull << lane) - 1ull));
            recs[slot] = r;
            acc[slot] = 0ull;
        }
        __syncthreads(); // wcount and block_base are reused by the next pass
    }
}

__global__ __launch_bounds__(kClassifyBlock) void k_classify(
    int n_new, const int64_t *__restrict__ wp_offsets, int w_fixed,
    const int *__restrict__ wf_edge, const int *__restrict__ wf_prev,
    const float *__restrict__ bs, const uint64_t *__restrict__ edge_mask, int max_hs,
    const float *__restrict__ times, const float *__restrict__ topts, float dt, int tmax_steps,
    uint8_t *__restrict__ has_hs, HsPath *__restrict__ recs, int *__restrict__ ctl,
    int *__restrict__ hist, unsigned long long *__restrict__ acc,
    unsigned long long *__restrict__ clamped, int *__restrict__ row_src,
    const float *__restrict__ parent_cp, float *__restrict__ path_cp, int64_t cp_base)
{
    classify_body(n_new, wp_offsets, w_fixed, wf_edge, wf_prev, bs, edge_mask, max_hs, times,
                  topts, dt, tmax_steps, has_hs, recs, ctl, hist, acc, clamped, row_src,
                  parent_cp, path_cp, cp_base, hw_grid());
}

// Copy unit per lane: 16-byte chunks (bit masks or byte rows), 8-byte words
// or 4/1-byte chunks of a byte row.  Each path is served by a segment of S
// lanes (S a power of two, at most 64); a segment takes kCopyRows paths per
// pass so that their loads are in flight together, and sums each path's count
// with shuffles of width S.  Listed paths' rows are copied too when the
// time-binned pass runs (ctl given and no overflow); their CP is written by
// that pass.
constexpr int kCopyRows = 4;

template <int kUnit>
struct CopyWord;
template <>
struct CopyWord<16> {
    using T = uint4;
};
template <>
struct CopyWord<8> {
    using T = uint64_t;
};
template <>
struct CopyWord<4> {
    using T = uint32_t;
};
template <>
struct CopyWord<1> {
    using T = uint8_t;
};

// byte rows: normalise to 0/1 and count zeros; bit rows: count ones
template <bool kBits>
__device__ __forceinline__ int copy_count(uint4 &x)
{
    if (kBits)
        return __popc(x.x) + __popc(x.y) + __popc(x.z) + __popc(x.w);
    x.x = nonzero_bytes01(x.x);
    x.y = nonzero_bytes01(x.y);
    x.z = nonzero_bytes01(x.z);
    x.w = nonzero_bytes01(x.w);
    return 16 - __popc(x.x) - __popc(x.y) - __popc(x.z) - __popc(x.w);
}
template <bool kBits>
__device__ __forceinline__ int copy_count(uint64_t &x)
{
    return __popcll(x); // 8-byte units are used for bit rows only
}
template <bool kBits>
__device__ __forceinline__ int copy_count(uint32_t &x)
{
    x = nonzero_bytes01(x);
    return 4 - __popc(x);
}
template <bool kBits>
__device__ __forceinline__ int copy_count(uint8_t &x)
{
    x = x ? 1 : 0;
    return 1 - x;
}

template <int kUnit, bool kBits>
__global__ __launch_bounds__(256) void k_copy_rows(int n_new, const int *__restrict__ wf_prev,
                                                   const uint8_t *__restrict__ has_hs,
                                                   const int *__restrict__ ctl,
                                                   const uint8_t *__restrict__ src,
                                                   uint8_t *__restrict__ dst, int64_t dst_base,
                                                   int64_t row_bytes, int n_p, int seg,
                                                   float *__restrict__ path_cp, int64_t cp_base)
{
    using T = typename CopyWord<kUnit>::T;
    constexpr int R = kCopyRows;
    const int lane = threadIdx.x & 63;
    const int units = (int)(row_bytes / kUnit);
    const int per_wave = 64 / seg;
    const int64_t n_waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    const int sub = lane / seg, j0 = lane % seg;
    const bool copy_listed = ctl && ctl[kCtlOverflow] == 0;
    for (int64_t wb = wave_index() * per_wave * R; wb < n_new; wb += n_waves * per_wave * R) {
        const T *sp[R];
        T *dp[R];
        bool mine[R], listed[R];
        int count[R];
        int64_t w[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            w[r] = wb + r * per_wave + sub;
            const bool in = w[r] < n_new;
            // flag and parent index load together (no dependent round trip)
            const int prev = in ? wf_prev[w[r]] : 0;
            listed[r] = in && has_hs[w[r]];
            mine[r] = in && (!listed[r] || copy_listed);
            sp[r] = reinterpret_cast<const T *>(src + (int64_t)prev * row_bytes);
            dp[r] = reinterpret_cast<T *>(dst + (dst_base + (in ? w[r] : 0)) * row_bytes);
            count[r] = 0;
        }
        for (int j = j0; j < units; j += seg) {
            T x[R];
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (mine[r])
                    x[r] = sp[r][j];
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (mine[r]) {
                    count[r] += copy_count<kBits>(x[r]);
                    dp[r][j] = x[r];
                }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            for (int off = seg >> 1; off > 0; off >>= 1)
                count[r] += __shfl_down(count[r], off, seg);
            if (mine[r] && !listed[r] && j0 == 0 && path_cp) {
                const int invalid = kBits ? n_p - count[r] : count[r];
                path_cp[cp_base + w[r]] = ((float)invalid) / (float)n_p;
            }
        }
    }
}

// 64 new paths per wave iteration: lane j loads its path's edge data and
// finds whether it meets a stored half-space (the edge's summary, or its
// slots' b); the rows of the paths that meet none are copied eight at a time
// by the whole wave (their loads in flight together), and only the others
// run the per-path body.  The same rows and CPs as one wave per path,
// without a round of dependent loads per path.
template <bool kBits>
__global__ __launch_bounds__(256) void k_propagate_b64(
    int n_new, const int64_t *__restrict__ wp_offsets, int w_fixed,
    const int *__restrict__ wf_edge, const int *__restrict__ wf_prev,
    const float *__restrict__ as, const float *__restrict__ bs,
    const uint64_t *__restrict__ edge_mask, const float *__restrict__ times,
    const float *__restrict__ topts, float dt, int max_hs, const float *__restrict__ offsets_time,
    int tmax_steps, float om, int n_mc, const void *__restrict__ valid_src,
    void *__restrict__ valid_dst, int64_t dst_base, int n_p, float *__restrict__ path_cp,
    int64_t cp_base, unsigned long long *__restrict__ clamped, int unit)
{
    constexpr int R = 8;
    const int lane = threadIdx.x & 63;
    const int64_t n_waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    const int64_t row_bytes = kBits ? (int64_t)((n_p + 63) >> 6) * 8 : (int64_t)n_p;
    const int units = (int)(row_bytes / unit);
    // wave b takes paths b, b + nb, b + 2 nb, ...: the paths that meet a
    // half-space come in runs (edges near boxes), and a stride spreads each
    // run over many waves instead of queueing it on one
    const int64_t nb = (n_new + 63) / 64;
    for (int64_t bt = wave_index(); bt < nb; bt += n_waves) {
        const int64_t w = bt + (int64_t)lane * nb;
        const bool act = w < n_new;
        int e = 0, prev = 0;
        PathInfo pi{0, 2, 0.0f, 0.0f};
        bool has = false;
        if (act) {
            e = wf_edge[w];
            prev = wf_prev[w];
            pi = path_info(e, prev, wp_offsets, w_fixed, topts, times);
            if (edge_mask) {
                has = mask_bits(edge_mask[e], pi.W) != 0ull;
            } else {
                // the edge's slots are contiguous: 16 independent loads per
                // round, so a long edge costs a few rounds of latency
                const int64_t n_slots = (int64_t)pi.W * max_hs;
                const float *b0 = bs + pi.g0 * (int64_t)max_hs;
                for (int64_t c = 0; c < n_slots && !has; c += 16) {
                    float b[16];
#pragma unroll
                    for (int u = 0; u < 16; ++u)
                        b[u] = c + u < n_slots ? b0[c + u] : 1111.0f;
                    bool any = false;
#pragma unroll
                    for (int u = 0; u < 16; ++u)
                        any |= !(b[u] >= 1000.0f);
                    has = any;
                }
            }
        }
        // rows of the paths that meet no half-space: copies of their parents
        uint64_t m = __ballot(act && !has);
        while (m) {
            int l[R];
            for (int r = 0; r < R; ++r) {
                l[r] = m ? __builtin_ctzll(m) : -1;
                m &= m ? m - 1 : 0ull;
            }
            const uint8_t *sp[R];
            uint8_t *dp[R];
            int count[R];
            for (int r = 0; r < R; ++r) {
                const int pr = l[r] >= 0 ? __builtin_amdgcn_readlane(prev, l[r]) : 0;
                sp[r] = (const uint8_t *)valid_src + (int64_t)pr * row_bytes;
                dp[r] = (uint8_t *)valid_dst +
                        (dst_base + bt + (int64_t)(l[r] >= 0 ? l[r] : 0) * nb) * row_bytes;
                count[r] = 0;
            }
            if (unit == 16) {
                for (int j = lane; j < units; j += 64) {
                    uint4 x[R];
                    for (int r = 0; r < R; ++r)
                        if (l[r] >= 0)
                            x[r] = ((const uint4 *)sp[r])[j];
                    for (int r = 0; r < R; ++r)
                        if (l[r] >= 0) {
                            count[r] += copy_count<kBits>(x[r]);
                            ((uint4 *)dp[r])[j] = x[r];
                        }
                }
            } else if (kBits) { // 8-byte words
                for (int j = lane; j < units; j += 64) {
                    uint64_t x[R];
                    for (int r = 0; r < R; ++r)
                        if (l[r] >= 0)
                            x[r] = ((const uint64_t *)sp[r])[j];
                    for (int r = 0; r < R; ++r)
                        if (l[r] >= 0) {
                            count[r] += __popcll(x[r]);
                            ((uint64_t *)dp[r])[j] = x[r];
                        }
                }
            } else { // single bytes
                for (int j = lane; j < units; j += 64) {
                    uint8_t x[R];
                    for (int r = 0; r < R; ++r)
                        if (l[r] >= 0)
                            x[r] = sp[r][j];
                    for (int r = 0; r < R; ++r)
                        if (l[r] >= 0) {
                            count[r] += copy_count<false>(x[r]);
                            dp[r][j] = x[r];
                        }
                }
            }
            if (path_cp)
                for (int r = 0; r < R; ++r) {
                    const int c = wave_sum(count[r]);
                    if (l[r] >= 0 && lane == 0) {
                        const int invalid = kBits ? n_p - c : c;
                        path_cp[cp_base + bt + (int64_t)l[r] * nb] = ((float)invalid) / (float)n_p;
                    }
                }
        }
        // the paths that meet a half-space, one at a time
        m = __ballot(act && has);
        while (m) {
            const int l = __builtin_ctzll(m);
            m &= m - 1;
            PathInfo q;
            q.g0 = ((int64_t)__builtin_amdgcn_readlane((int)(pi.g0 >> 32), l) << 32) |
                   (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)pi.g0, l);
            q.W = __builtin_amdgcn_readlane(pi.W, l);
            q.dtl = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(pi.dtl), l));
            q.tbase = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(pi.tbase), l));
            propagate_path<kBits>(bt + (int64_t)l * nb, __builtin_amdgcn_readlane(e, l),
                                  __builtin_amdgcn_readlane(prev, l), q, as, bs, edge_mask, dt,
                                  max_hs, offsets_time, tmax_steps, om, n_mc, valid_src, valid_dst,
                                  dst_base, n_p, path_cp, cp_base, clamped);
        }
    }
}

// Copy-on-write form: the listed paths' child rows start as their parents'
// rows (the half-space pass ANDs into them).  One wave per listed path.
template <int kUnit>
__device__ __forceinline__ void copy_listed_body(const HsPath *__restrict__ recs,
                                                 const int *__restrict__ ctl,
                                                 const uint8_t *__restrict__ src,
                                                 uint8_t *__restrict__ dst, int64_t dst_base,
                                                 int64_t row_bytes, const VGrid &G)
{
    using T = typename CopyWord<kUnit>::T;
    if (ctl[kCtlOverflow]) // the group form writes whole rows itself
        return;
    const int lane = threadIdx.x & 63;
    const int n = ctl[kCtlListed];
    const int units = (int)(row_bytes / kUnit);
    const int64_t n_waves = wave_count(G);
    for (int64_t slot = wave_index(G); slot < n; slot += n_waves) {
        const HsPath r = recs[slot];
        const T *sp = reinterpret_cast<const T *>(src + (int64_t)r.prev * row_bytes);
        T *dp = reinterpret_cast<T *>(dst + (dst_base + r.w) * row_bytes);
        for (int j = lane; j < units; j += 64)
            dp[j] = sp[j];
    }
}

template <int kUnit>
__global__ __launch_bounds__(256) void k_copy_listed(const HsPath *__restrict__ recs,
                                                     const int *__restrict__ ctl,
                                                     const uint8_t *__restrict__ src,
                                                     uint8_t *__restrict__ dst, int64_t dst_base,
                                                     int64_t row_bytes)
{
    copy_listed_body<kUnit>(recs, ctl, src, dst, dst_base, row_bytes, hw_grid());
}

// The classification, row copies and workspace clearing of many problems'
// copy-on-write steps (see McJob); workgroup b serves problem map[b] >> 16.
__device__ __forceinline__ const McJob &mc_job(const McJob *__restrict__ jobs,
                                               const uint32_t *__restrict__ map, int stage,
                                               VGrid &G)
{
    const uint32_t m = map[blockIdx.x];
    const McJob &J = jobs[m >> 16];
    G = VGrid{(int64_t)(m & 0xffffu), (int64_t)J.nb[stage]};
    return J;
}

__global__ __launch_bounds__(256) void k_zero_many(const McJob *__restrict__ jobs,
                                                   const uint32_t *__restrict__ map)
{
    VGrid G;
    const McJob J = global_job(mc_job(jobs, map, kMcZero, G));
    int *p = J.B.ctl;
    for (int64_t i = G.bid * blockDim.x + threadIdx.x; i < J.n_zero; i += G.nb * blockDim.x)
        p[i] = 0;
}

__global__ __launch_bounds__(kClassifyBlock) void k_classify_many(
    const McJob *__restrict__ jobs, const uint32_t *__restrict__ map)
{
    VGrid G;
    const McJob J = global_job(mc_job(jobs, map, kMcClassify, G));
    const BinArgs &a = J.B;
    classify_body(a.n_new, J.wp_offsets, 0, J.wf_edge, J.wf_prev, a.bs, J.edge_mask, a.max_hs,
                  J.times, J.topts, a.dt, a.tmax_steps, J.has, const_cast<HsPath *>(a.recs),
                  a.ctl, a.hist, a.acc, J.clamped, J.row_src, J.parent_cp, a.path_cp, a.cp_base,
                  G);
}

__global__ __launch_bounds__(256) void k_copy_listed_many(const McJob *__restrict__ jobs,
                                                          const uint32_t *__restrict__ map)
{
    VGrid G;
    const McJob J = global_job(mc_job(jobs, map, kMcCopy, G));
    const BinArgs &a = J.B;
    copy_listed_body<16>(a.recs, a.ctl, (const uint8_t *)a.valid_src, (uint8_t *)a.valid_dst,
                         a.dst_base, J.row_bytes, G);
}

__global__ __launch_bounds__(256) void k_row_src_iota(int n, int *__restrict__ row_src)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        row_src[i] = (int)i;
}

// Materialises selected children of a copy-on-write step: child sel[j] goes
// to row dst_base + j, from its own row (row_src >= 0) or its parent's.
// Byte rows are normalised to 0/1 as the copy form writes them.  Several rows
// per workgroup, one segment of lanes per row.
__global__ __launch_bounds__(256) void k_gather_cow(const uint8_t *__restrict__ children,
                                                    const uint8_t *__restrict__ parents,
                                                    const int *__restrict__ row_src,
                                                    const int *__restrict__ sel, int n_sel,
                                                    uint8_t *__restrict__ dst, int64_t dst_base,
                                                    int64_t row_bytes, int bits, int seg)
{
    const int per_block = blockDim.x / seg;
    const int sub = threadIdx.x / seg, j0 = threadIdx.x % seg;
    const bool wide = ((row_bytes & 15) == 0) &&
                      ((((uintptr_t)children) | ((uintptr_t)parents) | ((uintptr_t)dst)) & 15) == 0;
    for (int64_t k = (int64_t)blockIdx.x * per_block + sub; k < n_sel;
         k += (int64_t)gridDim.x * per_block) {
        const int i = sel ? sel[k] : (int)k;
        const int r = row_src[i];
        const uint8_t *s = r >= 0 ? children + (int64_t)r * row_bytes
                                  : parents + (int64_t)(-1 - r) * row_bytes;
        uint8_t *d = dst + (dst_base + k) * row_bytes;
        if (wide) {
            const uint4 *s4 = reinterpret_cast<const uint4 *>(s);
            uint4 *d4 = reinterpret_cast<uint4 *>(d);
            for (int64_t q = j0; q < row_bytes / 16; q += seg) {
                uint4 x = s4[q];
                if (!bits)
                    (void)copy_count<false>(x);
                d4[q] = x;
            }
        } else {
            for (int64_t q = j0; q < row_bytes; q += seg)
                d[q] = bits ? s[q] : (s[q] ? 1 : 0);
        }
    }
}

__global__ __launch_bounds__(256) void k_setup_bits(uint64_t *__restrict__ v, int64_t n_paths,
                                                    int n_p)
{
    const int nwords = (n_p + 63) / 64;
    const int64_t total = n_paths * nwords;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        int wd = (int)(i % nwords);
        int rem = n_p - wd * 64;
        v[i] = rem >= 64 ? ~0ull : ((1ull << rem) - 1ull);
    }
}

} // namespace
} // namespace pump

using namespace pump;

extern "C" int pump_setup_as_bs(int64_t hs_count, float *d_bs, float *d_as, pump_stream_t stream)
{
    PUMP_REQUIRE(hs_count >= 0);
    if (hs_count == 0)
        return PUMP_OK;
    PUMP_REQUIRE(d_bs && d_as);
    hipLaunchKernelGGL(k_setup_as_bs, dim3(grid_for(hs_count, 256)), dim3(256), 0,
                       as_stream(stream), hs_count, d_bs, d_as);
    PUMP_LAUNCHED();
    return PUMP_OK;
}

extern "C" float pump_maxd2(float offset_mult)
{
    return (float)(0.1 * (double)offset_mult * (double)offset_mult);
}

extern "C" int pump_cache_halfspaces(int64_t n_edges, const float *d_disc_motions,
                                     const uint8_t *d_is_free_edges, int n_obs, int max_hs,
                                     const float *d_obstacles, int waypoints_count,
                                     const int64_t *d_wp_offsets, float *d_as, float *d_bs,
                                     int *d_count_stored, float maxd2, int write_sentinels,
                                     uint64_t *d_edge_hs_mask, pump_stream_t stream)
{
    PUMP_REQUIRE(n_edges >= 0 && n_obs >= 0 && max_hs >= 0);
    if (n_edges == 0)
        return PUMP_OK;
    PUMP_REQUIRE(d_disc_motions && d_is_free_edges && d_as && d_bs && d_count_stored);
    PUMP_REQUIRE(n_obs == 0 || d_obstacles);
    PUMP_REQUIRE(d_wp_offsets || waypoints_count >= 1);
    int64_t blocks = (n_edges + kEdgesPerBlock - 1) / kEdgesPerBlock;
    if (blocks > 0x7fffffff)
        return PUMP_E_RANGE;
    const int stage_hs = (write_sentinels && max_hs <= kStageHs) ? max_hs : 0;
    const size_t lds = (size_t)hs_lds(hs_obs_cap(n_obs), stage_hs).bytes;
    if (write_sentinels)
        hipLaunchKernelGGL(k_cache_hs<true>, dim3((unsigned)blocks), dim3(kHsBlock), lds,
                           as_stream(stream), n_edges, d_disc_motions, d_wp_offsets,
                           waypoints_count, d_is_free_edges, n_obs, max_hs, d_obstacles, d_as,
                           d_bs, d_count_stored, maxd2, d_edge_hs_mask, stage_hs);
    else
        hipLaunchKernelGGL(k_cache_hs<false>, dim3((unsigned)blocks), dim3(kHsBlock), lds,
                           as_stream(stream), n_edges, d_disc_motions, d_wp_offsets,
                           waypoints_count, d_is_free_edges, n_obs, max_hs, d_obstacles, d_as,
                           d_bs, d_count_stored, maxd2, d_edge_hs_mask, 0);
    PUMP_LAUNCHED();
    return PUMP_OK;
}

int pump::cache_hs_many(const HsJob *d_jobs, const int64_t *d_block_prefix, int n_inst,
                       int64_t n_blocks, int max_obs, int max_hs, hipStream_t s)
{
    PUMP_REQUIRE(d_jobs && d_block_prefix && n_inst > 0 && n_blocks >= 0 &&
                 n_blocks <= 0x7fffffff && max_obs >= 0 && max_hs >= 0);
    if (n_blocks == 0)
        return PUMP_OK;
    const int cap = hs_obs_cap(max_obs);
    const int stage_hs = max_hs <= kStageHs ? max_hs : 0;
    const size_t lds = (size_t)hs_lds(cap, stage_hs).bytes;
    hipLaunchKernelGGL(k_cache_hs_many, dim3((unsigned)n_blocks), dim3(kHsBlock), lds, s, d_jobs,
                       d_block_prefix, n_inst, cap, stage_hs);
    PUMP_LAUNCHED();
    return PUMP_OK;
}

extern "C" int pump_setup_valid_particles(uint8_t *d_valid, int n_p, int64_t max_path_count,
                                          float *d_path_cp, float *d_path_time,
                                          pump_stream_t stream)
{
    PUMP_REQUIRE(d_valid && d_path_cp && d_path_time && n_p > 0 && max_path_count >= 0);
    const int64_t nb = max_path_count * (int64_t)n_p;
    hipLaunchKernelGGL(k_setup_valid, dim3(grid_for(nb / 4 + 1, 256)), dim3(256), 0,
                       as_stream(stream), d_valid, nb, d_path_cp, d_path_time, max_path_count);
    PUMP_LAUNCHED();
    return PUMP_OK;
}

// Propagate kernel selection: the wave-per-path kernel (default) or the
// workgroup-per-path kernel with an LDS list (PUMP_HSMC_KERNEL=block), which
// suits edges that store many half-spaces.
static bool use_block_kernel()
{
    static int v = -1;
    if (v < 0) {
        const char *e = getenv("PUMP_HSMC_KERNEL");
        v = (e && e[0] == 'b') ? 1 : 0;
    }
    return v == 1;
}

static unsigned wave_grid(int64_t n_new)
{
    int dev = 0, cus = 256;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    int64_t blocks = ((int64_t)n_new + 3) / 4;
    const int64_t cap = (int64_t)cus * 32;
    return (unsigned)(blocks < cap ? (blocks > 0 ? blocks : 1) : cap);
}

static int check_prop_args(int num_new_paths, int waypoints_count, const int64_t *d_wp_offsets,
                           int max_hs, int tmax_steps, int n_p, int n_mc)
{
    if (num_new_paths < 0 || max_hs < 0 || n_p <= 0 || tmax_steps <= 0 || n_mc < n_p)
        return PUMP_E_INVALID;
    if (!d_wp_offsets && waypoints_count < 2)
        return PUMP_E_INVALID;
    if (max_hs > kListCap)
        return PUMP_E_RANGE;
    return PUMP_OK;
}

// workgroups of a grid-stride launch with one wave per item, at most 32 per CU
static unsigned wave_blocks_for(int64_t waves)
{
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int64_t cap = (int64_t)cus * 32;
    const int64_t blocks = (waves + 3) / 4;
    return (unsigned)(blocks < 1 ? 1 : (blocks < cap ? blocks : cap));
}

// The 64-paths-per-wave propagate: its copy unit and grid
template <bool kBits>
static int launch_b64(int n_new, const int64_t *wp_offsets, int W_fixed, const int *wf_edge,
                      const int *wf_prev, const float *as, const float *bs,
                      const uint64_t *edge_mask, const float *times, const float *topts, float dt,
                      int max_hs, const float *offsets_time, int tmax_steps, float om, int n_mc,
                      const void *src, void *dst, int64_t dst_base, int n_p, float *path_cp,
                      int64_t cp_base, unsigned long long *clamped, hipStream_t s)
{
    const int64_t row_bytes = kBits ? (int64_t)((n_p + 63) / 64) * 8 : (int64_t)n_p;
    const uintptr_t al = (uintptr_t)src | (uintptr_t)dst;
    const int unit = (row_bytes % 16 == 0 && (al & 15) == 0) ? 16 : (kBits ? 8 : 1);
    hipLaunchKernelGGL(k_propagate_b64<kBits>, dim3(wave_blocks_for((n_new + 63) / 64)),
                       dim3(256), 0, s, n_new, wp_offsets, W_fixed, wf_edge, wf_prev, as, bs,
                       edge_mask, times, topts, dt, max_hs, offsets_time, tmax_steps, om, n_mc,
                       src, dst, dst_base, n_p, path_cp, cp_base, clamped, unit);
    PUMP_LAUNCHED();
    return PUMP_OK;
}

// defined below, with the batched form
template <bool kBits>
static int launch_propagate(int n_new, int W_fixed, const int64_t *wp_offsets, const int *wf_edge,
                            const int *wf_prev, const float *as, const float *bs,
                            const uint64_t *edge_mask, const float *times, const float *topts,
                            float dt, int max_hs, const float *offsets_time, int tmax_steps,
                            float om, int n_mc, const void *src, void *dst, int64_t dst_base,
                            int n_p, float *path_cp, int64_t cp_base, unsigned long long *clamped,
                            const uint16_t *split, unsigned long long *band, void *ws,
                            size_t ws_bytes, hipStream_t s, const float *parent_cp = nullptr,
                            int *row_src = nullptr);

extern "C" int pump_propagate_valid_particles(
    int num_new_paths, int waypoints_count, const int64_t *d_wp_offsets, const int *d_wf_edge,
    const int *d_wf_prev, int64_t path_base, const float *d_as, const float *d_bs,
    const int *d_count_stored, const float *d_times, const float *d_topts, float dt, int max_hs,
    const float *d_offsets_time, int tmax_steps, float offset_mult, int num_mc_particles,
    uint8_t *d_valid, int n_p, unsigned long long *d_clamped, pump_stream_t stream)
{
    (void)d_count_stored; // passed but unused, as in the reference (PUMP.cu:640)
    int r = check_prop_args(num_new_paths, waypoints_count, d_wp_offsets, max_hs, tmax_steps, n_p,
                            num_mc_particles);
    if (r)
        return r;
    if (num_new_paths == 0)
        return PUMP_OK;
    PUMP_REQUIRE(d_wf_edge && d_wf_prev && d_as && d_bs && d_times && d_topts && d_offsets_time &&
                 d_valid);
    hipStream_t s = as_stream(stream);
    if (num_new_paths >= 4096 && !use_block_kernel()) {
        // A large wavefront runs the batched form in place (classify, row
        // copies, time-binned half-space tests) with a workspace from the
        // stream-ordered pool, released on the same stream; the reference
        // allocates scratch at this call site on every iteration too
        // (PUMP.cu:89).  Without one, the per-path form below runs.
        const size_t ws_bytes = pump_propagate_workspace(num_new_paths, tmax_steps);
        void *ws = nullptr;
        if (hipMallocAsync(&ws, ws_bytes, s) == hipSuccess && ws) {
            const int r = launch_propagate<false>(
                num_new_paths, waypoints_count, d_wp_offsets, d_wf_edge, d_wf_prev, d_as, d_bs,
                (const uint64_t *)nullptr, d_times, d_topts, dt, max_hs, d_offsets_time,
                tmax_steps, offset_mult, num_mc_particles, (const void *)d_valid, (void *)d_valid,
                path_base, n_p, (float *)nullptr, (int64_t)0, d_clamped,
                (const uint16_t *)nullptr, (unsigned long long *)nullptr, ws, ws_bytes, s);
            const hipError_t fe = hipFreeAsync(ws, s);
            if (r)
                return r;
            PUMP_HIP(fe);
            return PUMP_OK;
        }
        (void)hipGetLastError(); // the failed request is not an error of this call
    }
    if (use_block_kernel())
        hipLaunchKernelGGL(k_propagate, dim3(num_new_paths), dim3(256), 0, as_stream(stream),
                           num_new_paths, d_wp_offsets, waypoints_count, d_wf_edge, d_wf_prev,
                           d_as, d_bs, d_times, d_topts, dt, max_hs, d_offsets_time, tmax_steps,
                           offset_mult, num_mc_particles, d_valid, d_valid, path_base, n_p,
                           (float *)nullptr, (int64_t)0, d_clamped);
    else
        return launch_b64<false>(num_new_paths, d_wp_offsets, waypoints_count, d_wf_edge,
                                 d_wf_prev, d_as, d_bs, (const uint64_t *)nullptr, d_times,
                                 d_topts, dt, max_hs, d_offsets_time, tmax_steps, offset_mult,
                                 num_mc_particles, (const void *)d_valid, (void *)d_valid,
                                 path_base, n_p, (float *)nullptr, (int64_t)0, d_clamped,
                                 as_stream(stream));
    PUMP_LAUNCHED();
    return PUMP_OK;
}

extern "C" int pump_calculate_cp_hsmc(int num_new_paths, int64_t path_base, float *d_path_cp,
                                      const uint8_t *d_valid, int n_p, pump_stream_t stream)
{
    PUMP_REQUIRE(num_new_paths >= 0 && n_p > 0);
    if (num_new_paths == 0)
        return PUMP_OK;
    PUMP_REQUIRE(d_path_cp && d_valid);
    hipLaunchKernelGGL(k_cp, dim3(wave_grid(num_new_paths)), dim3(256), 0, as_stream(stream),
                       num_new_paths, path_base, d_path_cp, d_valid, n_p);
    PUMP_LAUNCHED();
    return PUMP_OK;
}

static size_t round256(size_t x)
{
    return (x + 255) & ~(size_t)255;
}

// Workspace of the batched form, in order: control words, the sub-binned
// per-time histogram and the sub-bin cursors (cleared together), time-index
// starts, per-path flags, listed-path records, CP accumulators of the group
// form, chunks and binned entries.  The entry capacity (one entry per
// (path, waypoint) pair that holds a half-space) is 4 per path + 65536 in
// the workspace of pump_propagate_workspace, which covers typical
// wavefronts; a larger workspace holds more (the entries fill it), and
// beyond the capacity the group form runs instead of the binned one.
constexpr int64_t kMaxEntries = int64_t(1) << 30; // entry positions are ints
struct BatchLayout {
    size_t hist, cursor, tstart, has, recs, acc, chunks, entries, total;
    int64_t entry_cap, chunk_cap;
    BatchLayout(int64_t n, int64_t T, int64_t cap = 0)
    {
        n = n > 0 ? n : 1;
        T = T > 0 ? T : 1;
        set(n, T, cap > 4 * n + 65536 ? cap : 4 * n + 65536);
    }
    // the layout whose entries fill ws_bytes (the default one if smaller)
    static BatchLayout fill(int64_t n, int64_t T, size_t ws_bytes)
    {
        BatchLayout L(n, T);
        if (ws_bytes <= L.total)
            return L;
        // the largest capacity that fits (total grows with it)
        int64_t lo = L.entry_cap, hi = kMaxEntries;
        if (BatchLayout(n, T, hi).total <= ws_bytes)
            return BatchLayout(n, T, hi);
        while (hi - lo > 1) {
            const int64_t mid = lo + (hi - lo) / 2;
            if (BatchLayout(n, T, mid).total <= ws_bytes)
                lo = mid;
            else
                hi = mid;
        }
        return BatchLayout(n, T, lo);
    }
    void set(int64_t n, int64_t T, int64_t cap)
    {
        entry_cap = cap < kMaxEntries ? cap : kMaxEntries;
        chunk_cap = entry_cap / kBinChunk + T + 1;
        const size_t bins = (size_t)T * kSubBins;
        hist = 256;
        cursor = hist + round256(bins * 4);
        tstart = cursor + round256(bins * 4);
        has = tstart + round256((size_t)(T + 1) * 4);
        recs = has + round256((size_t)n);
        acc = recs + round256((size_t)n * sizeof(HsPath));
        chunks = acc + round256((size_t)n * 8);
        entries = chunks + round256((size_t)chunk_cap * sizeof(BinChunk));
        total = entries + (size_t)entry_cap * sizeof(BinEntry);
    }
};

extern "C" size_t pump_propagate_workspace(int num_new_paths, int tmax_steps)
{
    return BatchLayout(num_new_paths, tmax_steps).total;
}

extern "C" size_t pump_propagate_workspace_pairs(int num_new_paths, int tmax_steps,
                                                 int64_t marked_pairs)
{
    return BatchLayout(num_new_paths, tmax_steps,
                       marked_pairs < kMaxEntries ? marked_pairs : kMaxEntries)
        .total;
}

__global__ __launch_bounds__(256) void k_zero_words(int *__restrict__ p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// Particle groups of 64 per wave of the batched form's half-space test,
// chosen by pump_hsmc_autotune_cow per configuration (validity form,
// precision, particle count, half-space slots) and kept for the process;
// 4 (256 particles) where nothing was tuned.
using PpwKey = std::array<int, 4>;
static std::mutex g_ppw_mu;
static std::map<PpwKey, int> g_ppw;
static thread_local int t_ppw_try = 0; // candidate being timed by the autotuner

static int ppw_groups(bool bits, bool split, int n_p, int max_hs)
{
    if (t_ppw_try)
        return t_ppw_try;
    std::lock_guard<std::mutex> lk(g_ppw_mu);
    const auto it = g_ppw.find(PpwKey{bits ? 1 : 0, split ? 1 : 0, n_p, max_hs});
    return it == g_ppw.end() ? 4 : it->second;
}

// Launches the propagate + CP step for a wavefront: the batched form when a
// workspace is given and the wavefront is large, else one wave-per-path (or,
// on request, workgroup-per-path) launch.
template <bool kBits>
static int launch_propagate(int n_new, int W_fixed, const int64_t *wp_offsets, const int *wf_edge,
                            const int *wf_prev, const float *as, const float *bs,
                            const uint64_t *edge_mask, const float *times, const float *topts,
                            float dt, int max_hs, const float *offsets_time, int tmax_steps,
                            float om, int n_mc, const void *src, void *dst, int64_t dst_base,
                            int n_p, float *path_cp, int64_t cp_base, unsigned long long *clamped,
                            const uint16_t *split, unsigned long long *band, void *ws,
                            size_t ws_bytes, hipStream_t s, const float *parent_cp,
                            int *row_src)
{
    const int64_t row_bytes = kBits ? (int64_t)((n_p + 63) / 64) * 8 : (int64_t)n_p;
    const BatchLayout L = BatchLayout::fill(n_new, tmax_steps, ws ? ws_bytes : 0);
    if (ws && ws_bytes >= L.total && n_new >= 4096 && !use_block_kernel()) {
        uint8_t *base = reinterpret_cast<uint8_t *>(ws);
        int *ctl = reinterpret_cast<int *>(base);
        int *hist = reinterpret_cast<int *>(base + L.hist);
        uint8_t *has = base + L.has;
        HsPath *recs = reinterpret_cast<HsPath *>(base + L.recs);
        unsigned long long *acc = reinterpret_cast<unsigned long long *>(base + L.acc);
        BinArgs B;
        B.n_new = n_new;
        B.n_p = n_p;
        B.n_mc = n_mc;
        B.tmax_steps = tmax_steps;
        B.max_hs = max_hs;
        B.known = edge_mask ? 1 : 0;
        B.dt = dt;
        B.om = om;
        B.as = as;
        B.bs = bs;
        B.offsets_time = offsets_time;
        B.recs = recs;
        B.acc = acc;
        B.ctl = ctl;
        B.hist = hist;
        B.cursor = reinterpret_cast<int *>(base + L.cursor);
        B.tstart = reinterpret_cast<int *>(base + L.tstart);
        B.chunks = reinterpret_cast<BinChunk *>(base + L.chunks);
        B.entries = reinterpret_cast<BinEntry *>(base + L.entries);
        B.entry_cap = L.entry_cap;
        B.chunk_cap = L.chunk_cap;
        B.valid_src = src;
        B.valid_dst = dst;
        B.dst_base = dst_base;
        B.path_cp = path_cp;
        B.cp_base = cp_base;
        B.split = split;
        B.band = band;
        B.nq = ppw_groups(kBits, split != nullptr, n_p, max_hs);
        // control words, histogram and cursors, cleared by a kernel (safe
        // under graph capture)
        const int64_t n_zero = (int64_t)(L.tstart / 4);
        hipLaunchKernelGGL(k_zero_words, dim3(grid_for(n_zero, 256, 1024)), dim3(256), 0, s, ctl,
                           n_zero);
        PUMP_LAUNCHED();
        hipLaunchKernelGGL(k_classify, dim3(grid_for(n_new, kClassifyBlock, 2048)),
                           dim3(kClassifyBlock), 0, s, n_new, wp_offsets, W_fixed, wf_edge,
                           wf_prev, bs, edge_mask, max_hs, times, topts, dt, tmax_steps, has,
                           recs, ctl, hist, acc, clamped, row_src, parent_cp, path_cp, cp_base);
        PUMP_LAUNCHED();
        int r = bins_scan(B, s);
        if (r)
            return r;
        if (row_src) {
            // copy-on-write: only the listed paths' rows are written
            const uintptr_t al = ((uintptr_t)src | (uintptr_t)dst);
            const unsigned g = wave_blocks_for(n_new);
            if (row_bytes % 16 == 0 && (al & 15) == 0)
                hipLaunchKernelGGL(k_copy_listed<16>, dim3(g), dim3(256), 0, s, recs, ctl,
                                   (const uint8_t *)src, (uint8_t *)dst, dst_base, row_bytes);
            else if (row_bytes % 8 == 0 && (al & 7) == 0)
                hipLaunchKernelGGL(k_copy_listed<8>, dim3(g), dim3(256), 0, s, recs, ctl,
                                   (const uint8_t *)src, (uint8_t *)dst, dst_base, row_bytes);
            else
                hipLaunchKernelGGL(k_copy_listed<1>, dim3(g), dim3(256), 0, s, recs, ctl,
                                   (const uint8_t *)src, (uint8_t *)dst, dst_base, row_bytes);
            PUMP_LAUNCHED();
            return bins_finish(B, kBits, s);
        }
        // copy unit and segment width
        const uintptr_t al = ((uintptr_t)src | (uintptr_t)dst);
        int unit = 1;
        if (row_bytes % 16 == 0 && (al & 15) == 0)
            unit = 16;
        else if (kBits)
            unit = 8;
        else if (row_bytes % 4 == 0 && (al & 3) == 0)
            unit = 4;
        const int units = (int)(row_bytes / unit);
        int seg = 1;
        while (seg < units && seg < 64)
            seg <<= 1;
        const int64_t waves = ((int64_t)n_new * seg + 64 * kCopyRows - 1) / (64 * kCopyRows);
        const unsigned grid = grid_for(waves * 64, 256, 65536);
        const uint8_t *sb = (const uint8_t *)src;
        uint8_t *db = (uint8_t *)dst;
        const int *cc = ctl;
        if (kBits && unit == 16)
            hipLaunchKernelGGL((k_copy_rows<16, true>), dim3(grid), dim3(256), 0, s, n_new,
                               wf_prev, has, cc, sb, db, dst_base, row_bytes, n_p, seg, path_cp,
                               cp_base);
        else if (kBits)
            hipLaunchKernelGGL((k_copy_rows<8, true>), dim3(grid), dim3(256), 0, s, n_new,
                               wf_prev, has, cc, sb, db, dst_base, row_bytes, n_p, seg, path_cp,
                               cp_base);
        else if (unit == 16)
            hipLaunchKernelGGL((k_copy_rows<16, false>), dim3(grid), dim3(256), 0, s, n_new,
                               wf_prev, has, cc, sb, db, dst_base, row_bytes, n_p, seg, path_cp,
                               cp_base);
        else if (unit == 4)
            hipLaunchKernelGGL((k_copy_rows<4, false>), dim3(grid), dim3(256), 0, s, n_new,
                               wf_prev, has, cc, sb, db, dst_base, row_bytes, n_p, seg, path_cp,
                               cp_base);
        else
            hipLaunchKernelGGL((k_copy_rows<1, false>), dim3(grid), dim3(256), 0, s, n_new,
                               wf_prev, has, cc, sb, db, dst_base, row_bytes, n_p, seg, path_cp,
                               cp_base);
        PUMP_LAUNCHED();
        return bins_finish(B, kBits, s);
    }
    if (use_block_kernel()) {
        if (kBits)
            hipLaunchKernelGGL(k_propagate_bits, dim3(n_new), dim3(256), 0, s, n_new, wp_offsets,
                               W_fixed, wf_edge, wf_prev, as, bs, times, topts, dt, max_hs,
                               offsets_time, tmax_steps, om, n_mc, (const uint64_t *)src,
                               (uint64_t *)dst, dst_base, n_p, path_cp, cp_base, clamped);
        else
            hipLaunchKernelGGL(k_propagate, dim3(n_new), dim3(256), 0, s, n_new, wp_offsets,
                               W_fixed, wf_edge, wf_prev, as, bs, times, topts, dt, max_hs,
                               offsets_time, tmax_steps, om, n_mc, (const uint8_t *)src,
                               (uint8_t *)dst, dst_base, n_p, path_cp, cp_base, clamped);
        PUMP_LAUNCHED();
    } else {
        int r = launch_b64<kBits>(n_new, wp_offsets, W_fixed, wf_edge, wf_prev, as, bs, edge_mask,
                                  times, topts, dt, max_hs, offsets_time, tmax_steps, om, n_mc,
                                  src, dst, dst_base, n_p, path_cp, cp_base, clamped, s);
        if (r)
            return r;
    }
    if (row_src) { // every child row was written
        hipLaunchKernelGGL(k_row_src_iota, dim3(grid_for(n_new, 256, 4096)), dim3(256), 0, s, n_new,
                           row_src);
        PUMP_LAUNCHED();
    }
    return PUMP_OK;
}

// ---------------------------------------------------------------------------
// The copy-on-write step of many problems, one launch per stage (McJob).
// ---------------------------------------------------------------------------
int pump::mc_cow_job(McJob &J, int n_new, const int64_t *wp_offsets, const int *wf_edge,
                     const int *wf_prev, const float *as, const float *bs,
                     const uint64_t *edge_mask, const float *times, const float *topts, float dt,
                     int max_hs, const float *offsets_time, int tmax_steps, float om, int n_mc,
                     const void *src, const float *parent_cp, void *dst, int *row_src, int n_p,
                     float *path_cp, void *ws, size_t ws_bytes)
{
    if (n_new < 1 || n_p < 1 || !wp_offsets || !ws || !src || !dst || !row_src || !path_cp ||
        !parent_cp)
        return PUMP_E_INVALID;
    const BatchLayout L = BatchLayout::fill(n_new, tmax_steps, ws_bytes);
    const int64_t row_bytes = (int64_t)((n_p + 63) / 64) * 8;
    if (ws_bytes < L.total || row_bytes % 16 != 0 ||
        (((uintptr_t)src | (uintptr_t)dst) & 15) != 0)
        return PUMP_E_INVALID;
    uint8_t *base = reinterpret_cast<uint8_t *>(ws);
    J = McJob{};
    BinArgs &B = J.B;
    B.n_new = n_new;
    B.n_p = n_p;
    B.n_mc = n_mc;
    B.tmax_steps = tmax_steps;
    B.max_hs = max_hs;
    B.known = edge_mask ? 1 : 0;
    B.dt = dt;
    B.om = om;
    B.as = as;
    B.bs = bs;
    B.offsets_time = offsets_time;
    B.recs = reinterpret_cast<HsPath *>(base + L.recs);
    B.acc = reinterpret_cast<unsigned long long *>(base + L.acc);
    B.ctl = reinterpret_cast<int *>(base);
    B.hist = reinterpret_cast<int *>(base + L.hist);
    B.cursor = reinterpret_cast<int *>(base + L.cursor);
    B.tstart = reinterpret_cast<int *>(base + L.tstart);
    B.chunks = reinterpret_cast<BinChunk *>(base + L.chunks);
    B.entries = reinterpret_cast<BinEntry *>(base + L.entries);
    B.entry_cap = L.entry_cap;
    B.chunk_cap = L.chunk_cap;
    B.valid_src = src;
    B.valid_dst = dst;
    B.dst_base = 0;
    B.path_cp = path_cp;
    B.cp_base = 0;
    B.split = nullptr;
    B.band = nullptr;
    B.nq = ppw_groups(true, false, n_p, max_hs);
    J.wp_offsets = wp_offsets;
    J.wf_edge = wf_edge;
    J.wf_prev = wf_prev;
    J.edge_mask = edge_mask;
    J.times = times;
    J.topts = topts;
    J.has = base + L.has;
    J.clamped = nullptr;
    J.row_src = row_src;
    J.parent_cp = parent_cp;
    J.row_bytes = row_bytes;
    J.n_zero = (int64_t)(L.tstart / 4);
    return PUMP_OK;
}

int pump::mc_plan(McJob *jobs, int n_jobs, McGrids &g, void **d_maps, hipStream_t s)
{
    *d_maps = nullptr;
    g = McGrids{};
    if (!jobs || n_jobs < 1 || n_jobs > 65536)
        return PUMP_E_INVALID;
    int64_t total_new = 0;
    for (int k = 0; k < n_jobs; ++k)
        total_new += jobs[k].B.n_new;
    int dev = 0, cus = 256;
    PUMP_HIP(hipGetDevice(&dev));
    PUMP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    // grid-stride stages share a grid in proportion to each problem's
    // wavefront: one full grid (32 workgroups per CU) for the binned tests,
    // four for the per-listed-path stages (emission, row copies, CP), whose
    // work items are short dependent load chains
    const int64_t budget = (int64_t)cus * 32;
    const auto share_of = [&](int64_t single, int64_t n, int64_t grids) {
        const int64_t part = (budget * grids * n + total_new - 1) / total_new;
        return std::max<int64_t>(1, std::min(single, part));
    };
    const auto share = [&](int64_t single, int64_t n) { return share_of(single, n, 1); };
    std::vector<uint32_t> maps[kMcStages];
    for (int k = 0; k < n_jobs; ++k) {
        McJob &J = jobs[k];
        const int64_t n = J.B.n_new;
        const int groups = (J.B.n_p + 63) / 64;
        const int nq = (J.B.nq == 1 || J.B.nq == 2 || J.B.nq == 8) ? J.B.nq : 4;
        const int hs = nq == 1 ? kMcHs1 : nq == 2 ? kMcHs2 : nq == 8 ? kMcHs8 : kMcHs4;
        int64_t nb[kMcStages] = {};
        nb[kMcZero] = std::max<int64_t>(1, std::min<int64_t>(32, (J.n_zero + 2047) / 2048));
        nb[kMcClassify] = grid_for(n, kClassifyBlock, 2048);
        nb[kMcScan] = 1;
        nb[kMcCopy] = share_of(wave_blocks_for(n), n, 4);
        nb[kMcEmit] = share_of(grid_for(n, 256, 4096), n, 4);
        nb[hs] = share(wave_blocks_for(J.B.chunk_cap * ((groups + nq - 1) / nq)), n);
        nb[kMcGroups] = share(
            std::min<int64_t>(wave_blocks_for(n * ((groups + kHsGroups - 1) / kHsGroups)), 1024),
            n);
        nb[kMcCp] = share_of(wave_blocks_for(n), n, 4);
        for (int st = 0; st < kMcStages; ++st) {
            if (nb[st] > 65536)
                return PUMP_E_INVALID;
            J.nb[st] = (int)nb[st];
            for (int64_t b = 0; b < nb[st]; ++b)
                maps[st].push_back(((uint32_t)k << 16) | (uint32_t)b);
        }
    }
    size_t total = 0;
    for (int st = 0; st < kMcStages; ++st)
        total += maps[st].size();
    uint32_t *d = nullptr;
    PUMP_HIP(hipMalloc((void **)&d, std::max<size_t>(total, 1) * sizeof(uint32_t)));
    size_t off = 0;
    for (int st = 0; st < kMcStages; ++st) {
        g.block_map[st] = d + off;
        g.blocks[st] = (int64_t)maps[st].size();
        if (!maps[st].empty()) {
            const hipError_t e = hipMemcpyAsync(d + off, maps[st].data(),
                                                maps[st].size() * sizeof(uint32_t),
                                                hipMemcpyHostToDevice, s);
            if (e != hipSuccess) {
                hipFree(d);
                PUMP_HIP(e);
            }
        }
        off += maps[st].size();
    }
    const hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        hipFree(d);
        PUMP_HIP(e);
    }
    *d_maps = d;
    return PUMP_OK;
}

int pump::mc_cow_many(const McJob *d_jobs, const McGrids &g, hipStream_t s)
{
    if (g.blocks[kMcZero] > 0) {
        hipLaunchKernelGGL(k_zero_many, dim3((unsigned)g.blocks[kMcZero]), dim3(256), 0, s,
                           d_jobs, g.block_map[kMcZero]);
        PUMP_LAUNCHED();
    }
    if (g.blocks[kMcClassify] > 0) {
        hipLaunchKernelGGL(k_classify_many, dim3((unsigned)g.blocks[kMcClassify]),
                           dim3(kClassifyBlock), 0, s, d_jobs, g.block_map[kMcClassify]);
        PUMP_LAUNCHED();
    }
    int r = bins_many(d_jobs, g, false, s);
    if (r)
        return r;
    if (g.blocks[kMcCopy] > 0) {
        hipLaunchKernelGGL(k_copy_listed_many, dim3((unsigned)g.blocks[kMcCopy]), dim3(256), 0,
                           s, d_jobs, g.block_map[kMcCopy]);
        PUMP_LAUNCHED();
    }
    return bins_many(d_jobs, g, true, s);
}

extern "C" int pump_propagate_cp(int num_new_paths, int waypoints_count,
                                 const int64_t *d_wp_offsets, const int *d_wf_edge,
                                 const int *d_wf_prev, const float *d_as, const float *d_bs,
                                 const uint64_t *d_edge_hs_mask, const float *d_times,
                                 const float *d_topts, float dt, int max_hs,
                                 const float *d_offsets_time, int tmax_steps, float offset_mult,
                                 int num_mc_particles, const uint8_t *d_valid_src,
                                 uint8_t *d_valid_dst, int64_t dst_base, int n_p,
                                 float *d_path_cp, int64_t cp_base, unsigned long long *d_clamped,
                                 void *d_workspace, size_t workspace_bytes, pump_stream_t stream)
{
    int r = check_prop_args(num_new_paths, waypoints_count, d_wp_offsets, max_hs, tmax_steps, n_p,
                            num_mc_particles);
    if (r)
        return r;
    if (num_new_paths == 0)
        return PUMP_OK;
    PUMP_REQUIRE(d_wf_edge && d_wf_prev && d_as && d_bs && d_times && d_topts && d_offsets_time &&
                 d_valid_src && d_valid_dst);
    return launch_propagate<false>(num_new_paths, waypoints_count, d_wp_offsets, d_wf_edge,
                                   d_wf_prev, d_as, d_bs, d_edge_hs_mask, d_times, d_topts, dt,
                                   max_hs, d_offsets_time, tmax_steps, offset_mult,
                                   num_mc_particles, d_valid_src, d_valid_dst, dst_base, n_p,
                                   d_path_cp, cp_base, d_clamped, (const uint16_t *)nullptr,
                                   (unsigned long long *)nullptr, d_workspace, workspace_bytes,
                                   as_stream(stream));
}

extern "C" int pump_gather_rows(const uint8_t *d_src, const int *d_src_rows, int n_rows,
                                uint8_t *d_dst, int64_t dst_base, int64_t row_bytes,
                                pump_stream_t stream)
{
    PUMP_REQUIRE(n_rows >= 0 && row_bytes > 0);
    if (n_rows == 0)
        return PUMP_OK;
    PUMP_REQUIRE(d_src && d_src_rows && d_dst && d_src != d_dst);
    hipLaunchKernelGGL(k_gather_rows, dim3(n_rows), dim3(256), 0, as_stream(stream), d_src,
                       d_src_rows, n_rows, d_dst, dst_base, row_bytes);
    PUMP_LAUNCHED();
    return PUMP_OK;
}

extern "C" int pump_propagate_cp_bits(int num_new_paths, int waypoints_count,
                                      const int64_t *d_wp_offsets, const int *d_wf_edge,
                                      const int *d_wf_prev, const float *d_as, const float *d_bs,
                                      const uint64_t *d_edge_hs_mask, const float *d_times,
                                      const float *d_topts, float dt, int max_hs,
                                      const float *d_offsets_time, int tmax_steps,
                                      float offset_mult, int num_mc_particles,
                                      const uint64_t *d_valid_src, uint64_t *d_valid_dst,
                                      int64_t dst_base, int n_p, float *d_path_cp,
                                      int64_t cp_base, int precision,
                                      const void *d_offsets_split,
                                      unsigned long long *d_band_count,
                                      unsigned long long *d_clamped, void *d_workspace,
                                      size_t workspace_bytes, pump_stream_t stream)
{
    int r = check_prop_args(num_new_paths, waypoints_count, d_wp_offsets, max_hs, tmax_steps, n_p,
                            num_mc_particles);
    if (r)
        return r;
    if (num_new_paths == 0)
        return PUMP_OK;
    PUMP_REQUIRE(d_wf_edge && d_wf_prev && d_as && d_bs && d_times && d_topts && d_offsets_time &&
                 d_valid_src && d_valid_dst && d_valid_src != d_valid_dst);
    PUMP_REQUIRE(precision == 0 || (precision == 1 && d_offsets_split));
    // the split-BF16 test needs om > 0 for its sign rule; otherwise FP32 runs
    // (the decisions are the same either way)
    const uint16_t *split = (precision == 1 && offset_mult > 0.0f && offset_mult < INFINITY)
                                ? (const uint16_t *)d_offsets_split
                                : nullptr;
    return launch_propagate<true>(num_new_paths, waypoints_count, d_wp_offsets, d_wf_edge,
                                  d_wf_prev, d_as, d_bs, d_edge_hs_mask, d_times, d_topts, dt,
                                  max_hs, d_offsets_time, tmax_steps, offset_mult,
                                  num_mc_particles, d_valid_src, d_valid_dst, dst_base, n_p,
                                  d_path_cp, cp_base, d_clamped, split, d_band_count, d_workspace,
                                  workspace_bytes, as_stream(stream));
}

extern "C" int pump_propagate_cp_cow(int validity_bits, int num_new_paths, int waypoints_count,
                                     const int64_t *d_wp_offsets, const int *d_wf_edge,
                                     const int *d_wf_prev, const float *d_as, const float *d_bs,
                                     const uint64_t *d_edge_hs_mask, const float *d_times,
                                     const float *d_topts, float dt, int max_hs,
                                     const float *d_offsets_time, int tmax_steps,
                                     float offset_mult, int num_mc_particles,
                                     const void *d_valid_src, const float *d_parent_cp,
                                     void *d_valid_dst, int64_t dst_base, int *d_row_src, int n_p,
                                     float *d_path_cp, int64_t cp_base, int precision,
                                     const void *d_offsets_split,
                                     unsigned long long *d_band_count,
                                     unsigned long long *d_clamped, void *d_workspace,
                                     size_t workspace_bytes, pump_stream_t stream)
{
    int r = check_prop_args(num_new_paths, waypoints_count, d_wp_offsets, max_hs, tmax_steps, n_p,
                            num_mc_particles);
    if (r)
        return r;
    if (num_new_paths == 0)
        return PUMP_OK;
    PUMP_REQUIRE(validity_bits == 0 || validity_bits == 1);
    PUMP_REQUIRE(d_wf_edge && d_wf_prev && d_as && d_bs && d_times && d_topts && d_offsets_time &&
                 d_valid_src && d_valid_dst && d_valid_src != d_valid_dst && d_parent_cp &&
                 d_row_src && d_path_cp);
    PUMP_REQUIRE(precision == 0 || (precision == 1 && validity_bits && d_offsets_split));
    if (validity_bits) {
        const uint16_t *split = (precision == 1 && offset_mult > 0.0f && offset_mult < INFINITY)
                                    ? (const uint16_t *)d_offsets_split
                                    : nullptr;
        return launch_propagate<true>(
            num_new_paths, waypoints_count, d_wp_offsets, d_wf_edge, d_wf_prev, d_as, d_bs,
            d_edge_hs_mask, d_times, d_topts, dt, max_hs, d_offsets_time, tmax_steps, offset_mult,
            num_mc_particles, d_valid_src, d_valid_dst, dst_base, n_p, d_path_cp, cp_base,
            d_clamped, split, d_band_count, d_workspace, workspace_bytes, as_stream(stream),
            d_parent_cp, d_row_src);
    }
    return launch_propagate<false>(
        num_new_paths, waypoints_count, d_wp_offsets, d_wf_edge, d_wf_prev, d_as, d_bs,
        d_edge_hs_mask, d_times, d_topts, dt, max_hs, d_offsets_time, tmax_steps, offset_mult,
        num_mc_particles, d_valid_src, d_valid_dst, dst_base, n_p, d_path_cp, cp_base, d_clamped,
        (const uint16_t *)nullptr, (unsigned long long *)nullptr, d_workspace, workspace_bytes,
        as_stream(stream), d_parent_cp, d_row_src);
}

extern "C" int pump_gather_rows_cow(const void *d_children, const void *d_parents,
                                    const int *d_row_src, const int *d_sel, int n_sel,
                                    void *d_dst, int64_t dst_base, int64_t row_bytes,
                                    int validity_bits, pump_stream_t stream)
{
    PUMP_REQUIRE(n_sel >= 0 && row_bytes > 0 && dst_base >= 0);
    if (n_sel == 0)
        return PUMP_OK;
    PUMP_REQUIRE(d_children && d_parents && d_row_src && d_dst && d_children != d_dst);
    int seg = 1;
    const int64_t units = (row_bytes % 16 == 0) ? row_bytes / 16 : row_bytes;
    while (seg < units && seg < 256)
        seg <<= 1;
    const int per_block = 256 / seg;
    hipLaunchKernelGGL(k_gather_cow, dim3(grid_for(n_sel, per_block, 65536)), dim3(256), 0,
                       as_stream(stream), (const uint8_t *)d_children, (const uint8_t *)d_parents,
                       d_row_src, d_sel, n_sel, (uint8_t *)d_dst, dst_base, row_bytes,
                       validity_bits ? 1 : 0, seg);
    PUMP_LAUNCHED();
    return PUMP_OK;
}

// initial bit masks: all n_p particles valid
extern "C" int pump_setup_valid_bits(uint64_t *d_valid, int n_p, int64_t n_paths,
                                     pump_stream_t stream)
{
    PUMP_REQUIRE(d_valid && n_p > 0 && n_paths >= 0);
    const int64_t total = n_paths * ((n_p + 63) / 64);
    hipLaunchKernelGGL(k_setup_bits, dim3(grid_for(total, 256)), dim3(256), 0, as_stream(stream),
                       d_valid, n_paths, n_p);
    PUMP_LAUNCHED();
    return PUMP_OK;
}

extern "C" int pump_hsmc_autotune_cow(
    int validity_bits, int num_new_paths, int waypoints_count, const int64_t *d_wp_offsets,
    const int *d_wf_edge, const int *d_wf_prev, const float *d_as, const float *d_bs,
    const uint64_t *d_edge_hs_mask, const float *d_times, const float *d_topts, float dt,
    int max_hs, const float *d_offsets_time, int tmax_steps, float offset_mult,
    int num_mc_particles, const void *d_valid_src, const float *d_parent_cp, void *d_valid_dst,
    int64_t dst_base, int *d_row_src, int n_p, float *d_path_cp, int64_t cp_base, int precision,
    const void *d_offsets_split, unsigned long long *d_band_count, unsigned long long *d_clamped,
    void *d_workspace, size_t workspace_bytes, pump_stream_t stream, int *particles_per_wave)
{
    hipStream_t s = as_stream(stream);
    auto run = [&](unsigned long long *band, unsigned long long *clamped) {
        return pump_propagate_cp_cow(validity_bits, num_new_paths, waypoints_count, d_wp_offsets,
                                     d_wf_edge, d_wf_prev, d_as, d_bs, d_edge_hs_mask, d_times,
                                     d_topts, dt, max_hs, d_offsets_time, tmax_steps, offset_mult,
                                     num_mc_particles, d_valid_src, d_parent_cp, d_valid_dst,
                                     dst_base, d_row_src, n_p, d_path_cp, cp_base, precision,
                                     d_offsets_split, band, clamped, d_workspace, workspace_bytes,
                                     stream);
    };
    static const int kCand[4] = {1, 2, 4, 8};
    int best = 4;
    float best_ms = 0;
    int r = PUMP_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t ce = hipEventCreate(&e0);
    if (ce == hipSuccess) {
        ce = hipEventCreate(&e1);
        if (ce != hipSuccess)
            hipEventDestroy(e0);
    }
    if (ce != hipSuccess) {
        note_hip_error(ce, __FILE__, __LINE__);
        return PUMP_E_HIP;
    }
    for (int c : kCand) {
        t_ppw_try = c;
        r = run(nullptr, nullptr); // warm-up
        if (r != PUMP_OK)
            break;
        hipEventRecord(e0, s);
        for (int rep = 0; rep < 3 && r == PUMP_OK; ++rep)
            r = run(nullptr, nullptr);
        hipEventRecord(e1, s);
        float ms = 0;
        if (r == PUMP_OK && (hipEventSynchronize(e1) != hipSuccess ||
                             hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            r = PUMP_E_HIP;
        if (r != PUMP_OK)
            break;
        if (best_ms == 0 || ms < best_ms) {
            best_ms = ms;
            best = c;
        }
    }
    t_ppw_try = 0;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (r != PUMP_OK)
        return r;
    const bool split = validity_bits && precision == 1 && offset_mult > 0.0f &&
                       offset_mult < INFINITY;
    {
        std::lock_guard<std::mutex> lk(g_ppw_mu);
        g_ppw[PpwKey{validity_bits ? 1 : 0, split ? 1 : 0, n_p, max_hs}] = best;
    }
    if (particles_per_wave)
        *particles_per_wave = 64 * best;
    // the outputs, and the counters, from the chosen shape
    return run(d_band_count, d_clamped);
}

extern "C" int pump_hsmc_particles_per_wave(int validity_bits, int precision, int n_p, int max_hs)
{
    return 64 * ppw_groups(validity_bits != 0, validity_bits && precision == 1, n_p, max_hs);
}

extern "C" int pump_hsmc_set_particles_per_wave(int validity_bits, int precision, int n_p,
                                                int max_hs, int particles_per_wave)
{
    if (particles_per_wave != 0 && particles_per_wave != 64 && particles_per_wave != 128 &&
        particles_per_wave != 256 && particles_per_wave != 512)
        return PUMP_E_INVALID;
    const PpwKey key{validity_bits ? 1 : 0, (validity_bits && precision == 1) ? 1 : 0, n_p, max_hs};
    std::lock_guard<std::mutex> lk(g_ppw_mu);
    if (particles_per_wave == 0)
        g_ppw.erase(key);
    else
        g_ppw[key] = particles_per_wave / 64;
    return PUMP_OK;
}
