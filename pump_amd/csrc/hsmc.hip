// hsmc.hip -- half-space cache and the half-space Monte Carlo step
// (reference: setupAsBs / cacheHalfspaces hsmc.cu:10-99,201-211,
// bvlsShortcut bvls.cu:259-273, setupValidParticles / propagateValidParticles /
// calculateCP_HSMC PUMP.cu:613-777).
//
// cacheHalfspaces: a workgroup owns a run of edges and walks their waypoints
// one per lane; obstacles are staged in LDS (any count: boxes past the LDS
// capacity are read from global memory).
//
// propagate: one workgroup per new path.  The (waypoint, half-space) pairs
// that actually hold a half-space are gathered once into LDS together with
// the waypoint's time index; the particles then stream over that short list.
// The half-space data is uniform across the workgroup, the offsets row of a
// time index is read coalesced (consecutive particles are consecutive
// 12-byte records), and the child validity is written coalesced.  Particle
// decisions use the reference's float operation order
//   LHS = ((0 + o0*a0) + o1*a1) + o2*a2;  valid &= LHS*offsetMult < b
// with no contraction, so they are bit-identical.  The AND over
// (waypoint, half-space) pairs is order-independent.
#include <cstdlib>

#include "pump_device.hpp"
#include "pump_internal.hpp"

#pragma clang fp contract(off)

namespace pump {
namespace {

constexpr int kEdgesPerBlock = 32;
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

template <bool kSentinels>
__global__ __launch_bounds__(kHsBlock) void k_cache_hs(
    int64_t n_edges, const float *__restrict__ disc, const int64_t *__restrict__ wp_offsets,
    int w_fixed, const uint8_t *__restrict__ is_free, int n_obs, int max_hs,
    const float *__restrict__ obstacles, float *__restrict__ as, float *__restrict__ bs,
    int *__restrict__ count_stored, float maxd2)
{
    __shared__ ObsLds o;
    __shared__ int64_t off[kEdgesPerBlock + 1];
    __shared__ uint8_t fr[kEdgesPerBlock];
    const int64_t e0 = (int64_t)blockIdx.x * kEdgesPerBlock;
    const int ne = (int)((n_edges - e0) < kEdgesPerBlock ? (n_edges - e0) : kEdgesPerBlock);
    const int t = threadIdx.x;
    stage_obstacles(o, obstacles, n_obs);
    for (int q = t; q <= ne; q += kHsBlock)
        off[q] = wp_offsets ? wp_offsets[e0 + q] : (e0 + q) * (int64_t)w_fixed;
    for (int q = t; q < ne; q += kHsBlock)
        fr[q] = is_free[e0 + q];
    __syncthreads();
    const int64_t g_begin = off[0], g_end = off[ne];
    for (int64_t g = g_begin + t; g < g_end; g += kHsBlock) {
        const int le = find_edge(off, ne, g);
        int64_t hs = g * max_hs;
        int stored = 0;
        if (fr[le]) {
            float pos[3], vel[3];
            for (int d = 0; d < 3; ++d) {
                pos[d] = disc[g * kDim + d];
                vel[d] = disc[g * kDim + 3 + d];
            }
            for (int b = 0; b < n_obs && stored < max_hs; ++b) {
                float lo[3], hi[3], x[3];
                box_bounds(o, obstacles, b, lo, hi);
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
                for (int d = 0; d < 3; ++d) {
                    as[hs * 3 + d] = x[d];
                    xx = xx + x[d] * x[d];
                }
                bs[hs] = xx;
                ++stored;
                ++hs;
            }
        }
        count_stored[g] = stored;
        if (kSentinels) {
            for (int h = stored; h < max_hs; ++h, ++hs) {
                as[hs * 3 + 0] = -1.0f;
                as[hs * 3 + 1] = -1.0f;
                as[hs * 3 + 2] = -1.0f;
                bs[hs] = 1111.0f;
            }
        }
    }
}

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
    float lhs = 0.0f;
    lhs = lhs + o[0] * en.a0;
    lhs = lhs + o[1] * en.a1;
    lhs = lhs + o[2] * en.a2;
    return lhs * om < en.b;
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
__global__ __launch_bounds__(256) void k_cp(int n_new, int64_t base, float *__restrict__ cp,
                                            const uint8_t *__restrict__ valid, int n_p)
{
    __shared__ int red[4];
    const int w = blockIdx.x;
    const uint8_t *row = valid + (base + w) * (int64_t)n_p;
    int c = 0;
    for (int p = threadIdx.x; p < n_p; p += blockDim.x)
        c += row[p] ? 0 : 1;
    for (int off = 32; off > 0; off >>= 1)
        c += __shfl_down(c, off, 64);
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0)
        cp[base + w] = ((float)(red[0] + red[1] + red[2] + red[3])) / (float)n_p;
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
// Wave-per-path form of propagate + CP.  One wave
// owns one new path: a first pass ballots, per chunk of 64 waypoints, which
// waypoints hold at least one stored half-space (b < 1000); the particles are
// then streamed 64 at a time (lane = particle) and tested only against those
// waypoints' half-spaces, whose data are wave-uniform loads.  A path whose
// edge stores no half-space (the common case) is a straight copy of the
// parent row plus a count.  No LDS, no workgroup barriers.
struct PathInfo {
    int64_t g0;
    int W;
    float dtl, tbase;
};

__device__ __forceinline__ PathInfo path_info(int e, int prev, const int64_t *wp_offsets,
                                              int w_fixed, const float *topts,
                                              const float *times)
{
    PathInfo pi;
    pi.g0 = wp_offsets ? wp_offsets[e] : (int64_t)e * w_fixed;
    pi.W = (int)((wp_offsets ? wp_offsets[e + 1] : pi.g0 + w_fixed) - pi.g0);
    pi.dtl = topts[e] / (float)(pi.W - 1);
    pi.tbase = times[prev];
    return pi;
}

// masks[c] bit l: waypoint i0+c*64+l holds a stored half-space (i0..i0+255);
// counts clamped time indices of those waypoints
__device__ __forceinline__ uint64_t hs_masks(uint64_t *masks, const PathInfo &pi, int i0,
                                             int max_hs, const float *__restrict__ bs, float dt,
                                             int tmax_steps, unsigned long long *clamped)
{
    const int lane = threadIdx.x & 63;
    uint64_t any = 0;
    for (int c = 0; c < 4; ++c) {
        const int i = i0 + c * 64 + lane;
        bool has = false;
        if (i < pi.W) {
            const int64_t head = (pi.g0 + i) * (int64_t)max_hs;
            for (int h = 0; h < max_hs; ++h)
                has |= bs[head + h] < 1000.0f;
            if (has && clamped) {
                float time = pi.tbase + (float)i * pi.dtl;
                int ti = (int)(time / dt);
                if (ti < 0 || ti >= tmax_steps)
                    atomicAdd(clamped, 1ull);
            }
        }
        const uint64_t m = __ballot(has);
        masks[c] = m;
        any |= m;
    }
    return any;
}

// AND of one particle's tests over every stored half-space of the marked
// waypoints (wave-uniform control flow over the masks)
__device__ __forceinline__ bool particle_walk(bool v, int p, const uint64_t *masks, int i0,
                                              const PathInfo &pi, int max_hs,
                                              const float *__restrict__ as,
                                              const float *__restrict__ bs,
                                              const float *__restrict__ off, int n_mc,
                                              int tmax_steps, float dt, float om)
{
    for (int c = 0; c < 4; ++c) {
        uint64_t m = masks[c];
        while (m) {
            const int l = __builtin_ctzll(m);
            m &= m - 1;
            const int i = i0 + c * 64 + l;
            float time = pi.tbase + (float)i * pi.dtl;
            int ti = (int)(time / dt);
            ti = ti < 0 ? 0 : (ti >= tmax_steps ? tmax_steps - 1 : ti);
            const int64_t head = (pi.g0 + i) * (int64_t)max_hs;
            const float *o = off + ((int64_t)ti * n_mc + p) * 3;
            float o0 = 0.0f, o1 = 0.0f, o2 = 0.0f;
            if (v) {
                o0 = o[0];
                o1 = o[1];
                o2 = o[2];
            }
            for (int h = 0; h < max_hs; ++h) {
                const float b = bs[head + h];
                if (b >= 1000.0f)
                    continue;
                const float a0 = as[(head + h) * 3 + 0];
                const float a1 = as[(head + h) * 3 + 1];
                const float a2 = as[(head + h) * 3 + 2];
                float lhs = 0.0f;
                lhs = lhs + o0 * a0;
                lhs = lhs + o1 * a1;
                lhs = lhs + o2 * a2;
                v = v && (lhs * om < b);
            }
        }
    }
    return v;
}

__device__ __forceinline__ int wave_sum(int x)
{
    for (int off = 32; off > 0; off >>= 1)
        x += __shfl_down(x, off, 64);
    return x;
}

// 0x01 in every byte of x that is nonzero, 0x00 elsewhere
__device__ __forceinline__ uint32_t nonzero_bytes01(uint32_t x)
{
    return ((((x & 0x7f7f7f7fu) + 0x7f7f7f7fu) | x) & 0x80808080u) >> 7;
}

template <bool kBits>
__global__ __launch_bounds__(256) void k_propagate_wave(
    int n_new, const int64_t *__restrict__ wp_offsets, int w_fixed,
    const int *__restrict__ wf_edge, const int *__restrict__ wf_prev,
    const float *__restrict__ as, const float *__restrict__ bs, const float *__restrict__ times,
    const float *__restrict__ topts, float dt, int max_hs, const float *__restrict__ offsets_time,
    int tmax_steps, float om, int n_mc, const void *__restrict__ valid_src,
    void *__restrict__ valid_dst, int64_t dst_base, int n_p, float *__restrict__ path_cp,
    int64_t cp_base, unsigned long long *__restrict__ clamped)
{
    const int lane = threadIdx.x & 63;
    const int64_t n_waves = (int64_t)gridDim.x * 4;
    for (int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); w < n_new; w += n_waves) {
        const int e = wf_edge[w];
        const int prev = wf_prev[w];
        const PathInfo pi = path_info(e, prev, wp_offsets, w_fixed, topts, times);
        int count = 0; // valid particles (bits) or invalid particles (bytes), per lane
        // waypoints in groups of 256; later groups continue from the child row
        for (int i0 = 0; i0 < pi.W; i0 += 256) {
            uint64_t masks[4];
            const uint64_t any = hs_masks(masks, pi, i0, max_hs, bs, dt, tmax_steps, clamped);
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
                    for (int wd = 0; wd < nwords; ++wd) {
                        const uint64_t word = src[wd];
                        const int p = wd * 64 + lane;
                        bool v = (p < n_p) && ((word >> lane) & 1ull);
                        if (__ballot(v) != 0ull)
                            v = particle_walk(v, p, masks, i0, pi, max_hs, as, bs, offsets_time,
                                              n_mc, tmax_steps, dt, om);
                        const uint64_t out = __ballot(v);
                        if (lane == 0) {
                            dst[wd] = out;
                            if (last)
                                count += __popcll(out);
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
                if (!any && words) {
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
                    for (int p0 = 0; p0 < n_p; p0 += 64) {
                        const int p = p0 + lane;
                        bool v = (p < n_p) && src[p] != 0;
                        if (any && __ballot(v) != 0ull)
                            v = particle_walk(v, p, masks, i0, pi, max_hs, as, bs, offsets_time,
                                              n_mc, tmax_steps, dt, om);
                        if (p < n_p) {
                            dst[p] = v ? 1 : 0;
                            if (last)
                                count += v ? 0 : 1;
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
                                     pump_stream_t stream)
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
    if (write_sentinels)
        hipLaunchKernelGGL(k_cache_hs<true>, dim3((unsigned)blocks), dim3(kHsBlock), 0,
                           as_stream(stream), n_edges, d_disc_motions, d_wp_offsets,
                           waypoints_count, d_is_free_edges, n_obs, max_hs, d_obstacles, d_as,
                           d_bs, d_count_stored, maxd2);
    else
        hipLaunchKernelGGL(k_cache_hs<false>, dim3((unsigned)blocks), dim3(kHsBlock), 0,
                           as_stream(stream), n_edges, d_disc_motions, d_wp_offsets,
                           waypoints_count, d_is_free_edges, n_obs, max_hs, d_obstacles, d_as,
                           d_bs, d_count_stored, maxd2);
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

static unsigned wave_grid(int n_new)
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
    if (use_block_kernel())
        hipLaunchKernelGGL(k_propagate, dim3(num_new_paths), dim3(256), 0, as_stream(stream),
                           num_new_paths, d_wp_offsets, waypoints_count, d_wf_edge, d_wf_prev,
                           d_as, d_bs, d_times, d_topts, dt, max_hs, d_offsets_time, tmax_steps,
                           offset_mult, num_mc_particles, d_valid, d_valid, path_base, n_p,
                           (float *)nullptr, (int64_t)0, d_clamped);
    else
        hipLaunchKernelGGL(k_propagate_wave<false>, dim3(wave_grid(num_new_paths)), dim3(256), 0,
                           as_stream(stream), num_new_paths, d_wp_offsets, waypoints_count,
                           d_wf_edge, d_wf_prev, d_as, d_bs, d_times, d_topts, dt, max_hs,
                           d_offsets_time, tmax_steps, offset_mult, num_mc_particles,
                           (const void *)d_valid, (void *)d_valid, path_base, n_p,
                           (float *)nullptr, (int64_t)0, d_clamped);
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
    hipLaunchKernelGGL(k_cp, dim3(num_new_paths), dim3(256), 0, as_stream(stream), num_new_paths,
                       path_base, d_path_cp, d_valid, n_p);
    PUMP_LAUNCHED();
    return PUMP_OK;
}

extern "C" int pump_propagate_cp(int num_new_paths, int waypoints_count,
                                 const int64_t *d_wp_offsets, const int *d_wf_edge,
                                 const int *d_wf_prev, const float *d_as, const float *d_bs,
                                 const float *d_times, const float *d_topts, float dt, int max_hs,
                                 const float *d_offsets_time, int tmax_steps, float offset_mult,
                                 int num_mc_particles, const uint8_t *d_valid_src,
                                 uint8_t *d_valid_dst, int64_t dst_base, int n_p,
                                 float *d_path_cp, int64_t cp_base, unsigned long long *d_clamped,
                                 pump_stream_t stream)
{
    int r = check_prop_args(num_new_paths, waypoints_count, d_wp_offsets, max_hs, tmax_steps, n_p,
                            num_mc_particles);
    if (r)
        return r;
    if (num_new_paths == 0)
        return PUMP_OK;
    PUMP_REQUIRE(d_wf_edge && d_wf_prev && d_as && d_bs && d_times && d_topts && d_offsets_time &&
                 d_valid_src && d_valid_dst);
    if (use_block_kernel())
        hipLaunchKernelGGL(k_propagate, dim3(num_new_paths), dim3(256), 0, as_stream(stream),
                           num_new_paths, d_wp_offsets, waypoints_count, d_wf_edge, d_wf_prev,
                           d_as, d_bs, d_times, d_topts, dt, max_hs, d_offsets_time, tmax_steps,
                           offset_mult, num_mc_particles, d_valid_src, d_valid_dst, dst_base, n_p,
                           d_path_cp, cp_base, d_clamped);
    else
        hipLaunchKernelGGL(k_propagate_wave<false>, dim3(wave_grid(num_new_paths)), dim3(256), 0,
                           as_stream(stream), num_new_paths, d_wp_offsets, waypoints_count,
                           d_wf_edge, d_wf_prev, d_as, d_bs, d_times, d_topts, dt, max_hs,
                           d_offsets_time, tmax_steps, offset_mult, num_mc_particles,
                           (const void *)d_valid_src, (void *)d_valid_dst, dst_base, n_p,
                           d_path_cp, cp_base, d_clamped);
    PUMP_LAUNCHED();
    return PUMP_OK;
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

// split-BF16 MFMA form, implemented in hsmc_mfma.hip
namespace pump {
int propagate_bits_mfma(int n_new, const int64_t *wp_offsets, int w_fixed, const int *wf_edge,
                        const int *wf_prev, const float *as, const float *bs, const float *times,
                        const float *topts, float dt, int max_hs, const float *offsets_time,
                        int tmax_steps, float om, int n_mc, const uint64_t *valid_src,
                        uint64_t *valid_dst, int64_t dst_base, int n_p, float *path_cp,
                        int64_t cp_base, unsigned long long *band, unsigned long long *clamped,
                        hipStream_t s);
}

extern "C" int pump_propagate_cp_bits(int num_new_paths, int waypoints_count,
                                      const int64_t *d_wp_offsets, const int *d_wf_edge,
                                      const int *d_wf_prev, const float *d_as, const float *d_bs,
                                      const float *d_times, const float *d_topts, float dt,
                                      int max_hs, const float *d_offsets_time, int tmax_steps,
                                      float offset_mult, int num_mc_particles,
                                      const uint64_t *d_valid_src, uint64_t *d_valid_dst,
                                      int64_t dst_base, int n_p, float *d_path_cp,
                                      int64_t cp_base, int precision,
                                      unsigned long long *d_band_count,
                                      unsigned long long *d_clamped, pump_stream_t stream)
{
    int r = check_prop_args(num_new_paths, waypoints_count, d_wp_offsets, max_hs, tmax_steps, n_p,
                            num_mc_particles);
    if (r)
        return r;
    if (num_new_paths == 0)
        return PUMP_OK;
    PUMP_REQUIRE(d_wf_edge && d_wf_prev && d_as && d_bs && d_times && d_topts && d_offsets_time &&
                 d_valid_src && d_valid_dst && d_valid_src != d_valid_dst);
    if (precision == 1)
        return propagate_bits_mfma(num_new_paths, d_wp_offsets, waypoints_count, d_wf_edge,
                                   d_wf_prev, d_as, d_bs, d_times, d_topts, dt, max_hs,
                                   d_offsets_time, tmax_steps, offset_mult, num_mc_particles,
                                   d_valid_src, d_valid_dst, dst_base, n_p, d_path_cp, cp_base,
                                   d_band_count, d_clamped, as_stream(stream));
    PUMP_REQUIRE(precision == 0);
    if (use_block_kernel())
        hipLaunchKernelGGL(k_propagate_bits, dim3(num_new_paths), dim3(256), 0,
                           as_stream(stream), num_new_paths, d_wp_offsets, waypoints_count,
                           d_wf_edge, d_wf_prev, d_as, d_bs, d_times, d_topts, dt, max_hs,
                           d_offsets_time, tmax_steps, offset_mult, num_mc_particles, d_valid_src,
                           d_valid_dst, dst_base, n_p, d_path_cp, cp_base, d_clamped);
    else
        hipLaunchKernelGGL(k_propagate_wave<true>, dim3(wave_grid(num_new_paths)), dim3(256), 0,
                           as_stream(stream), num_new_paths, d_wp_offsets, waypoints_count,
                           d_wf_edge, d_wf_prev, d_as, d_bs, d_times, d_topts, dt, max_hs,
                           d_offsets_time, tmax_steps, offset_mult, num_mc_particles,
                           (const void *)d_valid_src, (void *)d_valid_dst, dst_base, n_p,
                           d_path_cp, cp_base, d_clamped);
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
