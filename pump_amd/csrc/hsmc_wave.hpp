// hsmc_wave.hpp -- wave-level building blocks of the half-space Monte Carlo
// step (PUMP.cu:613-705): per-path edge data, waypoint half-space masks and
// the particle test loop.  Used by the wave-per-path, batched and time-binned
// forms.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pump_internal.hpp"

#pragma clang fp contract(off)

namespace pump {
namespace wave {

// One particle's test against one stored half-space, PUMP.cu:645-651:
//   LHS = 0; LHS += o_d*a_d (d = 0, 1, 2); valid iff LHS*offsetMult < b.
// The leading 0 + o0*a0 is dropped: it differs from o0*a0 only in the sign
// of a zero LHS, and (-0)*om < b has the same truth value as (+0)*om < b.
__device__ __forceinline__ bool hs_pass(float o0, float o1, float o2, float a0, float a1,
                                        float a2, float om, float b)
{
    float lhs = o0 * a0;
    lhs = lhs + o1 * a1;
    lhs = lhs + o2 * a2;
    return lhs * om < b;
}

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

// time index of waypoint i: (int)((times[prev] + i*dtLocal)/dt), PUMP.cu:637-638
__device__ __forceinline__ int time_index(const PathInfo &pi, int i, float dt)
{
    float time = pi.tbase + (float)i * pi.dtl;
    return (int)(time / dt);
}

__device__ __forceinline__ bool time_clamped(const PathInfo &pi, int i, float dt, int tmax_steps)
{
    const int ti = time_index(pi, i, dt);
    return ti < 0 || ti >= tmax_steps;
}

// the library reads row ti of the offsets table clamped into [0, tmax_steps)
__device__ __forceinline__ int clamp_time(int ti, int tmax_steps)
{
    return ti < 0 ? 0 : (ti >= tmax_steps ? tmax_steps - 1 : ti);
}

// An edge's waypoint mask (bit k: waypoint k holds a stored half-space, bit
// 63: waypoint 63 or a later one does) is exact when the edge has at most 64
// waypoints or bit 63 is clear.
__device__ __forceinline__ bool mask_exact(uint64_t emask, int W)
{
    return W <= 64 || !(emask >> 63);
}

// the summary's bits that name waypoints of a W-waypoint edge
__device__ __forceinline__ uint64_t mask_bits(uint64_t emask, int W)
{
    return W >= 64 ? emask : emask & ((1ull << W) - 1ull);
}

// A stored half-space: the reference skips entries with b >= 1000
// (PUMP.cu:642), so "holds a half-space" means some b that is not >= 1000.
__device__ __forceinline__ bool waypoint_has_hs(const float *__restrict__ bs, int64_t head,
                                                int max_hs)
{
    bool has = false;
    for (int h = 0; h < max_hs; ++h)
        has |= !(bs[head + h] >= 1000.0f);
    return has;
}

// Calls f(i), in increasing i, for every waypoint i of a path's edge that
// holds a stored half-space (lane-level: one path per lane).  The summary
// names waypoints 0..62 exactly, and 63 too when exact; later waypoints of
// an inexact summary, or all of them without one, are found by scanning bs.
template <class F>
__device__ __forceinline__ void for_each_marked(const PathInfo &pi, bool known, uint64_t emask,
                                                const float *__restrict__ bs, int max_hs, F f)
{
    int scan_from = 0;
    if (known) {
        const uint64_t m0 = mask_bits(emask, pi.W);
        const bool exact = mask_exact(m0, pi.W);
        for (uint64_t m = exact ? m0 : m0 & ~(1ull << 63); m; m &= m - 1)
            f(__builtin_ctzll(m));
        scan_from = exact ? pi.W : 63;
    }
    for (int i = scan_from; i < pi.W; ++i)
        if (waypoint_has_hs(bs, (pi.g0 + i) * (int64_t)max_hs, max_hs))
            f(i);
}

// masks[c] bit l: waypoint i0+c*64+l holds a stored half-space (i0..i0+255),
// from the edge's waypoint mask when it is known and exact, else by ballots
// over bs.  Counts clamped time indices of marked waypoints into *clamped
// when non-null.
__device__ __forceinline__ uint64_t hs_masks(uint64_t *masks, const PathInfo &pi, int i0,
                                             bool known, uint64_t emask, int max_hs,
                                             const float *__restrict__ bs, float dt,
                                             int tmax_steps, unsigned long long *clamped)
{
    const int lane = threadIdx.x & 63;
    if (known && mask_exact(emask, pi.W)) {
        const uint64_t m = i0 == 0 ? mask_bits(emask, pi.W) : 0ull;
        masks[0] = m;
        masks[1] = masks[2] = masks[3] = 0;
        if (clamped && ((m >> lane) & 1ull) && time_clamped(pi, lane, dt, tmax_steps))
            atomicAdd(clamped, 1ull);
        return m;
    }
    uint64_t any = 0;
    for (int c = 0; c < 4; ++c) {
        const int i = i0 + c * 64 + lane;
        bool has = false;
        if (i < pi.W) {
            has = waypoint_has_hs(bs, (pi.g0 + i) * (int64_t)max_hs, max_hs);
            if (has && clamped && time_clamped(pi, i, dt, tmax_steps))
                atomicAdd(clamped, 1ull);
        }
        const uint64_t m = __ballot(has);
        masks[c] = m;
        any |= m;
    }
    return any;
}

__device__ __forceinline__ float lane_f(float x, int s)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), s));
}

// Tests NQ particle groups (lane = particle of each group) against every
// stored half-space of the marked waypoints; v[q] becomes the AND of the
// particle's tests (PUMP.cu:636-656, same float order).  Per waypoint, the
// slots are loaded lane-parallel (lane h: slot h), the live ones found by a
// ballot and broadcast by readlane, and the offsets of the NQ particles are
// loaded together: one round of memory latency per marked waypoint.
template <int NQ>
__device__ __forceinline__ void particle_walk(bool (&v)[NQ], const int (&p)[NQ], int n_p,
                                              const uint64_t *masks, int i0, const PathInfo &pi,
                                              int max_hs, const float *__restrict__ as,
                                              const float *__restrict__ bs,
                                              const float *__restrict__ off, int n_mc,
                                              int tmax_steps, float dt, float om)
{
    const int lane = threadIdx.x & 63;
    float o[NQ][3];
    int last_ti = -1; // consecutive waypoints often share a time index
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
            if (ti != last_ti) {
                last_ti = ti;
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    // p < n_p <= n_mc: in bounds whatever the particle's validity
                    const float *op = off + ((int64_t)ti * n_mc + p[q]) * 3;
                    const bool in = p[q] < n_p;
                    o[q][0] = in ? op[0] : 0.0f;
                    o[q][1] = in ? op[1] : 0.0f;
                    o[q][2] = in ? op[2] : 0.0f;
                }
            }
            for (int h0 = 0; h0 < max_hs; h0 += 64) {
                const int h = h0 + lane;
                float b = 1111.0f, a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
                if (h < max_hs) {
                    b = bs[head + h];
                    a0 = as[(head + h) * 3 + 0];
                    a1 = as[(head + h) * 3 + 1];
                    a2 = as[(head + h) * 3 + 2];
                }
                // the reference skips slots with b >= 1000 (PUMP.cu:642)
                uint64_t sm = __ballot(!(b >= 1000.0f));
                while (sm) {
                    const int sl = __builtin_ctzll(sm);
                    sm &= sm - 1;
                    const float B = lane_f(b, sl), A0 = lane_f(a0, sl), A1 = lane_f(a1, sl),
                                A2 = lane_f(a2, sl);
#pragma unroll
                    for (int q = 0; q < NQ; ++q)
                        v[q] = v[q] && hs_pass(o[q][0], o[q][1], o[q][2], A0, A1, A2, om, B);
                }
            }
        }
    }
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

// invalid (zero) bytes of a byte-validity row, summed over the wave's lanes
// by the caller; 16-byte loads where the row allows them
__device__ __forceinline__ int invalid_bytes(const uint8_t *__restrict__ r, int n_p, int lane)
{
    int count = 0;
    if (((n_p & 15) == 0) && (((uintptr_t)r) & 15) == 0) {
        const uint4 *r16 = reinterpret_cast<const uint4 *>(r);
        for (int k = lane; k < (n_p >> 4); k += 64) {
            const uint4 x = r16[k];
            count += 16 - __popc(nonzero_bytes01(x.x)) - __popc(nonzero_bytes01(x.y)) -
                     __popc(nonzero_bytes01(x.z)) - __popc(nonzero_bytes01(x.w));
        }
    } else if (((n_p & 3) == 0) && (((uintptr_t)r) & 3) == 0) {
        const uint32_t *r4 = reinterpret_cast<const uint32_t *>(r);
        for (int k = lane; k < (n_p >> 2); k += 64)
            count += 4 - __popc(nonzero_bytes01(r4[k]));
    } else {
        for (int k = lane; k < n_p; k += 64)
            count += r[k] ? 0 : 1;
    }
    return count;
}

// index of this wave within the grid, as a wave-uniform (scalar) value
__device__ __forceinline__ int64_t wave_index()
{
    return (int64_t)blockIdx.x * (blockDim.x >> 6) +
           __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

// The workgroup index and grid size a kernel body works with: the
// hardware's for a one-problem launch, or the problem's own slice of the
// grid when one launch serves many problems (the body's grid-stride loops
// then cover that problem's work only).
struct VGrid {
    int64_t bid, nb;
};
__device__ __forceinline__ VGrid hw_grid() { return VGrid{(int64_t)blockIdx.x, (int64_t)gridDim.x}; }
__device__ __forceinline__ int64_t wave_index(const VGrid &G)
{
    return G.bid * (blockDim.x >> 6) + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}
__device__ __forceinline__ int64_t wave_count(const VGrid &G) { return G.nb * (blockDim.x >> 6); }

constexpr int kWalkGroups = 4; // particle groups of 64 per walk, wave-per-path kernel
constexpr int kHsGroups = 4;   // the same for the group form of the batched pass

// A path that meets at least one stored half-space, as listed by the
// batched classification pass.
struct HsPath {
    int64_t g0;     // first waypoint of the edge
    uint64_t emask; // the edge's waypoint summary (when known)
    float dtl, tbase;
    int W, prev, w, pad;
};

// One (listed path, waypoint with a stored half-space) pair of the
// time-binned form; wp is the waypoint's global index (as/bs row).
struct BinEntry {
    int w;  // path index within the wavefront
    int ti; // clamped time index
    int64_t wp;
};

// A run of at most kBinChunk entries sharing one time index.
struct BinChunk {
    int ti, begin, end;
};

constexpr int kBinChunk = 16;
// Each time index is split into kSubBins counters (by path index) so that
// the counting sort's atomics are spread over more addresses.
constexpr int kSubBins = 16;

// control words at the head of the batched workspace
enum : int {
    kCtlListed = 0,   // listed paths
    kCtlEntries = 1,  // binned entries
    kCtlOverflow = 2, // entries or chunks exceed the workspace: the group form runs
    kCtlChunks = 3,   // binned chunks
};

} // namespace wave

// Listed-path half-space pass of the batched form (hsmc_bins.hip).
// bins_scan turns the sub-binned per-time histogram into time-index starts
// and chunks and sets the overflow flag; bins_finish emits the entries,
// tests them (or, on overflow, runs the group form) and writes the listed
// paths' CP.
struct BinArgs {
    int n_new, n_p, n_mc, tmax_steps, max_hs, known;
    float dt, om;
    const float *as, *bs, *offsets_time;
    const wave::HsPath *recs;
    unsigned long long *acc;
    int *ctl, *hist, *cursor, *tstart;
    wave::BinChunk *chunks;
    wave::BinEntry *entries;
    int64_t entry_cap, chunk_cap;
    const void *valid_src;
    void *valid_dst;
    int64_t dst_base;
    float *path_cp;
    int64_t cp_base;
    // split-BF16 MFMA test (precision 1): particle vectors of the offsets
    // table, and the count of FP32 re-tests in the margin band
    const uint16_t *split;
    unsigned long long *band;
    // particle groups of 64 per wave of the binned test (1, 2, 4 or 8)
    int nq;
};
int bins_scan(const BinArgs &a, hipStream_t s);
int bins_finish(const BinArgs &a, bool bits, hipStream_t s);

// The copy-on-write MC step (bit validity, FP32 tests) of many problems in
// one launch per stage.  A problem's job holds everything its stages read;
// stage j of the launch gives problem k nb[j] workgroups, and block_map[j][b]
// names the problem (high 16 bits) and the problem-local workgroup index
// (low 16 bits) of workgroup b.
enum : int {
    kMcZero = 0,
    kMcClassify,
    kMcScan,
    kMcCopy,
    kMcEmit,
    kMcHs1, // binned tests with 1, 2, 4, 8 particle groups per wave
    kMcHs2,
    kMcHs4,
    kMcHs8,
    kMcGroups,
    kMcCp,
    kMcStages
};
struct McJob {
    BinArgs B;
    // classification inputs
    const int64_t *wp_offsets;
    const int *wf_edge, *wf_prev;
    const uint64_t *edge_mask;
    const float *times, *topts;
    uint8_t *has;
    unsigned long long *clamped;
    int *row_src;
    const float *parent_cp;
    int64_t row_bytes, n_zero;
    int nb[kMcStages];
};
// copies of a job whose pointers are read as global ones (gptr)
__device__ __forceinline__ BinArgs global_args(const BinArgs &a)
{
    BinArgs g = a;
    g.as = gptr(a.as);
    g.bs = gptr(a.bs);
    g.offsets_time = gptr(a.offsets_time);
    g.recs = gptr(a.recs);
    g.acc = gptr(a.acc);
    g.ctl = gptr(a.ctl);
    g.hist = gptr(a.hist);
    g.cursor = gptr(a.cursor);
    g.tstart = gptr(a.tstart);
    g.chunks = gptr(a.chunks);
    g.entries = gptr(a.entries);
    g.valid_src = gptr(a.valid_src);
    g.valid_dst = gptr(a.valid_dst);
    g.path_cp = gptr(a.path_cp);
    g.split = gptr(a.split);
    g.band = gptr(a.band);
    return g;
}
__device__ __forceinline__ McJob global_job(const McJob &j)
{
    McJob g = j;
    g.B = global_args(j.B);
    g.wp_offsets = gptr(j.wp_offsets);
    g.wf_edge = gptr(j.wf_edge);
    g.wf_prev = gptr(j.wf_prev);
    g.edge_mask = gptr(j.edge_mask);
    g.times = gptr(j.times);
    g.topts = gptr(j.topts);
    g.has = gptr(j.has);
    g.clamped = gptr(j.clamped);
    g.row_src = gptr(j.row_src);
    g.parent_cp = gptr(j.parent_cp);
    return g;
}

struct McGrids {
    const uint32_t *block_map[kMcStages]; // device arrays
    int64_t blocks[kMcStages];
};
int bins_many(const McJob *d_jobs, const McGrids &g, bool finish, hipStream_t s);

// The many-problem copy-on-write MC step (hsmc.hip).  mc_cow_job fills one
// problem's job from the arguments of pump_propagate_cp_cow (validity_bits
// 1, precision 0, dst_base and cp_base 0); it fails with PUMP_E_INVALID where
// the one-launch form does not apply (no workspace of the binned layout, no
// waypoint offsets, rows not in 16-byte units).  mc_plan sizes every stage
// for every job (setting their nb) and builds the block maps in one device
// allocation (*d_maps, freed by the caller); mc_cow_many launches the
// stages for jobs already in device memory.
int mc_cow_job(McJob &J, int n_new, const int64_t *wp_offsets, const int *wf_edge,
               const int *wf_prev, const float *as, const float *bs, const uint64_t *edge_mask,
               const float *times, const float *topts, float dt, int max_hs,
               const float *offsets_time, int tmax_steps, float om, int n_mc, const void *src,
               const float *parent_cp, void *dst, int *row_src, int n_p, float *path_cp,
               void *ws, size_t ws_bytes);
int mc_plan(McJob *jobs, int n_jobs, McGrids &g, void **d_maps, hipStream_t s);
int mc_cow_many(const McJob *d_jobs, const McGrids &g, hipStream_t s);

} // namespace pump
