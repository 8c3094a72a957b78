// hsmc_bins.hip -- the listed paths' half-space pass of the batched HSMC
// step (reference: propagateValidParticles PUMP.cu:613-705, calculateCP_HSMC
// PUMP.cu:745-758).
//
// The particle test of PUMP.cu:641-651 reads row ti of the offsets table for
// every (path, waypoint) pair that holds a stored half-space.  Many such pairs
// of a wavefront share a time index, so they are sorted by it (counting sort:
// histogram in the classification pass, scan and scatter here) and cut into
// chunks of at most kBinChunk pairs.  A wave takes one (chunk, 256
// particles): it loads the particles' offsets of row ti once into registers
// and the chunk's half-spaces lane-parallel (lane j: pair j), then tests
// every pair against its particles.  A pair's result is ANDed into its
// path's child row, which already holds the parent row (the copy pass copies
// listed rows too in this form); most pairs invalidate nothing and issue no
// store.  The AND is order-independent and each particle decision uses the
// reference's float order, so rows and CPs are bit-identical to the
// per-path forms.  When the pairs overflow the workspace, the same launch
// runs the group form instead (one wave per listed path and 512 particles).
#include <algorithm>

#include "hsmc_mfma.hpp"
#include "hsmc_wave.hpp"
#include "pump_internal.hpp"

#pragma clang fp contract(off)

namespace pump {
namespace {

using namespace wave;

constexpr int kScanBlock = 1024;
constexpr int kTestGroups = 4; // particle groups of 64 per test wave, unless tuned
constexpr int kSlotTile = 8;   // half-space slots staged per pass

// exclusive scan of (x, y) over the block; returns the totals in *tx, *ty
__device__ __forceinline__ void block_scan2(int &x, int &y, int *tx, int *ty)
{
    __shared__ int sx[kScanBlock / 64], sy[kScanBlock / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int ix = x, iy = y;
    for (int off = 1; off < 64; off <<= 1) {
        const int ux = __shfl_up(ix, off, 64), uy = __shfl_up(iy, off, 64);
        if (lane >= off) {
            ix += ux;
            iy += uy;
        }
    }
    if (lane == 63) {
        sx[wid] = ix;
        sy[wid] = iy;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int ax = 0, ay = 0;
        for (int k = 0; k < kScanBlock / 64; ++k) {
            const int cx = sx[k], cy = sy[k];
            sx[k] = ax;
            sy[k] = ay;
            ax += cx;
            ay += cy;
        }
        *tx = ax;
        *ty = ay;
    }
    __syncthreads();
    x = sx[wid] + ix - x;
    y = sy[wid] + iy - y;
}

// total of time index t over its kSubBins counters, and the part of it held
// by sub-bins below `sub`
__device__ __forceinline__ int time_total(const int *__restrict__ hist, int t, int sub,
                                          int *below)
{
    const int4 *h4 = reinterpret_cast<const int4 *>(hist + (int64_t)t * kSubBins);
    int total = 0, pre = 0;
#pragma unroll
    for (int q = 0; q < kSubBins / 4; ++q) {
        const int4 x = h4[q];
        const int k = q * 4;
        total += x.x + x.y + x.z + x.w;
        pre += (k < sub ? x.x : 0) + (k + 1 < sub ? x.y : 0) + (k + 2 < sub ? x.z : 0) +
               (k + 3 < sub ? x.w : 0);
    }
    *below = pre;
    return total;
}

// One workgroup: time-index starts, chunks, entry/chunk totals, overflow.
// The pairs of time index t occupy [tstart[t], tstart[t+1]), sub-bin by
// sub-bin (sub-bin k: paths with w % kSubBins == k).
__device__ __forceinline__ void bin_scan_body(int T, const int *__restrict__ hist,
                                              int *__restrict__ tstart,
                                              BinChunk *__restrict__ chunks, int *__restrict__ ctl,
                                              int64_t entry_cap, int64_t chunk_cap)
{
    __shared__ int tot_e, tot_c;
    constexpr int kKeep = 4; // time indices per thread whose totals stay in registers
    const int per = (T + kScanBlock - 1) / kScanBlock;
    const int t0 = threadIdx.x * per;
    const int t1 = (t0 + per) < T ? (t0 + per) : T;
    int se = 0, sc = 0, unused;
    int keep[kKeep];
#pragma unroll
    for (int j = 0; j < kKeep; ++j)
        keep[j] = (t0 + j < t1) ? time_total(hist, t0 + j, 0, &unused) : 0;
    for (int t = t0; t < t1; ++t) {
        const int j = t - t0;
        const int h = j < kKeep ? keep[j] : time_total(hist, t, 0, &unused);
        se += h;
        sc += (h + kBinChunk - 1) / kBinChunk;
    }
    block_scan2(se, sc, &tot_e, &tot_c);
    const bool overflow = (int64_t)tot_e > entry_cap || (int64_t)tot_c > chunk_cap;
    for (int t = t0; t < t1; ++t) {
        const int j = t - t0;
        const int h = j < kKeep ? keep[j] : time_total(hist, t, 0, &unused);
        tstart[t] = se;
        if (!overflow)
            for (int j = 0; j * kBinChunk < h; ++j) {
                BinChunk c;
                c.ti = t;
                c.begin = se + j * kBinChunk;
                c.end = se + ((j + 1) * kBinChunk < h ? (j + 1) * kBinChunk : h);
                chunks[sc + j] = c;
            }
        se += h;
        sc += (h + kBinChunk - 1) / kBinChunk;
    }
    if (threadIdx.x == 0) {
        tstart[T] = tot_e;
        ctl[kCtlEntries] = tot_e;
        ctl[kCtlChunks] = tot_c;
        ctl[kCtlOverflow] = overflow ? 1 : 0;
    }
}

__global__ __launch_bounds__(kScanBlock) void k_bin_scan(int T, const int *__restrict__ hist,
                                                         int *__restrict__ tstart,
                                                         BinChunk *__restrict__ chunks,
                                                         int *__restrict__ ctl, int64_t entry_cap,
                                                         int64_t chunk_cap)
{
    bin_scan_body(T, hist, tstart, chunks, ctl, entry_cap, chunk_cap);
}

// lane = listed path: its (waypoint, time index) pairs into the sorted array
// at tstart[ti] + (pairs of lower sub-bins) + (rank within the sub-bin)
__device__ __forceinline__ void bin_emit_body(const HsPath *__restrict__ recs,
                                              const int *__restrict__ ctl, int known,
                                              const float *__restrict__ bs, int max_hs, float dt,
                                              int tmax_steps, const int *__restrict__ hist,
                                              const int *__restrict__ tstart,
                                              int *__restrict__ cursor,
                                              BinEntry *__restrict__ entries, const VGrid &G)
{
    if (ctl[kCtlOverflow])
        return;
    const int n = ctl[kCtlListed];
    for (int64_t slot = G.bid * blockDim.x + threadIdx.x; slot < n;
         slot += G.nb * blockDim.x) {
        const HsPath r = recs[slot];
        PathInfo pi;
        pi.g0 = r.g0;
        pi.W = r.W;
        pi.dtl = r.dtl;
        pi.tbase = r.tbase;
        const int sub = r.w & (kSubBins - 1);
        for_each_marked(pi, known != 0, r.emask, bs, max_hs, [&](int i) {
            const int ti = clamp_time(time_index(pi, i, dt), tmax_steps);
            int below;
            time_total(hist, ti, sub, &below);
            const int pos =
                tstart[ti] + below + atomicAdd(&cursor[(int64_t)ti * kSubBins + sub], 1);
            BinEntry en;
            en.w = r.w;
            en.ti = ti;
            en.wp = r.g0 + i;
            entries[pos] = en;
        });
    }
}

__global__ __launch_bounds__(256) void k_bin_emit(const HsPath *__restrict__ recs,
                                                  const int *__restrict__ ctl, int known,
                                                  const float *__restrict__ bs, int max_hs,
                                                  float dt, int tmax_steps,
                                                  const int *__restrict__ hist,
                                                  const int *__restrict__ tstart,
                                                  int *__restrict__ cursor,
                                                  BinEntry *__restrict__ entries)
{
    bin_emit_body(recs, ctl, known, bs, max_hs, dt, tmax_steps, hist, tstart, cursor, entries,
                  hw_grid());
}

// Binned test: one wave per (chunk, 64*NQ particles).
template <bool kBits, int NQ>
__device__ __forceinline__ void bin_test(const BinArgs &a, int64_t n_chunks, const VGrid &G)
{
    const int lane = threadIdx.x & 63;
    const int n_p = a.n_p, max_hs = a.max_hs;
    const float om = a.om;
    const int n_groups = (n_p + 63) >> 6;
    const int n_sg = (n_groups + NQ - 1) / NQ;
    const int64_t n_items = n_chunks * n_sg;
    const int64_t n_waves = wave_count(G);
    for (int64_t item = wave_index(G); item < n_items; item += n_waves) {
        const int ch = (int)(item / n_sg);
        const int sg = (int)(item - (int64_t)ch * n_sg);
        const BinChunk c = a.chunks[ch];
        const int ne = c.end - c.begin; // 1..kBinChunk
        int p[NQ];
        float o[NQ][3];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            p[q] = (sg * NQ + q) * 64 + lane;
            // p < n_p <= n_mc keeps the row read in bounds
            const float *op = a.offsets_time + ((int64_t)c.ti * a.n_mc + p[q]) * 3;
            const bool in = p[q] < n_p;
            o[q][0] = in ? op[0] : 0.0f;
            o[q][1] = in ? op[1] : 0.0f;
            o[q][2] = in ? op[2] : 0.0f;
        }
        int my_w = 0;
        int64_t my_wp = 0;
        if (lane < ne) {
            const BinEntry en = a.entries[c.begin + lane];
            my_w = en.w;
            my_wp = en.wp;
        }
        for (int h0 = 0; h0 < max_hs; h0 += kSlotTile) {
            // lane j stages slots h0..h0+kSlotTile-1 of pair j and marks the
            // live ones (the reference skips b >= 1000, PUMP.cu:642); the
            // normals are loaded for live slots only
            float b[kSlotTile], av[kSlotTile][3];
            int live = 0;
#pragma unroll
            for (int h = 0; h < kSlotTile; ++h) {
                b[h] = 1111.0f;
                if (lane < ne && h0 + h < max_hs)
                    b[h] = a.bs[my_wp * max_hs + h0 + h];
                live |= !(b[h] >= 1000.0f) ? (1 << h) : 0;
            }
#pragma unroll
            for (int h = 0; h < kSlotTile; ++h) {
                av[h][0] = av[h][1] = av[h][2] = 0.0f;
                if (live & (1 << h)) {
                    const int64_t k = my_wp * max_hs + h0 + h;
                    av[h][0] = a.as[k * 3 + 0];
                    av[h][1] = a.as[k * 3 + 1];
                    av[h][2] = a.as[k * 3 + 2];
                }
            }
            for (int j = 0; j < ne; ++j) {
                const int lm = __builtin_amdgcn_readlane(live, j);
                if (lm == 0)
                    continue;
                bool v[NQ];
#pragma unroll
                for (int q = 0; q < NQ; ++q)
                    v[q] = true;
#pragma unroll
                for (int h = 0; h < kSlotTile; ++h) {
                    if (!(lm & (1 << h)))
                        continue;
                    const float B = lane_f(b[h], j), A0 = lane_f(av[h][0], j),
                                A1 = lane_f(av[h][1], j), A2 = lane_f(av[h][2], j);
#pragma unroll
                    for (int q = 0; q < NQ; ++q)
                        v[q] = v[q] && hs_pass(o[q][0], o[q][1], o[q][2], A0, A1, A2, om, B);
                }
                const int64_t row = a.dst_base + __builtin_amdgcn_readlane(my_w, j);
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    const int grp = sg * NQ + q;
                    if (grp >= n_groups)
                        break;
                    if (kBits) {
                        // bits past n_p stay as they are (pass = 1)
                        const uint64_t pass = __ballot(v[q] || p[q] >= n_p);
                        if (pass != ~0ull && lane == 0)
                            atomicAnd((unsigned long long *)a.valid_dst + row * n_groups + grp,
                                      (unsigned long long)pass);
                    } else if (!v[q] && p[q] < n_p) {
                        ((uint8_t *)a.valid_dst)[row * n_p + p[q]] = 0;
                    }
                }
            }
        }
    }
}

// Split-BF16 MFMA form of bin_test (bit-identical decisions; see
// hsmc_mfma.hpp for the operand layout and the margin).  One wave per
// (chunk, 256 particles): the chunk's live (pair, slot) half-spaces become
// rows of 16-row tiles staged in the wave's LDS slice; particles are taken 16
// at a time from the precomputed particle vectors; each MFMA tile decides 16
// x 16 tests with two compares per element against the row's margin (built
// from the table-wide bound on |o|), band cases are re-tested in FP32, and a
// row's failures in a 16-particle group are ANDed into its path's word with
// one atomic.  A half-space with a = 0 and b not > 0 fails every particle
// (LHS is 0 or NaN): it clears the item's particle range of its path at once.
constexpr int kMfmaRows = kBinChunk * kSlotTile; // rows per chunk at most

struct MfmaRow {
    float a0, a1, a2, b;
    int w;
};

template <bool kBits, int NQ>
__device__ __forceinline__ void clear_range(const BinArgs &a, int w, int p_begin, int lane)
{
    const int n_p = a.n_p, n_groups = (n_p + 63) >> 6;
    if (kBits) {
        // whole 64-bit words: bits past n_p are zero already
        const int g0 = p_begin >> 6;
        for (int g = g0; g < g0 + NQ && g < n_groups; ++g)
            atomicAnd((unsigned long long *)a.valid_dst + (a.dst_base + w) * n_groups + g, 0ull);
    } else {
        uint8_t *r = (uint8_t *)a.valid_dst + (a.dst_base + w) * n_p;
        for (int p = p_begin; p < p_begin + 64 * NQ && p < n_p; ++p)
            r[p] = 0;
    }
    (void)lane;
}

template <bool kBits, int NQ>
__device__ __forceinline__ void bin_test_mfma(const BinArgs &a, int64_t n_chunks, MfmaRow *rows)
{
    using namespace mfma;
    // NQ: 64-particle groups per item (16-particle tiles: 4*NQ)
    constexpr int NT = 4 * NQ;
    const int lane = threadIdx.x & 63, q = lane >> 4, col = lane & 15;
    const int n_p = a.n_p, max_hs = a.max_hs;
    const float om = a.om;
    const int n_groups = (n_p + 63) >> 6;
    const int n_sg = (n_groups + NQ - 1) / NQ;
    const int64_t n_items = n_chunks * n_sg;
    const int64_t n_waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    // max |o_d| over the whole offsets table, stored after the vectors
    const float omax = *reinterpret_cast<const float *>(
        a.split + (int64_t)a.tmax_steps * a.n_mc * kSplitK);
    unsigned n_band = 0; // per lane
    for (int64_t item = wave_index(); item < n_items; item += n_waves) {
        const int ch = (int)(item / n_sg);
        const int sg = (int)(item - (int64_t)ch * n_sg);
        const BinChunk c = a.chunks[ch];
        const int ne = c.end - c.begin;
        const int pbase = sg * 64 * NQ;
        // stage the chunk's live half-spaces as rows, pair by pair
        int my_w = 0, live = 0, dead = 0;
        float b[kSlotTile], av[kSlotTile][3];
        if (lane < ne) {
            const BinEntry en = a.entries[c.begin + lane];
            my_w = en.w;
#pragma unroll
            for (int h = 0; h < kSlotTile; ++h) {
                b[h] = 1111.0f;
                av[h][0] = av[h][1] = av[h][2] = 0.0f;
                if (h < max_hs) {
                    const int64_t k = en.wp * max_hs + h;
                    b[h] = a.bs[k];
                    av[h][0] = a.as[k * 3 + 0];
                    av[h][1] = a.as[k * 3 + 1];
                    av[h][2] = a.as[k * 3 + 2];
                }
                // the reference skips slots with b >= 1000 (PUMP.cu:642)
                if (!(b[h] >= 1000.0f)) {
                    if (av[h][0] == 0.0f && av[h][1] == 0.0f && av[h][2] == 0.0f &&
                        !(b[h] > 0.0f))
                        dead = 1;
                    else
                        live |= 1 << h;
                }
            }
        }
        if (dead) {
            clear_range<kBits, NQ>(a, my_w, pbase, lane);
            live = 0; // every particle of the range has failed already
        }
        const int cnt = __popc(live);
        int pre = cnt; // inclusive scan over lanes
        for (int off = 1; off < 64; off <<= 1) {
            const int u = __shfl_up(pre, off, 64);
            if (lane >= off)
                pre += u;
        }
        const int n_rows = __builtin_amdgcn_readlane(pre, 63);
        pre -= cnt;
        if (lane < ne) {
            int r = pre;
#pragma unroll
            for (int h = 0; h < kSlotTile; ++h)
                if (live & (1 << h)) {
                    MfmaRow e;
                    e.a0 = av[h][0];
                    e.a1 = av[h][1];
                    e.a2 = av[h][2];
                    e.b = b[h];
                    e.w = my_w;
                    rows[r++] = e;
                }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // this lane's 4 bf16 of particle pbase + col; tile nt is nt*16 vectors on
        const bf16x4 *sb = reinterpret_cast<const bf16x4 *>(
            a.split + ((int64_t)c.ti * a.n_mc + pbase + col) * kSplitK + 4 * q);
        const int p_left = n_p - pbase - col; // tile nt holds this lane's particle iff nt*16 < p_left
        auto load_b = [&](int nt) {
            bf16x4 v = pack4(0, 0, 0, 0);
            if (nt < NT && nt * 16 < p_left)
                v = sb[nt * 16 * (kSplitK / 4)];
            return v;
        };
        for (int m0 = 0; m0 < n_rows; m0 += 16) {
            // A operand: row m0 + col; margins of the rows this lane's D holds
            bf16x4 aop = pack4(0, 0, 0, 0);
            int wcol = 0; // path of row m0 + col
            {
                const int r = m0 + col;
                if (r < n_rows) {
                    const MfmaRow e = rows[r];
                    aop = halfspace_operand(q, e.a0, e.a1, e.a2, e.b / om);
                    wcol = e.w;
                }
            }
            float mg[4];
            uint64_t rv[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int r = m0 + 4 * q + j;
                mg[j] = 0.0f;
                if (r < n_rows) {
                    const MfmaRow e = rows[r];
                    mg[j] = fmaf(omax, kMarginDot * (fabsf(e.a0) + fabsf(e.a1) + fabsf(e.a2)),
                                 kMarginB * fabsf(e.b / om) + 1e-30f);
                }
                rv[j] = __ballot(r < n_rows);
            }
            // one 64-particle word (four 16-particle tiles) per step: the four
            // MFMAs issue back to back, the next word's operands load meanwhile
            bf16x4 nb[4];
#pragma unroll
            for (int t = 0; t < 4; ++t)
                nb[t] = load_b(t);
            for (int wd = 0; wd < NQ; ++wd) {
                const int pw = pbase + wd * 64;
                if (pw >= n_p)
                    break;
                bf16x4 cb[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    cb[t] = nb[t];
                    nb[t] = load_b((wd + 1) * 4 + t);
                }
                f32x4 d[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
                    d[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(aop, cb[t], zero, 0, 0, 0);
                }
                // lane r < 16: failures of row m0 + r in this word
                uint64_t acc = 0;
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int p0 = pw + t * 16;
                    const int p = p0 + col;
                    const uint64_t pv = __ballot(p < n_p);
                    uint64_t fail[4], band = 0, any_fail = 0;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const uint64_t use = pv & rv[j];
                        const uint64_t clear = __ballot(fabsf(d[t][j]) > mg[j]); // not NaN
                        const uint64_t pos = __ballot(d[t][j] > 0.0f);
                        fail[j] = use & clear & pos;
                        any_fail |= fail[j];
                        band |= (use & ~clear) ? (1ull << j) : 0ull; // rows with band lanes
                    }
                    if (band) {
                        // FP32 re-test in the reference's order for band cases
                        const float *op = a.offsets_time + ((int64_t)c.ti * a.n_mc + p) * 3;
                        for (int j = 0; j < 4; ++j) {
                            if (!((band >> j) & 1))
                                continue;
                            const uint64_t use = pv & rv[j];
                            const bool mine = (use >> lane) & 1ull && !(fabsf(d[t][j]) > mg[j]);
                            bool f = false;
                            if (mine) {
                                const MfmaRow e = rows[m0 + 4 * q + j];
                                f = !hs_pass(op[0], op[1], op[2], e.a0, e.a1, e.a2, om, e.b);
                                ++n_band;
                            }
                            fail[j] |= __ballot(f);
                            any_fail |= fail[j];
                        }
                    }
                    if (!any_fail)
                        continue;
                    if (kBits) {
                        // row m0+4g+j, particles p0..p0+15: bits 16g.. of fail[j]
                        const int j = col & 3, g = col >> 2;
                        const uint64_t fj = j == 0 ? fail[0] : j == 1 ? fail[1]
                                          : j == 2 ? fail[2] : fail[3];
                        acc |= ((fj >> (16 * g)) & 0xffffull) << (16 * t);
                    } else {
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            if ((fail[j] >> lane) & 1ull)
                                ((uint8_t *)a.valid_dst)[(a.dst_base + rows[m0 + 4 * q + j].w) *
                                                             n_p +
                                                         p] = 0;
                    }
                }
                if (kBits && lane < 16 && acc)
                    atomicAnd((unsigned long long *)a.valid_dst +
                                  (a.dst_base + wcol) * n_groups + (pw >> 6),
                              ~(unsigned long long)acc);
            }
        }
        // the next item overwrites the rows
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    const int total_band = wave_sum((int)n_band);
    if (a.band && lane == 0 && total_band)
        atomicAdd(a.band, (unsigned long long)total_band);
}

// Group form (used when the binned pairs overflow the workspace): one wave
// per (listed path, 64*NQ particles) walks the path's marked waypoints and
// writes the child words from the parent row; the path's CP is finished by
// the last of its waves (a 64-bit atomic ticket packing the finished-wave
// count and the particle count).
template <bool kBits, int NQ>
__device__ __forceinline__ void group_pass(const BinArgs &a, int64_t n_listed, const VGrid &G)
{
    const int lane = threadIdx.x & 63;
    const int n_p = a.n_p;
    const int n_groups = (n_p + 63) >> 6;
    const int n_sg = (n_groups + NQ - 1) / NQ; // waves per listed path
    const int64_t n_items = n_listed * n_sg;
    const int64_t n_waves = wave_count(G);
    for (int64_t item = wave_index(G); item < n_items; item += n_waves) {
        const int slot = (int)(item / n_sg);
        const int sg = (int)(item - (int64_t)slot * n_sg);
        const HsPath r = a.recs[slot];
        PathInfo pi;
        pi.g0 = r.g0;
        pi.W = r.W;
        pi.dtl = r.dtl;
        pi.tbase = r.tbase;
        bool v[NQ];
        int p[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int grp = sg * NQ + q;
            p[q] = grp * 64 + lane;
            if (kBits) {
                const uint64_t word =
                    grp < n_groups
                        ? ((const uint64_t *)a.valid_src)[(int64_t)r.prev * n_groups + grp]
                        : 0ull;
                v[q] = (p[q] < n_p) && ((word >> lane) & 1ull);
            } else {
                v[q] = (p[q] < n_p) &&
                       ((const uint8_t *)a.valid_src)[(int64_t)r.prev * n_p + p[q]] != 0;
            }
        }
        for (int i0 = 0; i0 < pi.W; i0 += 256) {
            uint64_t masks[4];
            const uint64_t any = hs_masks(masks, pi, i0, a.known != 0, r.emask, a.max_hs, a.bs,
                                          a.dt, a.tmax_steps, (unsigned long long *)nullptr);
            uint64_t alive = 0;
#pragma unroll
            for (int q = 0; q < NQ; ++q)
                alive |= __ballot(v[q]);
            if (any && alive)
                particle_walk<NQ>(v, p, n_p, masks, i0, pi, a.max_hs, a.as, a.bs, a.offsets_time,
                                  a.n_mc, a.tmax_steps, a.dt, a.om);
        }
        int count = 0; // valid (bits) or invalid (bytes) particles of this wave
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int grp = sg * NQ + q;
            const uint64_t out = __ballot(v[q]);
            if (grp < n_groups) {
                const int64_t row = a.dst_base + r.w;
                if (kBits) {
                    if (lane == q)
                        ((uint64_t *)a.valid_dst)[row * n_groups + grp] = out;
                    count += __popcll(out);
                } else {
                    if (p[q] < n_p)
                        ((uint8_t *)a.valid_dst)[row * n_p + p[q]] = v[q] ? 1 : 0;
                    const int in_group = (n_p - grp * 64) < 64 ? (n_p - grp * 64) : 64;
                    count += in_group - __popcll(out);
                }
            }
        }
        if (lane == 0) {
            const unsigned long long old =
                atomicAdd(&a.acc[slot], (1ull << 32) | (unsigned long long)count);
            if ((int)(old >> 32) + 1 == n_sg && a.path_cp) {
                const int total = (int)(old & 0xffffffffull) + count;
                const int invalid = kBits ? n_p - total : total;
                a.path_cp[a.cp_base + r.w] = ((float)invalid) / (float)n_p;
            }
        }
    }
}

// The listed paths' half-space tests, binned (FP32, or split-BF16 MFMA when
// a particle-vector table is given); k_hs_groups runs instead when the
// binned pairs overflowed.  (Separate kernels keep each one's registers.)
template <bool kBits, bool kMfma, int NQ>
__global__ __launch_bounds__(256) void k_hs_pass(const BinArgs a)
{
    if (a.ctl[kCtlOverflow] != 0)
        return;
    if (kMfma) {
        __shared__ MfmaRow rows[256 / 64][kMfmaRows];
        bin_test_mfma<kBits, NQ>(a, a.ctl[kCtlChunks], rows[threadIdx.x >> 6]);
    } else {
        bin_test<kBits, NQ>(a, a.ctl[kCtlChunks], hw_grid());
    }
}

// k_hs_pass with the particle groups per wave given at run time (1, 2, 4, 8)
template <bool kBits, bool kMfma>
void launch_hs_pass(const BinArgs &a, int nq, unsigned grid, hipStream_t s)
{
    switch (nq) {
    case 1:
        hipLaunchKernelGGL((k_hs_pass<kBits, kMfma, 1>), dim3(grid), dim3(256), 0, s, a);
        break;
    case 2:
        hipLaunchKernelGGL((k_hs_pass<kBits, kMfma, 2>), dim3(grid), dim3(256), 0, s, a);
        break;
    case 8:
        hipLaunchKernelGGL((k_hs_pass<kBits, kMfma, 8>), dim3(grid), dim3(256), 0, s, a);
        break;
    default:
        hipLaunchKernelGGL((k_hs_pass<kBits, kMfma, kTestGroups>), dim3(grid), dim3(256), 0, s,
                           a);
    }
}

template <bool kBits>
__global__ __launch_bounds__(256) void k_hs_groups(const BinArgs a)
{
    if (a.ctl[kCtlOverflow] == 0)
        return;
    group_pass<kBits, kHsGroups>(a, a.ctl[kCtlListed], hw_grid());
}

// One wave per listed path: CP from its finished child row.
template <bool kBits>
__device__ __forceinline__ void cp_listed_body(const HsPath *__restrict__ recs,
                                               const int *__restrict__ ctl,
                                               const void *__restrict__ valid_dst,
                                               int64_t dst_base, int n_p,
                                               float *__restrict__ path_cp, int64_t cp_base,
                                               const VGrid &G)
{
    if (ctl[kCtlOverflow])
        return;
    const int lane = threadIdx.x & 63;
    const int n = ctl[kCtlListed];
    const int64_t n_waves = wave_count(G);
    for (int64_t slot = wave_index(G); slot < n; slot += n_waves) {
        const int64_t row = dst_base + recs[slot].w;
        int count = 0; // valid (bits) or invalid (bytes) particles
        if (kBits) {
            const int nw = (n_p + 63) >> 6;
            const uint64_t *r = (const uint64_t *)valid_dst + row * nw;
            for (int k = lane; k < nw; k += 64)
                count += __popcll(r[k]);
        } else {
            count = invalid_bytes((const uint8_t *)valid_dst + row * n_p, n_p, lane);
        }
        count = wave_sum(count);
        if (lane == 0 && path_cp) {
            const int invalid = kBits ? n_p - count : count;
            path_cp[cp_base + recs[slot].w] = ((float)invalid) / (float)n_p;
        }
    }
}

template <bool kBits>
__global__ __launch_bounds__(256) void k_cp_listed(const HsPath *__restrict__ recs,
                                                   const int *__restrict__ ctl,
                                                   const void *__restrict__ valid_dst,
                                                   int64_t dst_base, int n_p,
                                                   float *__restrict__ path_cp, int64_t cp_base)
{
    cp_listed_body<kBits>(recs, ctl, valid_dst, dst_base, n_p, path_cp, cp_base, hw_grid());
}

// ---------------------------------------------------------------------------
// The same stages for many problems per launch (copy-on-write step, bit
// validity, FP32 tests): workgroup b serves problem map[b] >> 16 as its
// workgroup map[b] & 0xffff of nb[stage].
// ---------------------------------------------------------------------------
__device__ __forceinline__ const McJob &mc_job(const McJob *__restrict__ jobs,
                                               const uint32_t *__restrict__ map, int stage,
                                               VGrid &G)
{
    const uint32_t m = map[blockIdx.x];
    const McJob &J = jobs[m >> 16];
    G = VGrid{(int64_t)(m & 0xffffu), (int64_t)J.nb[stage]};
    return J;
}

__global__ __launch_bounds__(kScanBlock) void k_bin_scan_many(const McJob *__restrict__ jobs,
                                                              const uint32_t *__restrict__ map)
{
    VGrid G;
    const BinArgs a = global_args(mc_job(jobs, map, kMcScan, G).B);
    bin_scan_body(a.tmax_steps, a.hist, a.tstart, a.chunks, a.ctl, a.entry_cap, a.chunk_cap);
}

__global__ __launch_bounds__(256) void k_bin_emit_many(const McJob *__restrict__ jobs,
                                                       const uint32_t *__restrict__ map)
{
    VGrid G;
    const BinArgs a = global_args(mc_job(jobs, map, kMcEmit, G).B);
    bin_emit_body(a.recs, a.ctl, a.known, a.bs, a.max_hs, a.dt, a.tmax_steps, a.hist, a.tstart,
                  a.cursor, a.entries, G);
}

template <int NQ>
__global__ __launch_bounds__(256) void k_hs_pass_many(const McJob *__restrict__ jobs,
                                                      const uint32_t *__restrict__ map, int stage)
{
    VGrid G;
    const BinArgs a = global_args(mc_job(jobs, map, stage, G).B);
    if (a.ctl[kCtlOverflow] != 0)
        return;
    bin_test<true, NQ>(a, a.ctl[kCtlChunks], G);
}

__global__ __launch_bounds__(256) void k_hs_groups_many(const McJob *__restrict__ jobs,
                                                        const uint32_t *__restrict__ map)
{
    VGrid G;
    const BinArgs a = global_args(mc_job(jobs, map, kMcGroups, G).B);
    if (a.ctl[kCtlOverflow] == 0)
        return;
    group_pass<true, kHsGroups>(a, a.ctl[kCtlListed], G);
}

__global__ __launch_bounds__(256) void k_cp_listed_many(const McJob *__restrict__ jobs,
                                                        const uint32_t *__restrict__ map)
{
    VGrid G;
    const BinArgs a = global_args(mc_job(jobs, map, kMcCp, G).B);
    cp_listed_body<true>(a.recs, a.ctl, a.valid_dst, a.dst_base, a.n_p, a.path_cp, a.cp_base, G);
}

unsigned wave_blocks(int64_t waves)
{
    int dev = 0, cus = 256;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int64_t cap = (int64_t)cus * 32;
    int64_t blocks = (waves + 3) / 4;
    return (unsigned)(blocks < 1 ? 1 : (blocks < cap ? blocks : cap));
}

} // namespace

// many problems' scan stage (finish = false, after their classification),
// or their emit, test, group and CP stages (finish = true, after the row
// copies)
int bins_many(const McJob *d_jobs, const McGrids &g, bool finish, hipStream_t s)
{
    const auto launch = [&](int stage, auto kernel, int block) {
        if (g.blocks[stage] > 0)
            hipLaunchKernelGGL(kernel, dim3((unsigned)g.blocks[stage]), dim3(block), 0, s, d_jobs,
                               g.block_map[stage]);
    };
    if (!finish) {
        launch(kMcScan, k_bin_scan_many, kScanBlock);
        PUMP_LAUNCHED();
        return PUMP_OK;
    }
    launch(kMcEmit, k_bin_emit_many, 256);
    PUMP_LAUNCHED();
    const int hs_stage[4] = {kMcHs1, kMcHs2, kMcHs4, kMcHs8};
    for (int i = 0; i < 4; ++i) {
        const int st = hs_stage[i];
        if (g.blocks[st] <= 0)
            continue;
        const dim3 grid((unsigned)g.blocks[st]);
        if (i == 0)
            hipLaunchKernelGGL(k_hs_pass_many<1>, grid, dim3(256), 0, s, d_jobs, g.block_map[st], st);
        else if (i == 1)
            hipLaunchKernelGGL(k_hs_pass_many<2>, grid, dim3(256), 0, s, d_jobs, g.block_map[st], st);
        else if (i == 2)
            hipLaunchKernelGGL(k_hs_pass_many<4>, grid, dim3(256), 0, s, d_jobs, g.block_map[st], st);
        else
            hipLaunchKernelGGL(k_hs_pass_many<8>, grid, dim3(256), 0, s, d_jobs, g.block_map[st], st);
        PUMP_LAUNCHED();
    }
    launch(kMcGroups, k_hs_groups_many, 256);
    PUMP_LAUNCHED();
    launch(kMcCp, k_cp_listed_many, 256);
    PUMP_LAUNCHED();
    return PUMP_OK;
}

int bins_scan(const BinArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(kScanBlock), 0, s, a.tmax_steps,
                       (const int *)a.hist, a.tstart, a.chunks, a.ctl, a.entry_cap, a.chunk_cap);
    PUMP_LAUNCHED();
    return PUMP_OK;
}

int bins_finish(const BinArgs &a, bool bits, hipStream_t s)
{
    // device-side counts; the grids are sized for their upper bounds
    hipLaunchKernelGGL(k_bin_emit, dim3(grid_for(a.n_new, 256, 4096)), dim3(256), 0, s, a.recs,
                       (const int *)a.ctl, a.known, a.bs, a.max_hs, a.dt, a.tmax_steps,
                       (const int *)a.hist, (const int *)a.tstart, a.cursor, a.entries);
    PUMP_LAUNCHED();
    const int64_t groups = (a.n_p + 63) / 64;
    const int nq = (a.nq == 1 || a.nq == 2 || a.nq == 8) ? a.nq : kTestGroups;
    const unsigned g_hs = wave_blocks(a.chunk_cap * ((groups + nq - 1) / nq));
    // the group form runs only after an overflow (its waves stride over the
    // work): a small grid keeps the usual early exit cheap
    const unsigned g_gr = std::min(
        wave_blocks((int64_t)a.n_new * ((groups + kHsGroups - 1) / kHsGroups)), 1024u);
    const unsigned g_cp = wave_blocks(a.n_new);
    const bool mfma = a.split && a.max_hs <= kSlotTile;
    if (bits) {
        if (mfma)
            launch_hs_pass<true, true>(a, nq, g_hs, s);
        else
            launch_hs_pass<true, false>(a, nq, g_hs, s);
        PUMP_LAUNCHED();
        hipLaunchKernelGGL(k_hs_groups<true>, dim3(g_gr), dim3(256), 0, s, a);
        PUMP_LAUNCHED();
        hipLaunchKernelGGL(k_cp_listed<true>, dim3(g_cp), dim3(256), 0, s, a.recs,
                           (const int *)a.ctl, (const void *)a.valid_dst, a.dst_base, a.n_p,
                           a.path_cp, a.cp_base);
    } else {
        if (mfma)
            launch_hs_pass<false, true>(a, nq, g_hs, s);
        else
            launch_hs_pass<false, false>(a, nq, g_hs, s);
        PUMP_LAUNCHED();
        hipLaunchKernelGGL(k_hs_groups<false>, dim3(g_gr), dim3(256), 0, s, a);
        PUMP_LAUNCHED();
        hipLaunchKernelGGL(k_cp_listed<false>, dim3(g_cp), dim3(256), 0, s, a.recs,
                           (const int *)a.ctl, (const void *)a.valid_dst, a.dst_base, a.n_p,
                           a.path_cp, a.cp_base);
    }
    PUMP_LAUNCHED();
    return PUMP_OK;
}

} // namespace pump
