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
#include "hsmc_wave.hpp"
#include "pump_internal.hpp"

#pragma clang fp contract(off)

namespace pump {
namespace {

using namespace wave;

constexpr int kScanBlock = 1024;
constexpr int kTestGroups = 4; // particle groups of 64 per test wave
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
__global__ __launch_bounds__(kScanBlock) void k_bin_scan(int T, const int *__restrict__ hist,
                                                         int *__restrict__ tstart,
                                                         BinChunk *__restrict__ chunks,
                                                         int *__restrict__ ctl, int64_t entry_cap,
                                                         int64_t chunk_cap)
{
    __shared__ int tot_e, tot_c;
    const int per = (T + kScanBlock - 1) / kScanBlock;
    const int t0 = threadIdx.x * per;
    const int t1 = (t0 + per) < T ? (t0 + per) : T;
    int se = 0, sc = 0, unused;
    for (int t = t0; t < t1; ++t) {
        const int h = time_total(hist, t, 0, &unused);
        se += h;
        sc += (h + kBinChunk - 1) / kBinChunk;
    }
    block_scan2(se, sc, &tot_e, &tot_c);
    const bool overflow = (int64_t)tot_e > entry_cap || (int64_t)tot_c > chunk_cap;
    for (int t = t0; t < t1; ++t) {
        const int h = time_total(hist, t, 0, &unused);
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

// lane = listed path: its (waypoint, time index) pairs into the sorted array
// at tstart[ti] + (pairs of lower sub-bins) + (rank within the sub-bin)
__global__ __launch_bounds__(256) void k_bin_emit(const HsPath *__restrict__ recs,
                                                  const int *__restrict__ ctl, int known,
                                                  const float *__restrict__ bs, int max_hs,
                                                  float dt, int tmax_steps,
                                                  const int *__restrict__ hist,
                                                  const int *__restrict__ tstart,
                                                  int *__restrict__ cursor,
                                                  BinEntry *__restrict__ entries)
{
    if (ctl[kCtlOverflow])
        return;
    const int n = ctl[kCtlListed];
    for (int slot = blockIdx.x * blockDim.x + threadIdx.x; slot < n;
         slot += gridDim.x * blockDim.x) {
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

// Binned test: one wave per (chunk, 64*kTestGroups particles).
template <bool kBits>
__device__ __forceinline__ void bin_test(const BinArgs &a, int64_t n_chunks)
{
    constexpr int NQ = kTestGroups;
    const int lane = threadIdx.x & 63;
    const int n_p = a.n_p, max_hs = a.max_hs;
    const float om = a.om;
    const int n_groups = (n_p + 63) >> 6;
    const int n_sg = (n_groups + NQ - 1) / NQ;
    const int64_t n_items = n_chunks * n_sg;
    const int64_t n_waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t item = wave_index(); item < n_items; item += n_waves) {
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
            // live ones (the reference skips b >= 1000, PUMP.cu:642)
            float b[kSlotTile], av[kSlotTile][3];
            int live = 0;
#pragma unroll
            for (int h = 0; h < kSlotTile; ++h) {
                b[h] = 1111.0f;
                av[h][0] = av[h][1] = av[h][2] = 0.0f;
                if (lane < ne && h0 + h < max_hs) {
                    const int64_t k = my_wp * max_hs + h0 + h;
                    b[h] = a.bs[k];
                    av[h][0] = a.as[k * 3 + 0];
                    av[h][1] = a.as[k * 3 + 1];
                    av[h][2] = a.as[k * 3 + 2];
                }
                live |= !(b[h] >= 1000.0f) ? (1 << h) : 0;
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

// Group form (used when the binned pairs overflow the workspace): one wave
// per (listed path, 64*NQ particles) walks the path's marked waypoints and
// writes the child words from the parent row; the path's CP is finished by
// the last of its waves (a 64-bit atomic ticket packing the finished-wave
// count and the particle count).
template <bool kBits, int NQ>
__device__ __forceinline__ void group_pass(const BinArgs &a, int64_t n_listed)
{
    const int lane = threadIdx.x & 63;
    const int n_p = a.n_p;
    const int n_groups = (n_p + 63) >> 6;
    const int n_sg = (n_groups + NQ - 1) / NQ; // waves per listed path
    const int64_t n_items = n_listed * n_sg;
    const int64_t n_waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t item = wave_index(); item < n_items; item += n_waves) {
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

// The listed paths' half-space tests: binned, or the group form on overflow.
template <bool kBits>
__global__ __launch_bounds__(256) void k_hs_pass(const BinArgs a)
{
    if (a.ctl[kCtlOverflow] == 0)
        bin_test<kBits>(a, a.ctl[kCtlChunks]);
    else
        group_pass<kBits, kHsGroups>(a, a.ctl[kCtlListed]);
}

// One wave per listed path: CP from its finished child row.
template <bool kBits>
__global__ __launch_bounds__(256) void k_cp_listed(const HsPath *__restrict__ recs,
                                                   const int *__restrict__ ctl,
                                                   const void *__restrict__ valid_dst,
                                                   int64_t dst_base, int n_p,
                                                   float *__restrict__ path_cp, int64_t cp_base)
{
    if (ctl[kCtlOverflow])
        return;
    const int lane = threadIdx.x & 63;
    const int n = ctl[kCtlListed];
    const int64_t n_waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t slot = wave_index(); slot < n; slot += n_waves) {
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
    const int64_t w_bins = a.chunk_cap * ((groups + kTestGroups - 1) / kTestGroups);
    const int64_t w_groups = (int64_t)a.n_new * ((groups + kHsGroups - 1) / kHsGroups);
    const unsigned g_hs = wave_blocks(w_bins > w_groups ? w_bins : w_groups);
    const unsigned g_cp = wave_blocks(a.n_new);
    if (bits) {
        hipLaunchKernelGGL(k_hs_pass<true>, dim3(g_hs), dim3(256), 0, s, a);
        PUMP_LAUNCHED();
        hipLaunchKernelGGL(k_cp_listed<true>, dim3(g_cp), dim3(256), 0, s, a.recs,
                           (const int *)a.ctl, (const void *)a.valid_dst, a.dst_base, a.n_p,
                           a.path_cp, a.cp_base);
    } else {
        hipLaunchKernelGGL(k_hs_pass<false>, dim3(g_hs), dim3(256), 0, s, a);
        PUMP_LAUNCHED();
        hipLaunchKernelGGL(k_cp_listed<false>, dim3(g_cp), dim3(256), 0, s, a.recs,
                           (const int *)a.ctl, (const void *)a.valid_dst, a.dst_base, a.n_p,
                           a.path_cp, a.cp_base);
    }
    PUMP_LAUNCHED();
    return PUMP_OK;
}

} // namespace pump
