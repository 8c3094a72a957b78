// roadmap.hip -- the roadmap glue of mainPUMP.cu:257-392 on the device:
//   * rn = k-th smallest pair cost (thrust::sort + index, mainPUMP.cu:303-304)
//     as a 3-pass radix select over order-preserving uint32 keys
//     (11 + 11 + 10 bits), histograms in LDS, no host round trips between
//     passes;
//   * edge numbering (copts < rn, row-major i then j, mainPUMP.cu:327-346):
//     per-row counts, an exclusive scan, then an order-preserving per-row
//     compaction with wave ballots;
//   * maxNNSize and the padded neighbour tables (mainPUMP.cu:362-392);
//   * waypoint offsets for a fixed or per-edge waypoint count.
// All of this is HBM-streaming integer/compare work over the pair arrays.
#include "pump_internal.hpp"

namespace pump {
namespace {

constexpr int kSelBins = 2048;

struct SelState {
    uint32_t prefix;
    uint32_t mask;
    unsigned long long k; // rank still to skip inside the current prefix
    float result;
    uint32_t pad;
};

__device__ __forceinline__ uint32_t order_key(float f)
{
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float key_float(uint32_t k)
{
    uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(u);
}

__global__ __launch_bounds__(256) void k_sel_init(SelState *st, unsigned long long *hist,
                                                  unsigned long long k)
{
    for (int b = threadIdx.x; b < kSelBins; b += blockDim.x)
        hist[b] = 0;
    if (threadIdx.x == 0) {
        st->prefix = 0;
        st->mask = 0;
        st->k = k;
        st->result = 0.0f;
    }
}

__global__ __launch_bounds__(256) void k_sel_hist(const float *__restrict__ keys, int64_t n,
                                                  const SelState *__restrict__ st, int shift,
                                                  uint32_t bin_mask,
                                                  unsigned long long *__restrict__ hist)
{
    __shared__ uint32_t h[kSelBins];
    for (int b = threadIdx.x; b < kSelBins; b += blockDim.x)
        h[b] = 0;
    __syncthreads();
    const uint32_t prefix = st->prefix, mask = st->mask;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // 4-wide unrolled stream
    for (; i + 3 * stride < n; i += 4 * stride) {
        uint32_t u0 = order_key(keys[i]);
        uint32_t u1 = order_key(keys[i + stride]);
        uint32_t u2 = order_key(keys[i + 2 * stride]);
        uint32_t u3 = order_key(keys[i + 3 * stride]);
        if ((u0 & mask) == prefix)
            atomicAdd(&h[(u0 >> shift) & bin_mask], 1u);
        if ((u1 & mask) == prefix)
            atomicAdd(&h[(u1 >> shift) & bin_mask], 1u);
        if ((u2 & mask) == prefix)
            atomicAdd(&h[(u2 >> shift) & bin_mask], 1u);
        if ((u3 & mask) == prefix)
            atomicAdd(&h[(u3 >> shift) & bin_mask], 1u);
    }
    for (; i < n; i += stride) {
        uint32_t u = order_key(keys[i]);
        if ((u & mask) == prefix)
            atomicAdd(&h[(u >> shift) & bin_mask], 1u);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < kSelBins; b += blockDim.x)
        if (h[b])
            atomicAdd(&hist[b], (unsigned long long)h[b]);
}

// one block: find the bin holding rank k, narrow the prefix, clear the
// histogram for the next pass
__global__ __launch_bounds__(256) void k_sel_pick(SelState *st, unsigned long long *hist,
                                                  int shift, uint32_t bin_mask, int last)
{
    __shared__ unsigned long long part[256];
    __shared__ int found_bin;
    __shared__ unsigned long long found_before;
    const int per = kSelBins / 256; // 8 bins per thread
    unsigned long long local[8];
    unsigned long long sum = 0;
    for (int q = 0; q < per; ++q) {
        local[q] = hist[threadIdx.x * per + q];
        sum += local[q];
    }
    part[threadIdx.x] = sum;
    __syncthreads();
    // exclusive scan (Hillis-Steele on 256 entries)
    for (int off = 1; off < 256; off <<= 1) {
        unsigned long long v = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    unsigned long long before = part[threadIdx.x] - sum;
    const unsigned long long k = st->k;
    if (threadIdx.x == 0)
        found_bin = -1;
    __syncthreads();
    if (k >= before && k < before + sum) {
        unsigned long long c = before;
        for (int q = 0; q < per; ++q) {
            if (k < c + local[q]) {
                found_bin = threadIdx.x * per + q;
                found_before = c;
                break;
            }
            c += local[q];
        }
    }
    __syncthreads();
    for (int q = 0; q < per; ++q)
        hist[threadIdx.x * per + q] = 0;
    if (threadIdx.x == 0 && found_bin >= 0) {
        st->prefix |= ((uint32_t)found_bin & bin_mask) << shift;
        st->mask |= bin_mask << shift;
        st->k = k - found_before;
        if (last)
            st->result = key_float(st->prefix);
    }
}

// ---- exclusive scan of int64 (in place, total appended at a[n]) ---------
constexpr int kScanTile = 2048; // 1024 threads x 2

__device__ __forceinline__ int64_t block_exclusive_scan_1024(int64_t v, int64_t *sh,
                                                             int64_t &total)
{
    // v: this thread's value; returns exclusive prefix within the block
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int64_t x = v;
    for (int off = 1; off < 64; off <<= 1) {
        int64_t y = __shfl_up(x, off, 64);
        if (lane >= off)
            x += y;
    }
    if (lane == 63)
        sh[wid] = x;
    __syncthreads();
    if (wid == 0) {
        int64_t w = lane < 16 ? sh[lane] : 0;
        for (int off = 1; off < 16; off <<= 1) {
            int64_t y = __shfl_up(w, off, 64);
            if (lane >= off)
                w += y;
        }
        if (lane < 16)
            sh[16 + lane] = w;
    }
    __syncthreads();
    int64_t wave_before = wid > 0 ? sh[16 + wid - 1] : 0;
    total = sh[16 + 15];
    __syncthreads();
    return wave_before + x - v;
}

__global__ __launch_bounds__(1024) void k_scan_tiles(const int64_t *__restrict__ a, int64_t n,
                                                     int64_t *__restrict__ tile_sums)
{
    __shared__ int64_t sh[32];
    int64_t base = (int64_t)blockIdx.x * kScanTile + 2 * threadIdx.x;
    int64_t v = (base < n ? a[base] : 0) + (base + 1 < n ? a[base + 1] : 0);
    int64_t total;
    block_exclusive_scan_1024(v, sh, total);
    if (threadIdx.x == 0)
        tile_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_scan_sums(int64_t *__restrict__ tile_sums, int64_t nt)
{
    __shared__ int64_t sh[32];
    int64_t carry = 0;
    for (int64_t b = 0; b < nt; b += 1024) {
        int64_t idx = b + threadIdx.x;
        int64_t v = idx < nt ? tile_sums[idx] : 0;
        int64_t total;
        int64_t ex = block_exclusive_scan_1024(v, sh, total);
        if (idx < nt)
            tile_sums[idx] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0)
        tile_sums[nt] = carry;
}

// the whole scan in one workgroup, tile after tile with a running carry
// (short arrays: one launch instead of three)
__global__ __launch_bounds__(1024) void k_scan_small(int64_t *__restrict__ a, int64_t n)
{
    __shared__ int64_t sh[32];
    int64_t carry = 0;
    for (int64_t t0 = 0; t0 < n; t0 += kScanTile) {
        const int64_t base = t0 + 2 * threadIdx.x;
        const int64_t v0 = base < n ? a[base] : 0;
        const int64_t v1 = base + 1 < n ? a[base + 1] : 0;
        int64_t total;
        const int64_t ex = block_exclusive_scan_1024(v0 + v1, sh, total);
        if (base < n)
            a[base] = carry + ex;
        if (base + 1 < n)
            a[base + 1] = carry + ex + v0;
        carry += total;
    }
    if (threadIdx.x == 0)
        a[n] = carry;
}

__global__ __launch_bounds__(1024) void k_scan_apply(int64_t *__restrict__ a, int64_t n,
                                                     const int64_t *__restrict__ tile_sums,
                                                     int64_t nt)
{
    __shared__ int64_t sh[32];
    int64_t base = (int64_t)blockIdx.x * kScanTile + 2 * threadIdx.x;
    int64_t v0 = base < n ? a[base] : 0;
    int64_t v1 = base + 1 < n ? a[base + 1] : 0;
    int64_t total;
    int64_t ex = block_exclusive_scan_1024(v0 + v1, sh, total);
    int64_t off = tile_sums[blockIdx.x] + ex;
    if (base < n)
        a[base] = off;
    if (base + 1 < n)
        a[base + 1] = off + v0;
    if (blockIdx.x == 0 && threadIdx.x == 0)
        a[n] = tile_sums[nt];
}

// ---- edges ----------------------------------------------------------------
__global__ __launch_bounds__(256) void k_count_rows(const float *__restrict__ copts, int num,
                                                    float rn, int64_t *__restrict__ counts)
{
    __shared__ int sh[4];
    const int i = blockIdx.x;
    const float *row = copts + (int64_t)i * (num - 1);
    int c = 0;
    for (int k = threadIdx.x; k < num - 1; k += blockDim.x)
        c += row[k] < rn;
    for (int off = 32; off > 0; off >>= 1)
        c += __shfl_down(c, off, 64);
    if ((threadIdx.x & 63) == 0)
        sh[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0)
        counts[i] = (int64_t)(sh[0] + sh[1] + sh[2] + sh[3]);
}

__global__ __launch_bounds__(256) void k_fill_rows(const float *__restrict__ copts,
                                                   const float *__restrict__ topts, int num,
                                                   float rn, const int64_t *__restrict__ row_start,
                                                   int *__restrict__ edge_src,
                                                   int *__restrict__ edge_dst,
                                                   float *__restrict__ copts_edge,
                                                   float *__restrict__ topts_edge)
{
    __shared__ int wave_cnt[4];
    const int i = blockIdx.x;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int64_t rbase = (int64_t)i * (num - 1);
    int64_t out = row_start[i];
    for (int k0 = 0; k0 < num - 1; k0 += 256) {
        int k = k0 + threadIdx.x;
        float c = 0.0f;
        bool keep = false;
        if (k < num - 1) {
            c = copts[rbase + k];
            keep = c < rn;
        }
        unsigned long long m = __ballot(keep);
        int before_in_wave = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0)
            wave_cnt[wid] = __popcll(m);
        __syncthreads();
        int wave_before = 0;
        for (int w = 0; w < wid; ++w)
            wave_before += wave_cnt[w];
        int chunk_total = wave_cnt[0] + wave_cnt[1] + wave_cnt[2] + wave_cnt[3];
        if (keep) {
            int64_t e = out + wave_before + before_in_wave;
            int j = k + (k >= i ? 1 : 0);
            edge_src[e] = i;
            edge_dst[e] = j;
            copts_edge[e] = c;
            topts_edge[e] = topts[rbase + k];
        }
        out += chunk_total;
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_zero_i32(int *__restrict__ a, int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        a[i] = 0;
}

__global__ __launch_bounds__(256) void k_fill_i32(int *__restrict__ a, int64_t n, int v)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        a[i] = v;
}

__global__ __launch_bounds__(256) void k_in_degree(const int *__restrict__ edge_dst, int64_t n,
                                                   int *__restrict__ indeg)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride)
        atomicAdd(&indeg[edge_dst[e]], 1);
}

__global__ __launch_bounds__(256) void k_max_deg(const int64_t *__restrict__ row_start,
                                                 const int *__restrict__ indeg, int num,
                                                 int *__restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int m = 0;
    if (i < num) {
        int go = (int)(row_start[i + 1] - row_start[i]);
        m = go > indeg[i] ? go : indeg[i];
    }
    for (int off = 32; off > 0; off >>= 1) {
        int o = __shfl_down(m, off, 64);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0)
        atomicMax(out, m);
}

__global__ __launch_bounds__(256) void k_nn_go(const int64_t *__restrict__ row_start,
                                               const int *__restrict__ edge_dst, int num,
                                               int max_nn, int *__restrict__ go)
{
    const int64_t n = (int64_t)num * max_nn;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
        int i = (int)(q / max_nn);
        int nn = (int)(q - (int64_t)i * max_nn);
        int64_t deg = row_start[i + 1] - row_start[i];
        go[q] = nn < deg ? edge_dst[row_start[i] + nn] : -1;
    }
}

// unordered in-lists, then each node's list sorted by source (sources of one
// destination are distinct, so rank = number of smaller sources)
__global__ __launch_bounds__(256) void k_come_scatter(const int *__restrict__ edge_src,
                                                      const int *__restrict__ edge_dst,
                                                      const float *__restrict__ copts_edge,
                                                      int64_t n, int max_nn,
                                                      int *__restrict__ fill,
                                                      int *__restrict__ come,
                                                      float *__restrict__ dist)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
        int j = edge_dst[e];
        int slot = atomicAdd(&fill[j], 1);
        come[(int64_t)j * max_nn + slot] = edge_src[e];
        dist[(int64_t)j * max_nn + slot] = copts_edge[e];
    }
}

__global__ __launch_bounds__(256) void k_come_sort(const int *__restrict__ fill, int max_nn,
                                                   int *__restrict__ come,
                                                   float *__restrict__ dist)
{
    extern __shared__ unsigned char smem[];
    int *src = reinterpret_cast<int *>(smem);
    float *dd = reinterpret_cast<float *>(src + max_nn);
    const int j = blockIdx.x;
    const int n = fill[j];
    int *row = come + (int64_t)j * max_nn;
    float *drow = dist + (int64_t)j * max_nn;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        src[q] = row[q];
        dd[q] = drow[q];
    }
    __syncthreads();
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        int s = src[q];
        int r = 0;
        for (int m = 0; m < n; ++m)
            r += src[m] < s;
        row[r] = s;
        drow[r] = dd[q];
    }
}

__global__ __launch_bounds__(256) void k_nn_idxs(const int *__restrict__ edge_src,
                                                 const int *__restrict__ edge_dst, int64_t n,
                                                 int num, int *__restrict__ nn_idxs)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride)
        nn_idxs[(int64_t)edge_dst[e] * num + edge_src[e]] = (int)e;
}

__global__ __launch_bounds__(256) void k_wcount(const float *__restrict__ topts_edge, int64_t n,
                                                int w_fixed, float wp_dt, int w_cap,
                                                int64_t *__restrict__ w)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
        int c;
        if (w_fixed > 0) {
            c = w_fixed;
        } else {
            c = (int)ceilf(topts_edge[e] / wp_dt);
            c = c < 2 ? 2 : c;
            c = c > w_cap ? w_cap : c;
        }
        w[e] = c;
    }
}

} // namespace

// in-place exclusive scan of a[0..n), total into a[n]
int scan_exclusive_i64(int64_t *d_a, int64_t n, void *d_ws, size_t ws_bytes, hipStream_t s)
{
    const int64_t nt = (n + kScanTile - 1) / kScanTile;
    if (n >= 0 && nt <= 8) {
        hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(1024), 0, s, d_a, n);
        PUMP_LAUNCHED();
        return PUMP_OK;
    }
    if (ws_bytes < (size_t)(nt + 1) * sizeof(int64_t) || !d_ws)
        return PUMP_E_CAPACITY;
    int64_t *sums = reinterpret_cast<int64_t *>(d_ws);
    if (nt > 0) {
        hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)nt), dim3(1024), 0, s, d_a, n, sums);
        PUMP_LAUNCHED();
    }
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, sums, nt);
    PUMP_LAUNCHED();
    if (nt > 0) {
        hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nt), dim3(1024), 0, s, d_a, n, sums, nt);
        PUMP_LAUNCHED();
    } else {
        PUMP_HIP(hipMemsetAsync(d_a, 0, sizeof(int64_t), s));
    }
    return PUMP_OK;
}

} // namespace pump

using namespace pump;

extern "C" int64_t pump_rn_index(int64_t num, float rn_perc)
{
    float prod = (float)((num - 1) * num) * rn_perc;
    return (int64_t)prod;
}

extern "C" size_t pump_select_rn_workspace(int64_t)
{
    return sizeof(SelState) + 256 + kSelBins * sizeof(unsigned long long);
}

extern "C" int pump_select_rn(const float *d_copts, int64_t n_pairs, int64_t rn_idx, float *rn_host,
                              void *d_ws, size_t ws_bytes, pump_stream_t stream)
{
    PUMP_REQUIRE(d_copts && rn_host && n_pairs > 0 && rn_idx >= 0 && rn_idx < n_pairs);
    if (!d_ws || ws_bytes < pump_select_rn_workspace(n_pairs))
        return PUMP_E_CAPACITY;
    hipStream_t s = as_stream(stream);
    SelState *st = reinterpret_cast<SelState *>(d_ws);
    unsigned long long *hist =
        reinterpret_cast<unsigned long long *>(reinterpret_cast<char *>(d_ws) + sizeof(SelState) + 256);
    hipLaunchKernelGGL(k_sel_init, dim3(1), dim3(256), 0, s, st, hist,
                       (unsigned long long)rn_idx);
    PUMP_LAUNCHED();
    const int shifts[3] = {21, 10, 0};
    const uint32_t masks[3] = {0x7ffu, 0x7ffu, 0x3ffu};
    int dev = 0, cus = 256;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    unsigned blocks = grid_for(n_pairs / 4 + 1, 256, (int64_t)cus * 8);
    for (int p = 0; p < 3; ++p) {
        hipLaunchKernelGGL(k_sel_hist, dim3(blocks), dim3(256), 0, s, d_copts, n_pairs, st,
                           shifts[p], masks[p], hist);
        PUMP_LAUNCHED();
        hipLaunchKernelGGL(k_sel_pick, dim3(1), dim3(256), 0, s, st, hist, shifts[p], masks[p],
                           p == 2 ? 1 : 0);
        PUMP_LAUNCHED();
    }
    float rn = 0.0f;
    PUMP_HIP(hipMemcpyAsync(&rn, &st->result, sizeof(float), hipMemcpyDeviceToHost, s));
    PUMP_HIP(hipStreamSynchronize(s));
    *rn_host = rn;
    return PUMP_OK;
}

extern "C" size_t pump_scan_workspace(int64_t n)
{
    return (size_t)((n + kScanTile - 1) / kScanTile + 2) * sizeof(int64_t);
}

extern "C" int pump_count_edges(const float *d_copts, int num, float rn, int64_t *d_row_start,
                                int64_t *n_edges_host, void *d_ws, size_t ws_bytes,
                                pump_stream_t stream)
{
    PUMP_REQUIRE(d_copts && d_row_start && n_edges_host && num >= 2);
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_count_rows, dim3(num), dim3(256), 0, s, d_copts, num, rn, d_row_start);
    PUMP_LAUNCHED();
    int r = scan_exclusive_i64(d_row_start, num, d_ws, ws_bytes, s);
    if (r != PUMP_OK)
        return r;
    int64_t total = 0;
    PUMP_HIP(hipMemcpyAsync(&total, d_row_start + num, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    PUMP_HIP(hipStreamSynchronize(s));
    *n_edges_host = total;
    return PUMP_OK;
}

extern "C" int pump_fill_edges(const float *d_copts, const float *d_topts, int num, float rn,
                               const int64_t *d_row_start, int *d_edge_src, int *d_edge_dst,
                               float *d_copts_edge, float *d_topts_edge, pump_stream_t stream)
{
    PUMP_REQUIRE(d_copts && d_topts && d_row_start && d_edge_src && d_edge_dst && d_copts_edge &&
                 d_topts_edge && num >= 2);
    hipLaunchKernelGGL(k_fill_rows, dim3(num), dim3(256), 0, as_stream(stream), d_copts, d_topts,
                       num, rn, d_row_start, d_edge_src, d_edge_dst, d_copts_edge, d_topts_edge);
    PUMP_LAUNCHED();
    return PUMP_OK;
}

extern "C" int pump_max_nn_size(const int64_t *d_row_start, const int *d_edge_dst, int64_t n_edges,
                                int num, int *d_in_degree, int *max_nn_host, pump_stream_t stream)
{
    PUMP_REQUIRE(d_row_start && d_in_degree && max_nn_host && num >= 1);
    PUMP_REQUIRE(n_edges == 0 || d_edge_dst);
    hipStream_t s = as_stream(stream);
    int *d_max = nullptr;
    PUMP_HIP(hipMalloc((void **)&d_max, sizeof(int)));
    hipLaunchKernelGGL(k_zero_i32, dim3(grid_for(num + 1, 256)), dim3(256), 0, s, d_in_degree,
                       (int64_t)num);
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, d_max, (int64_t)1);
    if (n_edges > 0)
        hipLaunchKernelGGL(k_in_degree, dim3(grid_for(n_edges, 256)), dim3(256), 0, s, d_edge_dst,
                           n_edges, d_in_degree);
    hipLaunchKernelGGL(k_max_deg, dim3((num + 255) / 256), dim3(256), 0, s, d_row_start,
                       d_in_degree, num, d_max);
    hipError_t le = hipGetLastError();
    int m = 0;
    hipError_t ce = hipMemcpyAsync(&m, d_max, sizeof(int), hipMemcpyDeviceToHost, s);
    hipError_t se = hipStreamSynchronize(s);
    hipFree(d_max);
    PUMP_HIP(le);
    PUMP_HIP(ce);
    PUMP_HIP(se);
    *max_nn_host = m;
    return PUMP_OK;
}

extern "C" int pump_nn_tables(const int64_t *d_row_start, const int *d_edge_src,
                              const int *d_edge_dst, const float *d_copts_edge, int64_t n_edges,
                              int num, int max_nn, int *d_nn_go_edges, int *d_nn_come_edges,
                              float *d_distances_come, int *d_nn_idxs, pump_stream_t stream)
{
    PUMP_REQUIRE(d_row_start && num >= 1 && max_nn >= 0);
    hipStream_t s = as_stream(stream);
    const int64_t tab = (int64_t)num * max_nn;
    if (d_nn_go_edges && tab > 0) {
        hipLaunchKernelGGL(k_nn_go, dim3(grid_for(tab, 256)), dim3(256), 0, s, d_row_start,
                           d_edge_dst, num, max_nn, d_nn_go_edges);
        PUMP_LAUNCHED();
    }
    if (d_nn_come_edges && d_distances_come && tab > 0) {
        int *fill = nullptr;
        PUMP_HIP(hipMalloc((void **)&fill, sizeof(int) * (size_t)num));
        hipLaunchKernelGGL(k_zero_i32, dim3(grid_for(num, 256)), dim3(256), 0, s, fill,
                           (int64_t)num);
        hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(tab, 256)), dim3(256), 0, s,
                           d_nn_come_edges, tab, -1);
        hipLaunchKernelGGL(k_zero_i32, dim3(grid_for(tab, 256)), dim3(256), 0, s,
                           reinterpret_cast<int *>(d_distances_come), tab);
        if (n_edges > 0) {
            hipLaunchKernelGGL(k_come_scatter, dim3(grid_for(n_edges, 256)), dim3(256), 0, s,
                               d_edge_src, d_edge_dst, d_copts_edge, n_edges, max_nn, fill,
                               d_nn_come_edges, d_distances_come);
            hipLaunchKernelGGL(k_come_sort, dim3(num), dim3(256),
                               (size_t)max_nn * (sizeof(int) + sizeof(float)), s, fill, max_nn,
                               d_nn_come_edges, d_distances_come);
        }
        hipError_t le = hipGetLastError();
        hipError_t se = hipStreamSynchronize(s);
        hipFree(fill);
        PUMP_HIP(le);
        PUMP_HIP(se);
    }
    if (d_nn_idxs) {
        const int64_t nn2 = (int64_t)num * num;
        hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(nn2, 256)), dim3(256), 0, s, d_nn_idxs, nn2,
                           -3);
        if (n_edges > 0)
            hipLaunchKernelGGL(k_nn_idxs, dim3(grid_for(n_edges, 256)), dim3(256), 0, s,
                               d_edge_src, d_edge_dst, n_edges, num, d_nn_idxs);
        PUMP_LAUNCHED();
    }
    return PUMP_OK;
}

extern "C" int pump_waypoint_offsets(const float *d_topts_edge, int64_t n_edges, int w_fixed,
                                     float wp_dt, int w_cap, int64_t *d_wp_offsets,
                                     int64_t *total_host, void *d_ws, size_t ws_bytes,
                                     pump_stream_t stream)
{
    PUMP_REQUIRE(d_wp_offsets && total_host && n_edges >= 0);
    PUMP_REQUIRE(w_fixed >= 2 || (w_fixed <= 0 && w_cap >= 2 && wp_dt > 0.0f && d_topts_edge));
    PUMP_REQUIRE(w_fixed < kMaxWaypointsPerEdge && w_cap < kMaxWaypointsPerEdge);
    hipStream_t s = as_stream(stream);
    if (n_edges > 0) {
        hipLaunchKernelGGL(k_wcount, dim3(grid_for(n_edges, 256)), dim3(256), 0, s, d_topts_edge,
                           n_edges, w_fixed, wp_dt, w_cap, d_wp_offsets);
        PUMP_LAUNCHED();
    }
    int r = scan_exclusive_i64(d_wp_offsets, n_edges, d_ws, ws_bytes, s);
    if (r != PUMP_OK)
        return r;
    int64_t total = 0;
    PUMP_HIP(hipMemcpyAsync(&total, d_wp_offsets + n_edges, sizeof(int64_t),
                            hipMemcpyDeviceToHost, s));
    PUMP_HIP(hipStreamSynchronize(s));
    *total_host = total;
    return PUMP_OK;
}
