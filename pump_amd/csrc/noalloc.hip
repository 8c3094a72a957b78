// noalloc.hip -- the batched HSMC step (propagate + CP, bit validity) for
// callers that cannot allocate device memory: code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc.  Its workspace is a
// file-scope device array of fixed size instead of a caller buffer.
//
// Each GPU the library runs on holds its own copy of the array until the
// process ends.  Every call on a GPU shares that GPU's copy, so calls that
// overlap on one GPU (two streams, two host threads) race; calls on one
// stream are ordered.  The kernels clear what must start at zero on every
// call, so the contents a previous call left behind are never read.
#include <mutex>

#include "../../include/pump_abi.h"
#include "pump_internal.hpp"

namespace {

// 64 MiB: the batched form's workspace for wavefronts of up to 566,564 paths at
// 2001 time indices (pump_propagate_workspace)
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr int kMaxDevices = 64;

} // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};

// The device address of this device's copy of the array, looked up once per
// device with that device current.  In host code the array's name is only a
// host-side stand-in, never a device address.
int scratch_address(void **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        pump::note_hip_error(e, __FILE__, __LINE__);
        return PUMP_E_HIP;
    }
    if (dev < 0 || dev >= kMaxDevices)
        return PUMP_E_RANGE;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            pump::note_hip_error(e != hipSuccess ? e : hipErrorInvalidSymbol, __FILE__, __LINE__);
            return PUMP_E_HIP;
        }
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return PUMP_OK;
}

} // namespace

extern "C" int pump_propagate_cp_bits_noalloc(
    int num_new_paths, int waypoints_count, const int64_t *d_wp_offsets, const int *d_wf_edge,
    const int *d_wf_prev, const float *d_as, const float *d_bs, const uint64_t *d_edge_hs_mask,
    const float *d_times, const float *d_topts, float dt, int max_hs,
    const float *d_offsets_time, int tmax_steps, float offset_mult, int num_mc_particles,
    const uint64_t *d_valid_src, uint64_t *d_valid_dst, int64_t dst_base, int n_p,
    float *d_path_cp, int64_t cp_base, int precision, const void *d_offsets_split,
    unsigned long long *d_band_count, unsigned long long *d_clamped, pump_stream_t stream)
{
    // the size check comes first: a wavefront whose workspace exceeds the
    // array is rejected before anything runs
    if (num_new_paths > 0 && tmax_steps > 0 &&
        pump_propagate_workspace(num_new_paths, tmax_steps) > kScratchBytes)
        return PUMP_E_CAPACITY;
    void *ws = nullptr;
    int r = scratch_address(&ws);
    if (r != PUMP_OK)
        return r;
    return pump_propagate_cp_bits(num_new_paths, waypoints_count, d_wp_offsets, d_wf_edge,
                                  d_wf_prev, d_as, d_bs, d_edge_hs_mask, d_times, d_topts, dt,
                                  max_hs, d_offsets_time, tmax_steps, offset_mult,
                                  num_mc_particles, d_valid_src, d_valid_dst, dst_base, n_p,
                                  d_path_cp, cp_base, precision, d_offsets_split, d_band_count,
                                  d_clamped, ws, kScratchBytes, stream);
}
