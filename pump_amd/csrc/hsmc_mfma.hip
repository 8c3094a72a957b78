// hsmc_mfma.hip -- particle-vector table of the split-BF16 MFMA half-space
// test (hsmc_mfma.hpp): for every (time index, particle) entry of the
// offsets table (PUMP.cu:646-647 layout, offsetsTime[(t*nMC+p)*3+d]), the
// 16 bf16 of its MFMA operand, so that the test loads a ready operand; after
// the vectors, max |o_d| over the whole table (float), which sets the margin.
#include "hsmc_mfma.hpp"
#include "pump_internal.hpp"

namespace pump {
namespace {

__global__ void k_zero_max(unsigned *p)
{
    if (threadIdx.x == 0)
        p[0] = 0u;
}

__global__ __launch_bounds__(256) void k_split_offsets(const float *__restrict__ off, int64_t n,
                                                       uint16_t *__restrict__ out,
                                                       unsigned *__restrict__ max_bits)
{
    float m = 0.0f;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        m = fmaxf(m, fmaxf(fmaxf(fabsf(off[i * 3 + 0]), fabsf(off[i * 3 + 1])),
                           fabsf(off[i * 3 + 2])));
        uint16_t k[mfma::kSplitK];
        mfma::particle_vector(off[i * 3 + 0], off[i * 3 + 1], off[i * 3 + 2], k);
        uint4 *o = reinterpret_cast<uint4 *>(out + i * mfma::kSplitK);
        o[0] = make_uint4(k[0] | ((uint32_t)k[1] << 16), k[2] | ((uint32_t)k[3] << 16),
                          k[4] | ((uint32_t)k[5] << 16), k[6] | ((uint32_t)k[7] << 16));
        o[1] = make_uint4(k[8] | ((uint32_t)k[9] << 16), k[10] | ((uint32_t)k[11] << 16),
                          k[12] | ((uint32_t)k[13] << 16), k[14] | ((uint32_t)k[15] << 16));
    }
    // one atomic per workgroup; non-negative floats order as their bits
    __shared__ float wm[4];
    for (int off2 = 32; off2 > 0; off2 >>= 1)
        m = fmaxf(m, __shfl_xor(m, off2, 64));
    if ((threadIdx.x & 63) == 0)
        wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        atomicMax(max_bits, __float_as_uint(fmaxf(fmaxf(wm[0], wm[1]), fmaxf(wm[2], wm[3]))));
}

} // namespace
} // namespace pump

using namespace pump;

extern "C" size_t pump_split_offsets_bytes(int tmax_steps, int n_mc)
{
    if (tmax_steps <= 0 || n_mc <= 0)
        return 0;
    return (size_t)tmax_steps * (size_t)n_mc * mfma::kSplitK * sizeof(uint16_t) + 16;
}

extern "C" int pump_split_offsets(const float *d_offsets_time, int tmax_steps, int n_mc,
                                  void *d_split, pump_stream_t stream)
{
    PUMP_REQUIRE(d_offsets_time && d_split && tmax_steps > 0 && n_mc > 0);
    PUMP_REQUIRE((((uintptr_t)d_split) & 15) == 0);
    const int64_t n = (int64_t)tmax_steps * n_mc;
    unsigned *max_bits = reinterpret_cast<unsigned *>((uint16_t *)d_split + n * mfma::kSplitK);
    hipLaunchKernelGGL(k_zero_max, dim3(1), dim3(64), 0, as_stream(stream), max_bits);
    PUMP_LAUNCHED();
    hipLaunchKernelGGL(k_split_offsets, dim3(grid_for(n, 256, 2048)), dim3(256), 0,
                       as_stream(stream), d_offsets_time, n, (uint16_t *)d_split, max_bits);
    PUMP_LAUNCHED();
    return PUMP_OK;
}
