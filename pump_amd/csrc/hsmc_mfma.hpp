// hsmc_mfma.hpp -- split-BF16 MFMA form of the time-binned half-space test
// (the chosen precision of the HSMC step; reference decisions of
// PUMP.cu:641-651 are kept exactly through an FP32 re-test band).
//
// For one time index, the test of particle p against stored half-space
// (a, b) is  RN(LHS*om) < b  with LHS the FP32 dot product o_p . a.  Here a
// v_mfma_f32_16x16x16_bf16 tile computes, for 16 half-spaces (rows) x 16
// particles (columns),
//     D = oh.ah + oh.al + ol.ah - bh - bl   ~=   o.a - b/om
// from bf16 hi/lo splits (x = xh + xl + e, |e| <= 2^-16 |x|; the products
// are exact in FP32).  With om > 0 and O >= max_d|o_d| (the table-wide
// bound, stored after the particle vectors),
//     |D - (o.a - b/om)| + |RN(LHS*om)/om - o.a|
//         <= 6e-5 * O * sum_d|a_d| + 2.5e-5 * |b/om|  =: margin,
// so a particle with |D| > margin is decided by the sign of D (pass iff
// D < 0) exactly as the reference decides it; the rest (|D| <= margin, and
// NaNs) are re-tested in FP32 in the reference's operation order and
// counted.  Error terms: split residuals and the dropped ol.al term, 3.02 *
// 2^-16 of |o_d a_d| per product; b/om rounding and split, 2^-24 + 2^-16 of
// |b/om|; accumulation of the 11 exact products in FP32, 10 * 2^-24 of their
// magnitude; the reference's own rounding, 4 * 2^-24 of sum_d|o_d a_d|.
//
// K layout (16 bf16 per operand row/column; lane l holds K 4*(l/16)..+3 of
// row/column l%16, the layout of the 16x16x16 bf16 MFMA):
//   particle (B):   oh0 oh1 oh2 oh0 | oh1 oh2 ol0 ol1 | ol2 -1 -1 0 | 0 0 0 0
//   half-space (A): ah0 ah1 ah2 al0 | al1 al2 ah0 ah1 | ah2 bh bl 0 | 0 0 0 0
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace pump {
namespace mfma {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kSplitK = 16;      // bf16 per particle vector
constexpr float kMarginDot = 6e-5f;
constexpr float kMarginB = 2.5e-5f;

__host__ __device__ __forceinline__ uint16_t bf16_rne(float x)
{
    uint32_t u;
    __builtin_memcpy(&u, &x, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__host__ __device__ __forceinline__ float bf16_f(uint16_t h)
{
    const uint32_t u = (uint32_t)h << 16;
    float x;
    __builtin_memcpy(&x, &u, 4);
    return x;
}

// hi/lo split: x = hi + lo + e, |e| <= 2^-16 |x| (x finite)
__host__ __device__ __forceinline__ void split(float x, uint16_t &hi, uint16_t &lo)
{
    hi = bf16_rne(x);
    lo = bf16_rne(x - bf16_f(hi));
}

constexpr uint16_t kBf16MinusOne = 0xbf80u;

// The particle vector of one offset row entry (o0, o1, o2).
__host__ __device__ __forceinline__ void particle_vector(float o0, float o1, float o2,
                                                         uint16_t *k)
{
    uint16_t h[3], l[3];
    split(o0, h[0], l[0]);
    split(o1, h[1], l[1]);
    split(o2, h[2], l[2]);
    k[0] = h[0];
    k[1] = h[1];
    k[2] = h[2];
    k[3] = h[0];
    k[4] = h[1];
    k[5] = h[2];
    k[6] = l[0];
    k[7] = l[1];
    k[8] = l[2];
    k[9] = kBf16MinusOne;
    k[10] = kBf16MinusOne;
    k[11] = 0;
    k[12] = 0;
    k[13] = 0;
    k[14] = 0;
    k[15] = 0;
}

__device__ __forceinline__ bf16x4 pack4(uint16_t a, uint16_t b, uint16_t c, uint16_t d)
{
    bf16x4 v;
    v[0] = (short)a;
    v[1] = (short)b;
    v[2] = (short)c;
    v[3] = (short)d;
    return v;
}

// The half-space operand of lane group q for row (a, b') (zeros for q = 3).
__device__ __forceinline__ bf16x4 halfspace_operand(int q, float a0, float a1, float a2,
                                                    float bp)
{
    uint16_t h0, l0, h1, l1, h2, l2, bh, bl;
    split(a0, h0, l0);
    split(a1, h1, l1);
    split(a2, h2, l2);
    split(bp, bh, bl);
    if (q == 0)
        return pack4(h0, h1, h2, l0);
    if (q == 1)
        return pack4(l1, l2, h0, h1);
    if (q == 2)
        return pack4(h2, bh, bl, 0);
    return pack4(0, 0, 0, 0);
}

} // namespace mfma
} // namespace pump
