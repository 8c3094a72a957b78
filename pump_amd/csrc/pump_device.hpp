// pump_device.hpp -- device arithmetic for PUMP's 2pBVP and collision tests
// on gfx950.
//
// Results are bit-identical to the reference's host build (x86-64, no FP
// contraction, pow(float,int) promoted to double):
//   * contraction is disabled for every function in this header (and the
//     library is compiled with -ffp-contract=off);
//   * pow(x,2) of a float is formed as x*x in double (exact); pow(x,3) and
//     pow(x,4) as (x*x)*x and (x*x)*(x*x), the correctly rounded powers.
//     glibc's pow differs from the correctly rounded value by one double ulp
//     for about 0.2% of arguments; after the final rounding to float the
//     outputs agree except in rare last-place cases, which the parity tests
//     count and bound;
//   * per-pair quantities that do not depend on tau are formed once (same
//     operations, same order) and reused by every cost evaluation.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace pump {

constexpr int kDim = 6;
constexpr float kGravity = -9.81f;  // 2pBVP.cu:3

// ---------------------------------------------------------------------------
// cost(tau, x0, x1), 2pBVP.cu:97-107, split into the tau-independent part
// (terms 1-7 of the numerator, summed in source order) and per-term factors.
// ---------------------------------------------------------------------------
struct CostTerms {
    double s7;                                     // 12a0^2+12a2^2+12a1^2+12b0^2+12b2^2-24a1b1+12b1^2
    float f8, f9, f10, f11, f12, f13, f14, f15;    // float products multiplied by tau
    double d16, d17, d18, d20, d22, d24;           // 4*pow(v,2), times tau^2
    double f19, f21, f23;                          // 4*v0*v1 (a float, held widened), times tau^2
    double f25, f26;                               // 2*g*vz (a float, held widened), times tau^3
    float f29a, f29b, f29c, f30a, f30b, f30c;      // 12*p0*(-2*p1 + (v0+v1)*tau)
};

__host__ __device__ __forceinline__ double sq(float x) { return (double)x * (double)x; }

__host__ __device__ __forceinline__ void cost_terms(CostTerms &c, const float *a, const float *b)
{
    const float g = kGravity;
    double s = 12.0 * sq(a[0]);
    s = s + 12.0 * sq(a[2]);
    s = s + 12.0 * sq(a[1]);
    s = s + 12.0 * sq(b[0]);
    s = s + 12.0 * sq(b[2]);
    s = s - (double)(24.0f * a[1] * b[1]);
    s = s + 12.0 * sq(b[1]);
    c.s7 = s;
    c.f8 = 12.0f * a[4] * a[1];
    c.f9 = 12.0f * a[1] * b[4];
    c.f10 = 12.0f * a[3] * b[0];
    c.f11 = 12.0f * b[3] * b[0];
    c.f12 = 12.0f * a[5] * b[2];
    c.f13 = 12.0f * b[5] * b[2];
    c.f14 = 12.0f * a[4] * b[1];
    c.f15 = 12.0f * b[4] * b[1];
    c.d16 = 4.0 * sq(a[4]);
    c.d17 = 4.0 * sq(a[3]);
    c.d18 = 4.0 * sq(a[5]);
    c.f19 = (double)(4.0f * a[4] * b[4]);
    c.d20 = 4.0 * sq(b[4]);
    c.f21 = (double)(4.0f * a[3] * b[3]);
    c.d22 = 4.0 * sq(b[3]);
    c.f23 = (double)(4.0f * a[5] * b[5]);
    c.d24 = 4.0 * sq(b[5]);
    c.f25 = (double)(2.0f * g * a[5]);
    c.f26 = (double)(2.0f * g * b[5]);
    c.f29a = 12.0f * a[0];
    c.f29b = -2.0f * b[0];
    c.f29c = a[3] + b[3];
    c.f30a = 12.0f * a[2];
    c.f30b = -2.0f * b[2];
    c.f30c = a[5] + b[5];
}

// 1/x by IEEE division (any x)
struct RecipDiv {
    __host__ __device__ double operator()(double x) const { return 1.0 / x; }
};

// 1/x for 2^-100 <= |x| <= 2^400, bit-identical to RecipDiv there.  The
// compiler's FP64 division n/d is: d' and n' from v_div_scale, r = rcp(d'),
// two Newton steps r1 = fma(r, fma(-d', r, 1), r) and r2 = fma(r1,
// fma(-d', r1, 1), r1), q = n' r2, then v_div_fmas(fma(-d', q, n'), r2, q)
// and v_div_fixup.  For n = 1 and such d the scale steps return their
// operands unchanged (they act only for subnormal or near-overflow operands
// and quotients, or exponent gaps of 768 and more), v_div_fmas is a plain
// fma and v_div_fixup changes only zeros, infinities and NaNs.  What is left
// is the sequence below, with q = r2 exactly.
struct RecipNewton {
    __device__ double operator()(double x) const
    {
        const double r0 = __builtin_amdgcn_rcp(x);
        const double r1 = fma(r0, fma(-x, r0, 1.0), r0);
        const double r2 = fma(r1, fma(-x, r1, 1.0), r1);
        return fma(fma(-x, r2, 1.0), r2, r2);
    }
};

template <class Recip>
__host__ __device__ __forceinline__ float cost_eval_with(const CostTerms &c, float tau, Recip recip)
{
    const double g2 = (double)kGravity * (double)kGravity;
    const double t = (double)tau;
    const double t2 = t * t;   // exact
    const double t3 = t2 * t;  // correctly rounded
    const double t4 = t2 * t2; // correctly rounded
    double s = c.s7;
    s = s + (double)(c.f8 * tau);
    s = s + (double)(c.f9 * tau);
    s = s - (double)(c.f10 * tau);
    s = s - (double)(c.f11 * tau);
    s = s - (double)(c.f12 * tau);
    s = s - (double)(c.f13 * tau);
    s = s - (double)(c.f14 * tau);
    s = s - (double)(c.f15 * tau);
    s = s + c.d16 * t2;
    s = s + c.d17 * t2;
    s = s + c.d18 * t2;
    s = s + c.f19 * t2;
    s = s + c.d20 * t2;
    s = s + c.f21 * t2;
    s = s + c.d22 * t2;
    s = s + c.f23 * t2;
    s = s + c.d24 * t2;
    s = s + c.f25 * t3;
    s = s - c.f26 * t3;
    s = s + t4;
    s = s + g2 * t4;
    s = s + (double)(c.f29a * (c.f29b + c.f29c * tau));
    s = s + (double)(c.f30a * (c.f30b + c.f30c * tau));
    return (float)(recip(t3) * s);
}

__host__ __device__ __forceinline__ float cost_eval(const CostTerms &c, float tau)
{
    return cost_eval_with(c, tau, RecipDiv{});
}

// |RN(d / 1e-6f)| > 0.0001 exactly when |d| >= kDcostBig, for every float d
// (the quotient is monotone in d; the one transition over all finite
// positive floats lies between 0x1.b7cdfcp-34 and 0x1.b7cdfep-34).  NaN
// fails both tests.  It lets the bisection test dcost's magnitude without
// forming the quotient: toptBisection's loop condition (2pBVP.cu:85) and
// the signs it branches on (2pBVP.cu:75,79,88) depend on dcost only
// through this test and the sign of d, which the division by a positive
// dtau keeps.
constexpr float kDcostBig = 0x1.b7cdfep-34f;

// dcost, 2pBVP.cu:109-115: float central difference, dtau = 1e-6f
__host__ __device__ __forceinline__ float dcost_eval(const CostTerms &c, float tau)
{
    const float h = 0.000001f;
    float up = cost_eval(c, tau + h / 2.0f);
    float dn = cost_eval(c, tau - h / 2.0f);
    return (up - dn) / h;
}

// toptBisection, 2pBVP.cu:71-95 (halving loop capped at 400: the reference
// would not terminate once tl underflows to zero).
__host__ __device__ __forceinline__ float topt_bisection(const CostTerms &c, float tmax)
{
    float tu = tmax;
    if (dcost_eval(c, tmax) < 0.0f)
        return tmax;
    float tl = 0.01f;
    for (int k = 0; k < 400 && dcost_eval(c, tl) > 0.0f; ++k)
        tl = tl / 2.0f;
    float topt = 0.0f;
    float dcval = 1.0f;
    int itr = 0;
    while ((double)fabsf(dcval) > 0.0001 && itr < 50) {
        ++itr;
        topt = (tu + tl) / 2.0f;
        dcval = dcost_eval(c, topt);
        if (dcval > 0.0f)
            tu = topt;
        else
            tl = topt;
    }
    return topt;
}

// ---------------------------------------------------------------------------
// pathPoint(t, tau, x0, x1), 2pBVP.cu:117-131
// ---------------------------------------------------------------------------
// Unsigned 32-bit division by a fixed d >= 1 (Granlund and Montgomery,
// PLDI 1994, the round-up form with a 33-bit multiplier 2^32 + m):
// l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1, and for every
// n < 2^32, n / d = (t + ((n - t) >> s1)) >> s2 with t = mulhi(m, n),
// s1 = min(l, 1), s2 = max(l - 1, 0).  m >= 1, so m = 0 can mark "not in use".
struct DivMagic {
    uint32_t m;
    uint32_t s1, s2;
};

inline DivMagic div_magic(uint32_t d)
{
    uint32_t l = 0;
    while (l < 32 && (1ull << l) < d)
        ++l;
    const uint64_t m = ((1ull << 32) * ((1ull << l) - d)) / d + 1;
    return DivMagic{(uint32_t)m, l ? 1u : 0u, l ? l - 1 : 0u};
}

__host__ __device__ __forceinline__ uint32_t div_magic_apply(uint32_t n, const DivMagic &dm)
{
    const uint32_t t = (uint32_t)(((uint64_t)dm.m * n) >> 32);
    return (t + ((n - t) >> dm.s1)) >> dm.s2;
}

struct TauPowers {
    float tau;
    double t3;     // tau^3 (tau^2 is exact and formed where it is used)
    double r2, r3; // correctly rounded 1/tau^2, 1/tau^3
};

__host__ __device__ __forceinline__ TauPowers tau_powers(float tau)
{
    TauPowers p;
    const double t2 = (double)tau * (double)tau;
    p.tau = tau;
    p.t3 = t2 * (double)tau;
    p.r2 = 1.0 / t2;
    p.r3 = 1.0 / p.t3;
    return p;
}

// Correctly rounded a/b from r = RN(1/b) (Markstein, IBM J. Res. Dev. 34(1),
// 1990): q0 = RN(a r) is within 1.5 ulp of a/b; one residual step
// q1 = RN(q0 + (a - b q0) r), with a - b q0 exact by FMA, brings it within
// 1 ulp, and a second step from a quotient within 1 ulp and a reciprocal
// within 1/2 ulp rounds correctly.  Valid while a/b, a and b stay clear of
// overflow and subnormals (|a| < 1e300, 1e-300 < |b| < 1e300 here).
// The residuals are formed as b q - a and subtracted, which keeps the sign
// of a zero quotient (a = -0 gives -0).
__host__ __device__ __forceinline__ double div_rn(double a, double b, double r)
{
    const double q0 = a * r;
    const double e0 = fma(b, q0, -a);
    const double q1 = fma(-e0, r, q0);
    const double e1 = fma(b, q1, -a);
    return fma(-e1, r, q1);
}

// Correctly rounded a/b for a divisor b = tau^2, the exact square of a float,
// from r = RN(1/b), in one residual step.  Write b = M^2 2^(2k) with M < 2^24
// an integer, so b is a multiple of u = 2^(2k) >= 2^-48 b, and q = a/b.
//   * q0 = RN(a r) is within 1.5 ulp(q) of q.  a - b q0 is a multiple of
//     u ulp(q0) and at most 1.5 ulp(q) b < 2^49 u ulp(q) in magnitude, so the
//     FMA forms it exactly.
//   * q0 + (a - b q0) r = q + (q - q0)(b r - 1), |b r - 1| <= 2^-53: the
//     exact value rounded by the second FMA is within 1.5 * 2^-53 ulp(q) of q.
//   * A rounding midpoint m (odd multiple of ulp(q)/2) is not q: m b has at
//     least 54 significant bits, a has 53.  And |q - m| = |a - m b| / b with
//     a - m b a non-zero multiple of min(ulp(a), u ulp(q) / 2), so
//     |q - m| >= 2^-49 ulp(q).
// Hence the rounding lands on RN(q).  Same range conditions as div_rn.
__host__ __device__ __forceinline__ double div_sq(double a, double b, double r)
{
    const double q0 = a * r;
    const double e0 = fma(b, q0, -a);
    return fma(-e0, r, q0);
}

// Per-edge constants of pathPoint (2pBVP.cu:117-131): every subexpression
// that depends only on (x0, x1, tau), formed with the same operations in the
// same order as the full expression, so a waypoint evaluated from them is
// bit-identical to the expression evaluated from scratch.  Per waypoint only
// the terms in t remain (about 100 FP64 operations instead of 165).  The
// products z0 tau^3 and vz0 tau^3 are one multiplication each and are
// formed where they are used.
struct EdgePoly {
    float c3[2], c2[2], cv[2]; // x, y: cubic and quadratic position terms, velocity term
    float e[2];                // x, y: velocity term of -(2 t e)
    double kz, dz;             // z position: bracket of s^2 tau, bracket of s^3
    double bv, cz;             // z velocity: bracket of t tau, bracket of s^2
};

__host__ __device__ __forceinline__ EdgePoly edge_poly(const float *x0, const float *x1,
                                                       const TauPowers &P)
{
    EdgePoly E;
    const float tau = P.tau;
    for (int d = 0; d < 2; ++d) {
        const float p0 = x0[d], v0 = x0[d + 3], p1 = x1[d], v1 = x1[d + 3];
        const float c3 = 2.0f * p0 - 2.0f * p1 + (v0 + v1) * tau;
        const float c2 = 3.0f * p0 - 3.0f * p1 + (2.0f * v0 + v1) * tau;
        const float c = 2.0f * p0 - 2.0f * p1 + v1 * tau;
        E.c3[d] = c3;
        E.c2[d] = c2;
        E.cv[d] = c;
        E.e[d] = 3.0f * p0 - 3.0f * p1 + v1 * tau;
    }
    const double z0 = (double)x0[2], z1 = (double)x1[2];
    const double vz0 = (double)x0[5], vz1 = (double)x1[5], td = (double)tau;
    E.kz = -3.0 * z0 + 3.0 * z1 - 2.0 * vz0 * td - 1.0 * vz1 * td;
    E.dz = 2.0 * z0 - 2.0 * z1 + (double)((x0[5] + x1[5]) * tau);
    E.bv = -6.0 * z0 + 6.0 * z1 - 4.0 * vz0 * td - 2.0 * vz1 * td;
    E.cz = 6.0 * z0 - 6.0 * z1 + 3.0 * vz0 * td + 3.0 * vz1 * td;
    return E;
}

// pathPoint at time t from the edge's constants; x0 supplies p0, v0 of x, y
// and z0, vz0
__host__ __device__ __forceinline__ void path_point(float t, const TauPowers &P, const EdgePoly &E,
                                                    const float *x0, float *x)
{
    const float tau = P.tau;
    const double t2 = (double)tau * (double)tau; // exact
    const double s2 = (double)t * (double)t;
    const double s3 = s2 * (double)t;
    const double q = 3.0 * s2 - (double)(4.0f * t * tau) + t2;
    for (int d = 0; d < 2; ++d) {
        const float p0 = x0[d], v0 = x0[d + 3];
        const float lin = t * v0 + p0;
        x[d] = (float)((double)lin + div_rn(s3 * (double)E.c3[d], P.t3, P.r3) -
                       div_sq(s2 * (double)E.c2[d], t2, P.r2));
        const float f = -(2.0f * t * E.e[d]);
        x[d + 3] = (float)(div_sq((double)f, t2, P.r2) + div_rn(3.0 * s2 * (double)E.cv[d], P.t3, P.r3) +
                           div_sq((double)v0 * q, t2, P.r2));
    }
    {
        const double a = (double)(t * x0[5]) * P.t3;
        const double c = s2 * (double)tau * E.kz;
        const double d = s3 * E.dz;
        const double bz = (double)x0[2] * P.t3;
        x[2] = (float)div_rn(a + bz + c + d, P.t3, P.r3);
    }
    {
        const double b = (double)(t * tau) * E.bv;
        const double c = s2 * E.cz;
        const double av = (double)x0[5] * P.t3;
        x[5] = (float)div_rn(av + b + c, P.t3, P.r3);
    }
}

__host__ __device__ __forceinline__ void path_point(float t, const TauPowers &P, const float *x0,
                                                    const float *x1, float *x)
{
    path_point(t, P, edge_poly(x0, x1, P), x0, x);
}

// Waypoint k of an edge with W waypoints (findDiscretizedPath,
// 2pBVP.cu:53-69): endpoints are copies, interior points at (tau/(W-1))*k.
// dt = tau/(W-1) is passed in when the caller has formed it once per edge.
__host__ __device__ __forceinline__ void waypoint_dt(int k, int W, const TauPowers &P, float dt,
                                            const float *x0, const float *x1, float *x)
{
    if (k == 0) {
        for (int d = 0; d < kDim; ++d)
            x[d] = x0[d];
    } else if (k == W - 1) {
        for (int d = 0; d < kDim; ++d)
            x[d] = x1[d];
    } else {
        path_point(dt * (float)k, P, x0, x1, x);
    }
}

__host__ __device__ __forceinline__ void waypoint_dt(int k, int W, const TauPowers &P,
                                                     const EdgePoly &E, float dt, const float *x0,
                                                     const float *x1, float *x)
{
    if (k == 0) {
        for (int d = 0; d < kDim; ++d)
            x[d] = x0[d];
    } else if (k == W - 1) {
        for (int d = 0; d < kDim; ++d)
            x[d] = x1[d];
    } else {
        path_point(dt * (float)k, P, E, x0, x);
    }
}

__host__ __device__ __forceinline__ void waypoint(int k, int W, const TauPowers &P, const float *x0,
                                         const float *x1, float *x)
{
    waypoint_dt(k, W, P, P.tau / (float)(W - 1), x0, x1, x);
}

// Bounding box of the positions of every waypoint of an edge.  The position
// part of pathPoint (2pBVP.cu:117-131) is the cubic Hermite segment through
// (x0, v0) and (x1, v1) over [0, tau], which lies in the convex hull of its
// Bezier control points x0, x0 + v0 tau/3, x1 - v1 tau/3, x1.  The waypoints
// are that cubic with FP32-rounded coefficients, rounded to FP32: their
// distance from it is below 30 * 2^-24 * S, S = |p0| + |p1| + |v0 tau| +
// |v1 tau|.  The hull is formed here in FP32 (within 4 * 2^-24 * S of the
// exact one, and lo/hi round by at most 2^-24 of their magnitude), so the
// margin 1e-5 (1 + S) covers every deviation with room to spare.
__host__ __device__ __forceinline__ void edge_bbox(const float *x0, const float *x1, float tau, float *lo,
                                          float *hi)
{
    const float t3 = tau * (1.0f / 3.0f);
    for (int d = 0; d < 3; ++d) {
        const float p0 = x0[d], p3 = x1[d];
        const float w0 = x0[d + 3] * t3, w1 = x1[d + 3] * t3;
        const float p1 = p0 + w0, p2 = p3 - w1;
        const float mn = fminf(fminf(p0, p1), fminf(p2, p3));
        const float mx = fmaxf(fmaxf(p0, p1), fmaxf(p2, p3));
        const float m = 1e-5f * (1.0f + fabsf(p0) + fabsf(p3) + 3.0f * (fabsf(w0) + fabsf(w1)));
        const float l = mn - m, h = mx + m;
        const bool finite = fabsf(l) < INFINITY && fabsf(h) < INFINITY; // false for NaN too
        lo[d] = finite ? l : -INFINITY;
        hi[d] = finite ? h : INFINITY;
    }
}

// ---------------------------------------------------------------------------
// Segment vs axis-aligned box, collisionCheck.cu:65-143 (isMotionValid,
// broadphaseValidQ, motionValidQ, faceContainsProjection).
// Boxes are given by their position bounds lo[3], hi[3].
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ bool face_contains(const float *v, const float *w, float lambda, int j,
                                              const float *lo, const float *hi)
{
    for (int d = 0; d < 3; ++d) {
        float proj = v[d] + (w[d] - v[d]) * lambda;
        if (d != j && !(lo[d] <= proj && proj <= hi[d]))
            return false;
    }
    return true;
}

// true if the segment v->w collides with the box (broad phase overlap and a
// face projection inside the box)
__host__ __device__ __forceinline__ bool segment_hits_box(const float *v, const float *w,
                                                 const float *bmin, const float *bmax,
                                                 const float *lo, const float *hi)
{
    for (int d = 0; d < 3; ++d)
        if (bmax[d] <= lo[d] || hi[d] <= bmin[d])
            return false;
    for (int d = 0; d < 3; ++d) {
        float vw = w[d] - v[d];
        float lambda = (v[d] < lo[d]) ? (lo[d] - v[d]) / vw : (hi[d] - v[d]) / vw;
        if (face_contains(v, w, lambda, d, lo, hi))
            return true;
    }
    return false;
}

__host__ __device__ __forceinline__ void seg_bbox(const float *v, const float *w, float *bmin, float *bmax)
{
    for (int d = 0; d < 3; ++d) {
        if (v[d] > w[d]) {
            bmin[d] = w[d];
            bmax[d] = v[d];
        } else {
            bmin[d] = v[d];
            bmax[d] = w[d];
        }
    }
}

// ---------------------------------------------------------------------------
// Obstacle boxes staged in LDS (position bounds only, structure of arrays);
// boxes past the capacity are read from global memory by the callers.
// ---------------------------------------------------------------------------
constexpr int kMaxObsLds = 512;

template <int N> struct ObsLdsN {
    float lo[3][N];
    float hi[3][N];
};
using ObsLds = ObsLdsN<kMaxObsLds>;

template <int N>
__device__ __forceinline__ void stage_obstacles(ObsLdsN<N> &o, const float *__restrict__ obs, int n)
{
    const int m = n < N ? n : N;
    for (int q = threadIdx.x; q < m * 3; q += blockDim.x) {
        int b = q / 3, d = q - 3 * (q / 3);
        o.lo[d][b] = obs[b * 12 + d];
        o.hi[d][b] = obs[b * 12 + 6 + d];
    }
}

// The staged rows are read through LDS-typed pointers: with two generic
// pointers the compiler merges the branches into one flat load through a
// selected address.
template <int N>
__device__ __forceinline__ void box_bounds(const ObsLdsN<N> &o, const float *__restrict__ obs, int b,
                                           float *lo, float *hi)
{
    typedef __attribute__((address_space(3))) const float LdsFloat;
    if (b < N) {
        for (int d = 0; d < 3; ++d) {
            lo[d] = ((LdsFloat *)o.lo[d])[b];
            hi[d] = ((LdsFloat *)o.hi[d])[b];
        }
    } else {
        for (int d = 0; d < 3; ++d) {
            lo[d] = obs[b * 12 + d];
            hi[d] = obs[b * 12 + 6 + d];
        }
    }
}

// true if segment v->w hits any of the n_obs boxes
__device__ __forceinline__ bool segment_blocked(const float *v, const float *w, const ObsLds &o,
                                                const float *__restrict__ obs, int n_obs)
{
    float bmin[3], bmax[3];
    seg_bbox(v, w, bmin, bmax);
    for (int b = 0; b < n_obs; ++b) {
        float lo[3], hi[3];
        box_bounds(o, obs, b, lo, hi);
        if (segment_hits_box(v, w, bmin, bmax, lo, hi))
            return true;
    }
    return false;
}

// index of the edge (within a block's edge run) holding global waypoint g,
// off[0] <= g < off[ne]
__device__ __forceinline__ int find_edge(const int64_t *off, int ne, int64_t g)
{
    int lo = 0, hi = ne;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (off[mid] <= g)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// index k of the instance whose blocks [prefix[k], prefix[k+1]) hold block b
// (prefix[0] = 0 <= b < prefix[n_inst]) in a batched launch over instances
__device__ __forceinline__ int instance_of(const int64_t *__restrict__ prefix, int n_inst,
                                           int64_t b)
{
    int lo = 0, hi = n_inst;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= b)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// DPP lane move: lane i of the result reads src from the lane that kCtrl
// names (wave_shl:1 = 0x130: from lane i + 1; row_ror:n = 0x120 + n within
// each row of 16); lanes without a source keep 0.
template <int kCtrl> __device__ __forceinline__ float dpp_move(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), kCtrl, 0xf, 0xf, false));
}

__device__ __forceinline__ float read_lane(float x, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}

// min / max over the 64 lanes of a wave (all lanes active): rotations
// within each row of 16, then the four row results
__device__ __forceinline__ float wave_min(float x)
{
    x = fminf(x, dpp_move<0x121>(x));
    x = fminf(x, dpp_move<0x122>(x));
    x = fminf(x, dpp_move<0x124>(x));
    x = fminf(x, dpp_move<0x128>(x));
    return fminf(fminf(read_lane(x, 0), read_lane(x, 16)), fminf(read_lane(x, 32), read_lane(x, 48)));
}

__device__ __forceinline__ float wave_max(float x)
{
    x = fmaxf(x, dpp_move<0x121>(x));
    x = fmaxf(x, dpp_move<0x122>(x));
    x = fmaxf(x, dpp_move<0x124>(x));
    x = fmaxf(x, dpp_move<0x128>(x));
    return fmaxf(fmaxf(read_lane(x, 0), read_lane(x, 16)), fmaxf(read_lane(x, 32), read_lane(x, 48)));
}

} // namespace pump


