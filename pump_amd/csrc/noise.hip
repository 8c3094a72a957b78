// noise.hip -- the Monte Carlo noise tables generated on the device.
//
// The reference reads its offset tables xcomb / xcombTime from a file that is
// not part of its repository (hardCoded.cuh:12-19); pump_make_noise builds
// them on the host and pump_make_noise_device here on the GPU, from the same
// definition: per axis the closed-loop double integrator
//   x_{t+1} = A_cl x_t + L z_t,  u = -kp p - kd v,  Q = q [[dt^3/3, dt^2/2], [dt^2/2, dt]],
// L the FP64 Cholesky factor of the 6x6 Q (one thread), z_t ~ N(0, I6) by
// Box-Muller from Philox4x32-10 draws (Salmon et al., SC'11) keyed by the
// seed, counter (particle, step, stream 3, block j), x_0 = 0, positions
// stored at unit scale.  One thread per particle runs its recurrence in
// FP64; the position covariance Sigma_{t+1} = A Sigma_t A^T + Q is
// propagated in FP32 (one thread).  Every statement keeps the host
// generator's order, without contraction, so the Cholesky factor and the
// covariance are bit-identical to the host's and the tables differ from the
// host's only where the device's FP64 log/sin/cos round differently from the
// host C library's.
#include "../../include/pump_host.h"
#include "pump_internal.hpp"

#pragma clang fp contract(off)

namespace pump {
namespace {

constexpr uint32_t kPhiloxM0 = 0xD2511F53u, kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u, kPhiloxW1 = 0xBB67AE85u;
constexpr uint32_t kStreamNoise = 3;

__device__ __forceinline__ void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                       uint32_t k0, uint32_t k1, uint32_t out[4])
{
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += kPhiloxW0;
            k1 += kPhiloxW1;
        }
        const uint32_t hi0 = __umulhi(kPhiloxM0, c0), lo0 = kPhiloxM0 * c0;
        const uint32_t hi1 = __umulhi(kPhiloxM1, c2), lo1 = kPhiloxM1 * c2;
        const uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

struct NoiseArgs {
    int n_mc, T;
    double dt, kp, kd, q;
    uint32_t key0, key1;
};

// the closed-loop transition A and the process covariance Q (6x6, FP64)
__device__ __forceinline__ void noise_system(const NoiseArgs &a, double *A, double *Q)
{
    const double dt = a.dt;
    const double a11 = 1 - a.kp * dt * dt / 2, a12 = dt - a.kd * dt * dt / 2;
    const double a21 = -a.kp * dt, a22 = 1 - a.kd * dt;
    for (int k = 0; k < 36; ++k) {
        A[k] = 0;
        Q[k] = 0;
    }
    for (int d = 0; d < 3; ++d) {
        A[d * 6 + d] = a11;
        A[d * 6 + d + 3] = a12;
        A[(d + 3) * 6 + d] = a21;
        A[(d + 3) * 6 + d + 3] = a22;
        Q[d * 6 + d] = a.q * dt * dt * dt / 3;
        Q[d * 6 + d + 3] = a.q * dt * dt / 2;
        Q[(d + 3) * 6 + d] = a.q * dt * dt / 2;
        Q[(d + 3) * 6 + d + 3] = a.q * dt;
    }
}

// L with L L^T = Q (FP64, one thread); a non-positive pivot writes NaN
__global__ void k_noise_cholesky(NoiseArgs a, double *__restrict__ L)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    double A[36], Q[36], M[36];
    noise_system(a, A, Q);
    for (int k = 0; k < 36; ++k)
        M[k] = 0;
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = Q[i * 6 + j];
            for (int k = 0; k < j; ++k)
                s -= M[i * 6 + k] * M[j * 6 + k];
            if (i == j)
                M[i * 6 + i] = s > 0 ? sqrt(s) : __longlong_as_double(0x7ff8000000000000ll);
            else
                M[i * 6 + j] = s / M[j * 6 + j];
        }
    for (int k = 0; k < 36; ++k)
        L[k] = M[k];
}

// One particle per thread: its offsets at every step, both table layouts.
__global__ __launch_bounds__(64) void k_noise_particles(NoiseArgs a, const double *__restrict__ Lg,
                                                        float *__restrict__ offsets_time,
                                                        float *__restrict__ xcomb)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n_mc)
        return;
    double A[36], Q[36], L[36];
    noise_system(a, A, Q);
    for (int k = 0; k < 36; ++k)
        L[k] = Lg[k];
    double x[6] = {0, 0, 0, 0, 0, 0};
    for (int t = 0; t < a.T; ++t) {
        for (int d = 0; d < 3; ++d) {
            const float v = (float)x[d];
            if (offsets_time)
                offsets_time[((int64_t)t * a.n_mc + p) * 3 + d] = v;
            if (xcomb)
                xcomb[((int64_t)p * a.T + t) * 3 + d] = v;
        }
        double z[6];
        for (int j = 0; j < 3; ++j) {
            uint32_t out[4];
            philox((uint32_t)p, (uint32_t)t, kStreamNoise, (uint32_t)j, a.key0, a.key1, out);
            const double u1 = ((double)out[0] + 0.5) * (1.0 / 4294967296.0);
            const double u2 = ((double)out[1] + 0.5) * (1.0 / 4294967296.0);
            const double r = sqrt(-2.0 * log(u1));
            z[2 * j] = r * cos(6.283185307179586 * u2);
            z[2 * j + 1] = r * sin(6.283185307179586 * u2);
        }
        double xn[6];
        for (int i = 0; i < 6; ++i) {
            double s = 0;
            for (int k = 0; k < 6; ++k)
                s += A[i * 6 + k] * x[k];
            for (int k = 0; k <= i; ++k)
                s += L[i * 6 + k] * z[k];
            xn[i] = s;
        }
        for (int i = 0; i < 6; ++i)
            x[i] = xn[i];
    }
}

// Position covariance, FP32: Sigma_{t+1} = A Sigma_t A^T + Q, Sigma_0 = 0
__global__ void k_noise_sigma(NoiseArgs a, float *__restrict__ sigma_pos)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    double A[36], Q[36];
    noise_system(a, A, Q);
    float S[36], Af[36], Qf[36], AS[36], N[36];
    for (int k = 0; k < 36; ++k) {
        S[k] = 0;
        Af[k] = (float)A[k];
        Qf[k] = (float)Q[k];
    }
    for (int t = 0; t < a.T; ++t) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                sigma_pos[(int64_t)t * 9 + i * 3 + j] = S[i * 6 + j];
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) {
                float s = 0;
                for (int k = 0; k < 6; ++k)
                    s += Af[i * 6 + k] * S[k * 6 + j];
                AS[i * 6 + j] = s;
            }
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) {
                float s = Qf[i * 6 + j];
                for (int k = 0; k < 6; ++k)
                    s += AS[i * 6 + k] * Af[j * 6 + k];
                N[i * 6 + j] = s;
            }
        for (int k = 0; k < 36; ++k)
            S[k] = N[k];
    }
}

} // namespace
} // namespace pump

using namespace pump;

extern "C" int pump_make_noise_device(const pump_noise_spec *sp, float *d_offsets_time,
                                      float *d_xcomb, float *d_sigma_pos, double *d_chol,
                                      pump_stream_t stream)
{
    if (!sp || sp->n_mc <= 0 || sp->tmax_steps <= 0 || !(sp->dt > 0) || !(sp->q > 0) || !d_chol)
        return PUMP_E_INVALID;
    NoiseArgs a;
    a.n_mc = sp->n_mc;
    a.T = sp->tmax_steps;
    a.dt = sp->dt;
    a.kp = sp->kp;
    a.kd = sp->kd;
    a.q = sp->q;
    a.key0 = (uint32_t)sp->seed;
    a.key1 = (uint32_t)(sp->seed >> 32);
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_noise_cholesky, dim3(1), dim3(64), 0, s, a, d_chol);
    PUMP_LAUNCHED();
    if (d_offsets_time || d_xcomb) {
        hipLaunchKernelGGL(k_noise_particles, dim3((a.n_mc + 63) / 64), dim3(64), 0, s, a,
                           (const double *)d_chol, d_offsets_time, d_xcomb);
        PUMP_LAUNCHED();
    }
    if (d_sigma_pos) {
        hipLaunchKernelGGL(k_noise_sigma, dim3(1), dim3(64), 0, s, a, d_sigma_pos);
        PUMP_LAUNCHED();
    }
    return PUMP_OK;
}
