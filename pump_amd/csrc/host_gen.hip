// host_gen.hip -- synthetic problem instances and the Monte Carlo noise
// tables (host C++).
//
// The reference samples with a Halton sequence (sampler.cu:20-41) and reads
// its noise tables xcomb / xcombTime from a pre-generated file that is not
// part of the repository (hardCoded.cuh:15-19).  The workloads here draw
// uniform states and boxes from a counter-based generator, Philox4x32-10
// (J. Salmon, M. Moraes, R. Dror, D. Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11), so any element can be produced independently and
// reproducibly; the noise tables are closed-loop double-integrator offsets
// generated the same way.
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/pump_host.h"

namespace {

constexpr uint32_t kPhiloxM0 = 0xD2511F53u, kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u, kPhiloxW1 = 0xBB67AE85u;

inline void mulhilo(uint32_t a, uint32_t b, uint32_t &hi, uint32_t &lo)
{
    uint64_t p = (uint64_t)a * (uint64_t)b;
    hi = (uint32_t)(p >> 32);
    lo = (uint32_t)p;
}

enum Stream : uint32_t { kStreamSamples = 1, kStreamBoxes = 2, kStreamNoise = 3 };

inline uint32_t draw(uint64_t seed, uint32_t a, uint32_t b, uint32_t stream, uint32_t word = 0)
{
    uint32_t ctr[4] = {a, b, stream, 0};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t out[4];
    pump_philox4x32(ctr, key, out);
    return out[word];
}

// uniform float in [0, 1) from the top 24 bits
inline float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

inline float uniform(float lo, float hi, uint32_t x) { return lo + u01(x) * (hi - lo); }

} // namespace

extern "C" void pump_philox4x32(const uint32_t ctr_in[4], const uint32_t key_in[2],
                                uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += kPhiloxW0;
            k1 += kPhiloxW1;
        }
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo(kPhiloxM0, c0, hi0, lo0);
        mulhilo(kPhiloxM1, c2, hi1, lo1);
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
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

extern "C" int pump_gen_problem(const pump_problem_spec *sp, float *samples, float *obstacles)
{
    if (!sp || sp->num < 2 || sp->n_obs < 0 || !samples || (sp->n_obs > 0 && !obstacles))
        return PUMP_E_INVALID;
    const int num = sp->num;
    for (int i = 0; i < num; ++i)
        for (int d = 0; d < 6; ++d)
            samples[(size_t)i * 6 + d] =
                uniform(sp->lo[d], sp->hi[d], draw(sp->seed, (uint32_t)i, (uint32_t)d, kStreamSamples));
    for (int d = 0; d < 6; ++d) {
        samples[d] = sp->start[d];
        samples[(size_t)(num - 1) * 6 + d] = sp->goal[d];
    }
    for (int o = 0; o < sp->n_obs; ++o) {
        float c[3], s[3];
        auto u = [&](uint32_t k) { return draw(sp->seed, (uint32_t)o, k, kStreamBoxes); };
        for (int d = 0; d < 3; ++d)
            c[d] = uniform(sp->centre_lo, sp->centre_hi, u((uint32_t)d));
        if (sp->obstacle_kind == PUMP_OBS_WALLS) {
            int thin = (int)(u01(u(6)) * 3.0f);
            thin = thin > 2 ? 2 : thin;
            int other = (int)(u01(u(7)) * 2.0f);
            other = other > 1 ? 1 : other;
            int rest[2], m = 0;
            for (int d = 0; d < 3; ++d)
                if (d != thin)
                    rest[m++] = d;
            int lng = rest[other], shrt = rest[1 - other];
            s[thin] = sp->wall_thickness;
            s[lng] = uniform(sp->wall_long_lo, sp->wall_long_hi, u(8));
            s[shrt] = uniform(sp->wall_short_lo, sp->wall_short_hi, u(9));
        } else {
            for (int d = 0; d < 3; ++d)
                s[d] = uniform(sp->side_lo, sp->side_hi, u((uint32_t)(3 + d)));
        }
        float *b = obstacles + (size_t)o * 12;
        for (int d = 0; d < 3; ++d) {
            b[d] = c[d] - s[d] / 2.0f;
            b[6 + d] = c[d] + s[d] / 2.0f;
            b[3 + d] = -sp->vel_span;
            b[9 + d] = sp->vel_span;
        }
    }
    return PUMP_OK;
}

extern "C" void pump_noise_defaults(pump_noise_spec *s)
{
    s->n_mc = 2048;   // numMCParticles (hardCoded.cuh:13)
    s->tmax_steps = 2001;
    s->dt = 0.05f;    // mainPUMP.cu:60
    s->kp = 4.0;      // natural frequency 2 rad/s, critically damped
    s->kd = 4.0;
    s->q = 0.1;       // steady-state position sigma ~0.05 per axis
    s->seed = 0x5eed;
}

// 6x6 Cholesky (FP64), lower-triangular L with L L^T = A
static bool cholesky6(const double *A, double *L)
{
    std::memset(L, 0, sizeof(double) * 36);
    for (int i = 0; i < 6; ++i) {
        for (int j = 0; j <= i; ++j) {
            double s = A[i * 6 + j];
            for (int k = 0; k < j; ++k)
                s -= L[i * 6 + k] * L[j * 6 + k];
            if (i == j) {
                if (s <= 0)
                    return false;
                L[i * 6 + i] = std::sqrt(s);
            } else {
                L[i * 6 + j] = s / L[j * 6 + j];
            }
        }
    }
    return true;
}

extern "C" int pump_make_noise(const pump_noise_spec *sp, float *offsets_time, float *xcomb,
                               float *sigma_pos)
{
    if (!sp || sp->n_mc <= 0 || sp->tmax_steps <= 0 || sp->dt <= 0 || sp->q <= 0)
        return PUMP_E_INVALID;
    const int n_mc = sp->n_mc, T = sp->tmax_steps;
    const double dt = sp->dt;
    // closed-loop per-axis transition
    const double a11 = 1 - sp->kp * dt * dt / 2, a12 = dt - sp->kd * dt * dt / 2;
    const double a21 = -sp->kp * dt, a22 = 1 - sp->kd * dt;
    double A[36] = {0}, Q[36] = {0}, L[36];
    for (int d = 0; d < 3; ++d) {
        A[d * 6 + d] = a11;
        A[d * 6 + d + 3] = a12;
        A[(d + 3) * 6 + d] = a21;
        A[(d + 3) * 6 + d + 3] = a22;
        Q[d * 6 + d] = sp->q * dt * dt * dt / 3;
        Q[d * 6 + d + 3] = sp->q * dt * dt / 2;
        Q[(d + 3) * 6 + d] = sp->q * dt * dt / 2;
        Q[(d + 3) * 6 + d + 3] = sp->q * dt;
    }
    if (!cholesky6(Q, L))
        return PUMP_E_RANGE;
    for (int p = 0; p < n_mc; ++p) {
        double x[6] = {0, 0, 0, 0, 0, 0};
        for (int t = 0; t < T; ++t) {
            for (int d = 0; d < 3; ++d) {
                float v = (float)x[d];
                if (offsets_time)
                    offsets_time[((size_t)t * n_mc + p) * 3 + d] = v;
                if (xcomb)
                    xcomb[((size_t)p * T + t) * 3 + d] = v;
            }
            // z ~ N(0, I6) by Box-Muller from three Philox blocks
            double z[6];
            for (int j = 0; j < 3; ++j) {
                uint32_t ctr[4] = {(uint32_t)p, (uint32_t)t, kStreamNoise, (uint32_t)j};
                uint32_t key[2] = {(uint32_t)sp->seed, (uint32_t)(sp->seed >> 32)};
                uint32_t out[4];
                pump_philox4x32(ctr, key, out);
                double u1 = ((double)out[0] + 0.5) * (1.0 / 4294967296.0);
                double u2 = ((double)out[1] + 0.5) * (1.0 / 4294967296.0);
                double r = std::sqrt(-2.0 * std::log(u1));
                z[2 * j] = r * std::cos(6.283185307179586 * u2);
                z[2 * j + 1] = r * std::sin(6.283185307179586 * u2);
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
            std::memcpy(x, xn, sizeof(x));
        }
    }
    if (sigma_pos) {
        // FP32 covariance propagation Sigma_{t+1} = A Sigma_t A^T + Q, Sigma_0 = 0
        float S[36] = {0}, Af[36], Qf[36];
        for (int k = 0; k < 36; ++k) {
            Af[k] = (float)A[k];
            Qf[k] = (float)Q[k];
        }
        for (int t = 0; t < T; ++t) {
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    sigma_pos[(size_t)t * 9 + i * 3 + j] = S[i * 6 + j];
            float AS[36], N[36];
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
            std::memcpy(S, N, sizeof(S));
        }
    }
    return PUMP_OK;
}
