// bvp.hip -- kernel (a): 2pBVP optimal arrival time and cost for every
// ordered sample pair (reference: fillCoptsTopts, 2pBVP.cu:133-150), the
// cost-to-goal heuristic (mainPUMP.cu:351-355) and the FP64 Newton polish of
// arrival times.
//
// One lane per pair.  Pairs are enumerated directly in the output order
// idx = i*(num-1) + k (k skips the diagonal), so stores of copts/topts are
// fully coalesced and the source state x_i is uniform across most of a wave.
// The work per pair is a few dozen FP64 cost evaluations (12.3 bisection
// steps of two evaluations per pair on average for 16000 uniform samples):
// the kernel is bound by VALU issue.  Lanes finish their bisection
// after different iteration counts; to keep them busy each wave pulls pairs
// from a shared counter when its lanes run dry (persistent waves).
#include "pump_device.hpp"
#include "pump_internal.hpp"

#pragma clang fp contract(off)

namespace pump {
namespace {

__device__ __forceinline__ void load_state(const float *__restrict__ s, int i, float *x)
{
    const float *p = s + (int64_t)i * kDim;
    for (int d = 0; d < kDim; ++d)
        x[d] = p[d];
}

__device__ __forceinline__ void pair_of(int64_t idx, int num, int &i, int &j)
{
    i = (int)(idx / (num - 1));
    int k = (int)(idx - (int64_t)i * (num - 1));
    j = k + (k >= i ? 1 : 0);
}

// Straight one-lane-per-pair form (used for small launches and as the
// simplest statement of the mapping).
__global__ __launch_bounds__(256) void k_fill_pairs(const float *__restrict__ samples, int num,
                                                    float tmax, float *__restrict__ copts,
                                                    float *__restrict__ topts, int64_t n_pairs)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n_pairs;
         idx += stride) {
        int i, j;
        pair_of(idx, num, i, j);
        float a[kDim], b[kDim];
        load_state(samples, i, a);
        load_state(samples, j, b);
        CostTerms c;
        cost_terms(c, a, b);
        float tau = topt_bisection(c, tmax);
        topts[idx] = tau;
        copts[idx] = cost_eval(c, tau);
    }
}

// Bisection as a per-lane state machine so that a lane that finishes takes
// the next pair while its wave-mates are still iterating.  Every step
// evaluates cost at two points per lane: the two of one dcost (tau +- h/2,
// as dcost forms them), or, in the final phase, cost(topt) for the output.
// The sequence of dcost evaluations per pair is the one toptBisection
// performs, so the results are identical.  Idle lanes are refilled only
// once kRefill of them have gathered (or none is busy), so that the
// wave-wide cost of setting up new pairs is paid once for many lanes.
enum : int { ST_TMAX = 0, ST_LOW = 1, ST_BIS = 2, ST_FIN = 3 };
constexpr int kRefill = 16;
constexpr int kChunk = 2048; // pairs per claim of the work counter

// Pair sources: map a global pair index to the pair's cost terms, tmax and
// output slots.  One instance (idx in the reference's output order) ...
// With fewer than 2^32 pairs the row i = g / (num - 1) is found by a
// multiply-high with the divisor's magic number (DivMagic).
struct OnePairSource {
    const float *samples;
    int num;
    float tmax;
    float *copts, *topts;
    DivMagic dm;
    __device__ void take(int64_t g, CostTerms &c, float &tm, float *&oc, float *&ot,
                         int64_t &oi) const
    {
        int i, j;
        if (dm.m) {
            const uint32_t n = (uint32_t)g;
            const uint32_t q = div_magic_apply(n, dm);
            const int k = (int)(n - q * (uint32_t)(num - 1));
            i = (int)q;
            j = k + (k >= i ? 1 : 0);
        } else {
            pair_of(g, num, i, j);
        }
        float a[kDim], b[kDim];
        load_state(samples, i, a);
        load_state(samples, j, b);
        cost_terms(c, a, b);
        tm = tmax;
        oc = copts;
        ot = topts;
        oi = g;
    }
};

// ... or many: instance k owns global indices [prefix[k], prefix[k+1]).
// A lane keeps the range of the instance its last pair came from, so the
// search over the prefix table runs only when a lane crosses into another
// instance (a wave's 2048-pair chunks mostly lie inside one).
struct ManyPairSource {
    const PairJob *jobs;
    const int64_t *prefix;
    int n_inst;
    mutable int cur = 0;
    mutable int64_t cur_lo = 0, cur_hi = 0;
    __device__ void take(int64_t g, CostTerms &c, float &tm, float *&oc, float *&ot,
                         int64_t &oi) const
    {
        if (!(g >= cur_lo && g < cur_hi)) {
            int lo = 0, hi = n_inst;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (prefix[mid] <= g)
                    lo = mid;
                else
                    hi = mid;
            }
            cur = lo;
            cur_lo = prefix[lo];
            cur_hi = prefix[lo + 1];
        }
        const PairJob J = global_job(jobs[cur]);
        const int64_t idx = g - cur_lo;
        int i, j;
        if (J.dm_m) {
            const uint32_t n = (uint32_t)idx;
            const uint32_t q = div_magic_apply(n, DivMagic{J.dm_m, J.dm_s1, J.dm_s2});
            const int k = (int)(n - q * (uint32_t)(J.num - 1));
            i = (int)q;
            j = k + (k >= i ? 1 : 0);
        } else {
            pair_of(idx, J.num, i, j);
        }
        float a[kDim], b[kDim];
        load_state(J.samples, i, a);
        load_state(J.samples, j, b);
        cost_terms(c, a, b);
        tm = J.tmax;
        oc = J.copts;
        ot = J.topts;
        oi = idx;
    }
};

// steps_out (may be null) receives the number of (lane, step) pairs
// executed, i.e. the bisection steps of all pairs: each step is two cost
// evaluations.
template <class Src>
__device__ __forceinline__ void pairs_persistent(const Src &src, int64_t n_pairs,
                                                 unsigned long long *__restrict__ next_chunk,
                                                 unsigned long long *__restrict__ steps_out)
{
    const int lane = threadIdx.x & 63;
    // each wave claims kChunk pairs at a time from the global counter (one
    // counter for the whole launch: small claims would serialise on it)

    int64_t my = -1;     // current pair of this lane, -1 when idle
    int64_t pool = 0;    // next unclaimed pair index in the wave's chunk
    int64_t pool_end = 0;
    CostTerms c;
    float tu = 0, tl = 0, topt = 0, tmax = 0;
    float dcval = 0; // numerator ca - cb of the last dcost (1 before the first)
    float *oc = nullptr, *ot = nullptr;
    int64_t oi = 0;
    int state = ST_TMAX, itr = 0, halvings = 0;
    bool done_all = false;
    unsigned long long steps = 0; // wave-uniform

    while (true) {
        // refill idle lanes from the wave's chunk, claiming a new chunk as needed
        unsigned long long idle = __ballot(my < 0);
        const bool refill = __popcll(idle) >= kRefill || idle == ~0ull;
        while (refill && idle != 0 && !done_all) {
            if (pool >= pool_end) {
                unsigned long long claimed = 0;
                if (lane == 0)
                    claimed = atomicAdd(next_chunk, (unsigned long long)kChunk);
                // every lane is active here, so the first active lane is lane 0
                const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)claimed);
                const uint32_t hi =
                    (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(claimed >> 32));
                int64_t b = (int64_t)(((uint64_t)hi << 32) | lo);
                if (b >= n_pairs) {
                    done_all = true;
                    break;
                }
                pool = b;
                pool_end = b + kChunk < n_pairs ? b + kChunk : n_pairs;
            }
            // rank of this lane among idle lanes
            int rank = __popcll(idle & ((1ull << lane) - 1ull));
            int64_t avail = pool_end - pool;
            bool take = (my < 0) && (rank < avail);
            if (take) {
                my = pool + rank;
                src.take(my, c, tmax, oc, ot, oi);
                state = ST_TMAX;
                tu = tmax;
                tl = 0.01f;
                topt = 0.0f;
                dcval = 1.0f;
                itr = 0;
                halvings = 0;
            }
            int took = __popcll(idle) < avail ? __popcll(idle) : (int)avail;
            pool += took;
            idle = __ballot(my < 0);
        }
        const unsigned long long busy = __ballot(my >= 0);
        if (busy == 0) {
            if (done_all)
                break;
            continue;
        }
        steps += (unsigned long long)__popcll(busy);
        if (my >= 0) {
            // the step's two cost evaluations
            const float x = (state == ST_FIN)    ? topt
                            : (state == ST_TMAX) ? tmax
                            : (state == ST_LOW)  ? tl
                                                 : (tu + tl) / 2.0f;
            const float h = 0.000001f;
            const float xa = state == ST_FIN ? x : x + h / 2.0f;
            const float xb = state == ST_FIN ? x : x - h / 2.0f;
            // For 2^-20 <= x <= 2^100 both points are at least 4.5e-7, so
            // tau^3 stays within the reciprocal's exact range.
            float ca, cb;
            if (x >= 0x1p-20f && x <= 0x1p100f) {
                ca = cost_eval_with(c, xa, RecipNewton{});
                cb = cost_eval_with(c, xb, RecipNewton{});
            } else {
                ca = cost_eval(c, xa);
                cb = cost_eval(c, xb);
            }
            if (state == ST_FIN) {
                ot[oi] = topt;
                oc[oi] = ca;
                my = -1;
            } else {
                // dcost = (ca - cb) / h (2pBVP.cu:109-115) enters only
                // through its sign, which is that of dc, and through
                // |dcost| > 0.0001, which is |dc| >= kDcostBig
                const float dc = ca - cb;
                if (state == ST_TMAX) {
                    if (dc < 0.0f) {
                        state = ST_FIN;
                        topt = tmax;
                    } else {
                        state = ST_LOW;
                    }
                } else if (state == ST_LOW) {
                    if (dc > 0.0f && halvings < 400) {
                        tl = tl / 2.0f;
                        ++halvings;
                    } else {
                        state = ST_BIS;
                        if (!(fabsf(dcval) >= kDcostBig && itr < 50))
                            state = ST_FIN; // topt stays 0, as the reference returns it
                    }
                } else {
                    ++itr;
                    topt = x;
                    dcval = dc;
                    if (dcval > 0.0f)
                        tu = topt;
                    else
                        tl = topt;
                    if (!(fabsf(dcval) >= kDcostBig && itr < 50))
                        state = ST_FIN;
                }
            }
        }
    }
    if (steps_out && lane == 0)
        atomicAdd(steps_out, steps);
}

__global__ __launch_bounds__(256) void k_fill_pairs_persistent(
    const float *__restrict__ samples, int num, float tmax, float *__restrict__ copts,
    float *__restrict__ topts, int64_t n_pairs, DivMagic dm,
    unsigned long long *__restrict__ next_chunk, unsigned long long *__restrict__ steps_out)
{
    const OnePairSource src{samples, num, tmax, copts, topts, dm};
    pairs_persistent(src, n_pairs, next_chunk, steps_out);
}

// all ordered pairs of many instances in one persistent launch
__global__ __launch_bounds__(256) void k_fill_pairs_many(const PairJob *__restrict__ jobs,
                                                         const int64_t *__restrict__ prefix,
                                                         int n_inst, int64_t n_pairs,
                                                         unsigned long long *__restrict__ next_chunk)
{
    const ManyPairSource src{jobs, prefix, n_inst};
    pairs_persistent(src, n_pairs, next_chunk, nullptr);
}

__global__ void k_zero_counter(unsigned long long *c)
{
    if (threadIdx.x == 0)
        *c = 0ull;
}

// h[i] = cost(toptBisection(x_i, goal)) for i < num-1, h[num-1] = 0
__global__ __launch_bounds__(256) void k_heuristic(const float *__restrict__ samples, int num,
                                                   float g0, float g1, float g2, float g3,
                                                   float g4, float g5, float tmax,
                                                   float *__restrict__ h)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num)
        return;
    if (i == num - 1) {
        h[i] = 0.0f;
        return;
    }
    float a[kDim], b[kDim] = {g0, g1, g2, g3, g4, g5};
    load_state(samples, i, a);
    CostTerms c;
    cost_terms(c, a, b);
    float tau = topt_bisection(c, tmax);
    h[i] = cost_eval(c, tau);
}

// FP64 polish of a bisected arrival time: the root of
//   q(t) = t^4 c'(t) = (1+g^2) t^4 - 4R t^2 + 24 S t - 36 D
// at the same local minimum.  A bracket [lo, hi] with q(lo) <= 0 <= q(hi) is
// grown geometrically around the bisection result (the float finite
// difference can stop anywhere within its noise, and at tmax it can read
// negative while c'(tmax) > 0), then safeguarded Newton steps converge to the
// last bit.  tmax is kept when c'(tmax) < 0 exactly.
__device__ __forceinline__ double q_of(double t, double k4, double D, double S, double R)
{
    return ((k4 * t * t - 4.0 * R) * t + 24.0 * S) * t - 36.0 * D;
}

__global__ __launch_bounds__(256) void k_polish(const float *__restrict__ samples, int num,
                                                float tmax, const float *__restrict__ topts,
                                                double *__restrict__ out, int64_t n_pairs)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const double k4 = 1.0 + (double)kGravity * (double)kGravity;
    const double tm = (double)tmax;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n_pairs;
         idx += stride) {
        int i, j;
        pair_of(idx, num, i, j);
        float a[kDim], b[kDim];
        load_state(samples, i, a);
        load_state(samples, j, b);
        double D = 0, S = 0, R = 0;
        for (int d = 0; d < 3; ++d) {
            double dp = (double)b[d] - (double)a[d];
            double v0 = a[d + 3], v1 = b[d + 3];
            D += dp * dp;
            S += dp * (v0 + v1);
            R += v0 * v0 + v0 * v1 + v1 * v1;
        }
        const double t0 = (double)topts[idx];
        double t = t0;
        if (t0 >= tm && q_of(tm, k4, D, S, R) < 0.0) {
            out[idx] = tm;
            continue;
        }
        double lo = t0, hi = t0;
        bool ok = false;
        for (int k = 0; k < 64; ++k) {
            double w = t0 * 1e-6 * (double)(1ull << (k < 60 ? k : 60));
            lo = fmax(t0 - w, 1e-9);
            hi = fmin(t0 + w, tm);
            if (q_of(lo, k4, D, S, R) <= 0.0 && q_of(hi, k4, D, S, R) >= 0.0) {
                ok = true;
                break;
            }
        }
        if (ok) {
            t = fmin(fmax(t0, lo), hi);
            for (int it = 0; it < 80; ++it) {
                double q = q_of(t, k4, D, S, R);
                if (q == 0.0)
                    break;
                if (q > 0.0)
                    hi = t;
                else
                    lo = t;
                double dq = (4.0 * k4 * t * t - 8.0 * R) * t + 24.0 * S;
                double tn = t - q / dq;
                if (!(tn > lo && tn < hi))
                    tn = 0.5 * (lo + hi);
                if (fabs(tn - t) <= 1e-16 * t) {
                    t = tn;
                    break;
                }
                t = tn;
            }
        }
        out[idx] = t;
    }
}

} // namespace
} // namespace pump

using namespace pump;

namespace pump {
// All ordered pairs of one instance; *steps (may be null) receives the
// bisection steps of all pairs, 0 for a launch below 65,536 pairs (the
// plain one-lane-per-pair kernel, which does not count them).
int fill_copts_topts(const float *d_samples, int num, float tmax, float *d_copts, float *d_topts,
                     hipStream_t s, int64_t *steps)
{
    PUMP_REQUIRE(d_samples && d_copts && d_topts && num >= 2);
    const int64_t n_pairs = (int64_t)num * (num - 1);
    if (steps)
        *steps = 0;
    if (n_pairs < (1 << 16)) {
        hipLaunchKernelGGL(k_fill_pairs, dim3(grid_for(n_pairs, 256)), dim3(256), 0, s, d_samples,
                           num, tmax, d_copts, d_topts, n_pairs);
        PUMP_LAUNCHED();
        return PUMP_OK;
    }
    // [0]: work counter, [1]: step count
    unsigned long long *counter = nullptr;
    PUMP_HIP(hipMalloc((void **)&counter, 2 * sizeof(unsigned long long)));
    hipError_t me = hipMemsetAsync(counter, 0, 2 * sizeof(unsigned long long), s);
    if (me != hipSuccess) {
        hipFree(counter);
        PUMP_HIP(me);
    }
    int dev = 0, cus = 256;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    // 8 resident 256-lane workgroups per CU
    int64_t blocks = (int64_t)cus * 8;
    int64_t need = (n_pairs + 255) / 256;
    if (blocks > need)
        blocks = need;
    const DivMagic dm = n_pairs < (1ll << 32) ? div_magic((uint32_t)(num - 1)) : DivMagic{0, 0, 0};
    hipLaunchKernelGGL(k_fill_pairs_persistent, dim3((unsigned)blocks), dim3(256), 0, s,
                       d_samples, num, tmax, d_copts, d_topts, n_pairs, dm, counter,
                       steps ? counter + 1 : nullptr);
    hipError_t le = hipGetLastError();
    hipError_t se = hipStreamSynchronize(s);
    unsigned long long n_steps = 0;
    if (le == hipSuccess && se == hipSuccess && steps)
        se = hipMemcpy(&n_steps, counter + 1, sizeof(n_steps), hipMemcpyDeviceToHost);
    hipFree(counter);
    PUMP_HIP(le);
    PUMP_HIP(se);
    if (steps)
        *steps = (int64_t)n_steps;
    return PUMP_OK;
}
} // namespace pump

extern "C" int pump_fill_copts_topts(const float *d_samples, int num, float tmax, float *d_copts,
                                     float *d_topts, pump_stream_t stream)
{
    return pump::fill_copts_topts(d_samples, num, tmax, d_copts, d_topts, as_stream(stream),
                                  nullptr);
}

extern "C" int pump_polish_topts(const float *d_samples, int num, float tmax, const float *d_topts,
                                 double *d_topts64, pump_stream_t stream)
{
    PUMP_REQUIRE(d_samples && d_topts && d_topts64 && num >= 2);
    const int64_t n_pairs = (int64_t)num * (num - 1);
    hipLaunchKernelGGL(k_polish, dim3(grid_for(n_pairs, 256)), dim3(256), 0, as_stream(stream),
                       d_samples, num, tmax, d_topts, d_topts64, n_pairs);
    PUMP_LAUNCHED();
    return PUMP_OK;
}

extern "C" int pump_heuristic(const float *d_samples, int num, const float *goal_host, float tmax,
                              float *d_h, pump_stream_t stream)
{
    PUMP_REQUIRE(d_samples && goal_host && d_h && num >= 1);
    const float *g = goal_host;
    hipLaunchKernelGGL(k_heuristic, dim3((num + 255) / 256), dim3(256), 0, as_stream(stream),
                       d_samples, num, g[0], g[1], g[2], g[3], g[4], g[5], tmax, d_h);
    PUMP_LAUNCHED();
    return PUMP_OK;
}

namespace pump {
// Many instances' pairs in one persistent launch (jobs and prefix in device
// memory; prefix[n_inst] = n_pairs).  The chunk counter is cleared by a
// kernel, so the sequence can be captured in a graph.
int fill_pairs_many(const PairJob *d_jobs, const int64_t *d_prefix, int n_inst, int64_t n_pairs,
                    int blocks_per_cu, unsigned long long *d_counter, hipStream_t s)
{
    PUMP_REQUIRE(d_jobs && d_prefix && d_counter && n_inst > 0 && n_pairs >= 0 &&
                 blocks_per_cu > 0);
    if (n_pairs == 0)
        return PUMP_OK;
    int dev = 0, cus = 256;
    PUMP_HIP(hipGetDevice(&dev));
    PUMP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    int64_t blocks = (int64_t)cus * blocks_per_cu;
    const int64_t need = (n_pairs + 255) / 256;
    if (blocks > need)
        blocks = need;
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, s, d_counter);
    PUMP_LAUNCHED();
    hipLaunchKernelGGL(k_fill_pairs_many, dim3((unsigned)blocks), dim3(256), 0, s, d_jobs,
                       d_prefix, n_inst, n_pairs, d_counter);
    PUMP_LAUNCHED();
    return PUMP_OK;
}
} // namespace pump

