// ref_harness.cpp -- C entry points around the reference's own 2pBVP.cu,
// compiled from the sources where they lie (see oracle/ref/Makefile).
//
// TEST INFRASTRUCTURE ONLY: used in this container to pin the oracle's
// 2pBVP restatement bit for bit and to write tests/golden fixtures.
// The reference's host functions are called unchanged; this file only loops
// over inputs and exposes a C ABI for ctypes.
#include "2pBVP.cuh"
#include <cstdint>
#include <cstring>
#include <vector>

extern "C" {

float ref_cost(float tau, float *x0, float *x1) { return cost(tau, x0, x1); }
float ref_dcost(float tau, float *x0, float *x1) { return dcost(tau, x0, x1); }
float ref_topt_bisection(float *x0, float *x1, float tmax) { return toptBisection(x0, x1, tmax); }
void ref_path_point(float t, float tau, float *x0, float *x1, float *x) { pathPoint(t, tau, x0, x1, x); }
float ref_find_discretized_path(float *wp, float *x0, float *x1, int num_disc)
{
    return findDiscretizedPath(wp, x0, x1, num_disc);
}

// findOptimalPath writes without a bound; run it into a large scratch buffer
// and copy out at most cap rows.
float ref_find_optimal_path(float dt, float *split, int cap, float *xs, int n_wp, int *len)
{
    std::vector<float> scratch((size_t)(n_wp + 1) * 2000 * DIM);
    int plen = 0;
    float tt = findOptimalPath(dt, scratch.data(), xs, n_wp, &plen);
    *len = plen;
    int rows = plen < cap ? plen : cap;
    std::memcpy(split, scratch.data(), sizeof(float) * (size_t)rows * DIM);
    return tt;
}

// all ordered pairs, in fillCoptsTopts' output order (2pBVP.cu:144-149),
// evaluated with the reference's host toptBisection/cost
void ref_fill_pairs(float *samples, int num, float tmax, float *copts, float *topts)
{
    for (int i = 0; i < num; ++i)
        for (int j = 0; j < num; ++j) {
            if (i == j)
                continue;
            int64_t idx = (int64_t)i * (num - 1) + j - (i < j ? 1 : 0);
            topts[idx] = toptBisection(&samples[i * DIM], &samples[j * DIM], tmax);
            copts[idx] = cost(topts[idx], &samples[i * DIM], &samples[j * DIM]);
        }
}

}  // extern "C"
