// problem_io.hip -- the reference's problem files and its sampler (host
// C++, no kernels):
//   * the problem file read by mainPUMP.cu:152-230: comma-separated values,
//     init (6), goal (6), lo (6), hi (6), offsetMult, the obstacle count and
//     12 values per obstacle (lowest corner, then highest); anything after
//     the last value is ignored.  Each value is the text between two commas
//     parsed from its start after white space, as `stringstream >> float`
//     does (correctly rounded, strtof);
//   * createSamplesHalton (sampler.cu:20-57): dimension d of sample n is
//     the radical inverse of n in base bases[d + skip], evaluated in float in
//     the reference's operation order, scaled to [lo, hi]; sample 0 is the
//     initial state and sample NUM-1 the goal.
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/pump_host.h"

#pragma clang fp contract(off)

namespace {

constexpr int kHaltonPrimes[25] = {2,  3,  5,  7,  11, 13, 17, 19, 23, 29, 31, 37, 41,
                                   43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97};

// localHaltonSingleNumber (sampler.cu:45-57), the same float/int steps
float halton(int n, int b)
{
    float hn = 0;
    int n0 = n;
    float f = 1 / ((float)b);
    while (n0 > 0) {
        float n1 = n0 / b;
        int r = n0 - n1 * b;
        hn += f * r;
        f = f / b;
        n0 = n1;
    }
    return hn;
}

struct Tokens {
    std::string text;
    size_t at = 0;
    // the next comma-separated field (the rest of the text for the last one)
    bool next(std::string &tok)
    {
        if (at > text.size())
            return false;
        const size_t c = text.find(',', at);
        if (c == std::string::npos) {
            tok = text.substr(at);
            at = text.size() + 1;
        } else {
            tok = text.substr(at, c - at);
            at = c + 1;
        }
        return true;
    }
    bool next_float(float &v)
    {
        std::string tok;
        if (!next(tok))
            return false;
        const char *s = tok.c_str();
        char *end = nullptr;
        errno = 0;
        v = std::strtof(s, &end);
        return end != s;
    }
    bool next_int(int &v)
    {
        std::string tok;
        if (!next(tok))
            return false;
        const char *s = tok.c_str();
        char *end = nullptr;
        const long x = std::strtol(s, &end, 10);
        if (end == s || x < 0 || x > 1000000)
            return false;
        v = (int)x;
        return true;
    }
};

} // namespace

extern "C" int pump_read_problem(const char *path, pump_problem_file *out, float *obstacles_host,
                                 int obs_cap)
{
    if (!path || !out)
        return PUMP_E_INVALID;
    std::memset(out, 0, sizeof(*out));
    FILE *f = std::fopen(path, "rb");
    if (!f)
        return PUMP_E_INVALID;
    Tokens T;
    char buf[4096];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0)
        T.text.append(buf, got);
    std::fclose(f);
    float *fields[4] = {out->init, out->goal, out->lo, out->hi};
    for (float *v : fields)
        for (int d = 0; d < 6; ++d)
            if (!T.next_float(v[d]))
                return PUMP_E_INVALID;
    if (!T.next_float(out->offset_mult) || !T.next_int(out->n_obs))
        return PUMP_E_INVALID;
    if (out->n_obs > 0 && (!obstacles_host || out->n_obs > obs_cap))
        return out->n_obs > obs_cap ? PUMP_E_CAPACITY : PUMP_E_INVALID;
    for (int i = 0; i < out->n_obs * 12; ++i)
        if (!T.next_float(obstacles_host[i]))
            return PUMP_E_INVALID;
    return PUMP_OK;
}

extern "C" int pump_halton_samples(int num, int skip, const float *init, const float *goal,
                                   const float *lo, const float *hi, float *samples_host)
{
    if (num < 2 || skip < 0 || skip + 6 > 25 || !init || !goal || !lo || !hi || !samples_host)
        return PUMP_E_INVALID;
    for (int d = 0; d < 6; ++d)
        for (int n = 0; n < num; ++n)
            samples_host[(size_t)n * 6 + d] = halton(n, kHaltonPrimes[d + skip]) * (hi[d] - lo[d]) +
                                              lo[d];
    for (int d = 0; d < 6; ++d) {
        samples_host[d] = init[d];
        samples_host[(size_t)(num - 1) * 6 + d] = goal[d];
    }
    return PUMP_OK;
}
