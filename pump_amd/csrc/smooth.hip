// smooth.hip -- smoothing of the search's solution path and its controller
// discretisation (PUMP.cu:406-568), host C++ over the batched full-MC
// entry point.
//
// The reference blends the solution's node states towards the states that
// the obstacle-free optimal path passes at the same time fractions,
// xsSmooth = (1-alpha)*xs + alpha*xsOpt, and bisects alpha on the full-MC CP
// of the blended path: up to 30 sequential collisionProbability calls.  Here
// the bisection is evaluated speculatively: from the current (alpha, lower,
// upper) the next kSmoothDepth steps can take 2^kSmoothDepth - 1 distinct
// alphas (both outcomes of every comparison), so all of them are built on
// the host and their CPs computed in one batched full-MC launch; then the
// sequential rule walks the realised branch.  The alphas, CPs, costs and
// the kept path are exactly those of the sequential loop.
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/pump_host.h"
#include "pump_device.hpp"
#include "pump_internal.hpp"
#include "roadmap_internal.hpp"

#pragma clang fp contract(off)

using namespace pump;
using namespace pump::rmb;

namespace {

constexpr int kSmoothItrs = 30;    // maxSmoothItrs (PUMP.cu:476)
constexpr int kSmoothDepth = 4;    // bisection steps evaluated per full-MC launch
constexpr int kDtMult = 10;        // dtMult (PUMP.cu:422)
constexpr float kEpsSmoothCP = 0.0002f;
constexpr float kControlDt = 0.05f; // controlDt (PUMP.cu:545)

struct Candidate {
    float alpha, lo, hi;
    std::vector<float> xs, path;
    int rows = 0;
    float cost = 0;
};

// xsSmooth, its findOptimalPath rows and its cost (PUMP.cu:486-494)
int build_candidate(Candidate &c, const std::vector<float> &xs, const std::vector<float> &xs_opt,
                    int n, float dt, int cap)
{
    c.xs.resize(xs.size());
    for (size_t i = 0; i < xs.size(); ++i)
        c.xs[i] = (1 - c.alpha) * xs[i] + c.alpha * xs_opt[i];
    c.path.resize((size_t)cap * 6);
    c.rows = pump_find_optimal_path(dt, c.xs.data(), n, c.path.data(), cap, nullptr);
    if (c.rows < 0)
        return PUMP_E_RANGE;
    c.cost = 0;
    for (int i = 0; i < n - 1; ++i) {
        const float tau = pump_host_topt(&c.xs[i * 6], &c.xs[(i + 1) * 6], 2.0f);
        c.cost += pump_host_cost(tau, &c.xs[i * 6], &c.xs[(i + 1) * 6]);
    }
    return PUMP_OK;
}

int smooth(pump_roadmap *rm, const pump_params &P, const pump_search_result &sres,
           pump_smooth_result *out, float *path_host, int path_cap, float *xs_smooth_host,
           float *disc_host, int disc_cap)
{
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t s = rm->stream;
    const int n = sres.soln_len;
    out->n_nodes = n;
    out->cost_before = sres.soln_cost;
    if (sres.status != 0 || n < 2) {
        out->status = 1; // no solution path: nothing to smooth
        return PUMP_OK;
    }
    if (sres.soln_nodes[n - 1] != rm->num - 1 || sres.soln_nodes[0] != 0)
        return PUMP_E_RANGE; // node list truncated
    const int cap = std::min(path_cap, rm->tmax_steps);
    std::vector<float> xs((size_t)n * 6);
    for (int i = 0; i < n; ++i)
        std::memcpy(&xs[(size_t)i * 6], &rm->samples_h[(size_t)sres.soln_nodes[i] * 6],
                    sizeof(float) * 6);

    // the kept path starts as the solution path (PUMP.cu:412-417)
    std::vector<float> best((size_t)cap * 6);
    int best_rows = pump_find_optimal_path(P.dt, xs.data(), n, best.data(), cap, nullptr);
    if (best_rows < 0) {
        out->status = -5;
        return PUMP_OK;
    }
    float best_cost = sres.soln_cost, best_cp = sres.soln_cp;

    // the obstacle-free optimal path and the states it passes at the solution
    // nodes' time fractions (PUMP.cu:425-472)
    const int opt_cap = cap * kDtMult;
    std::vector<float> opt((size_t)opt_cap * 6);
    {
        float ends[12];
        std::memcpy(ends, &xs[0], sizeof(float) * 6);
        std::memcpy(ends + 6, &xs[(size_t)(n - 1) * 6], sizeof(float) * 6);
        const int opt_rows = pump_find_optimal_path(P.dt / kDtMult, ends, 2, opt.data(), opt_cap,
                                                    nullptr);
        if (opt_rows < 1) {
            out->status = -5;
            return PUMP_OK;
        }
        opt.resize((size_t)opt_rows * 6);
    }
    const int opt_rows = (int)(opt.size() / 6);
    std::vector<float> cum(n, 0.0f), xs_opt((size_t)n * 6);
    for (int i = 0; i < n - 1; ++i)
        cum[i + 1] = cum[i] + pump_host_topt(&xs[(size_t)i * 6], &xs[(size_t)(i + 1) * 6], 2.0f);
    for (int i = 0; i < n; ++i) {
        const int idx = (int)((float)(opt_rows - 1) * cum[i] / cum[n - 1]);
        if (idx < 0 || idx >= opt_rows)
            return PUMP_E_RANGE;
        std::memcpy(&xs_opt[(size_t)i * 6], &opt[(size_t)idx * 6], sizeof(float) * 6);
    }

    // the alpha bisection (PUMP.cu:475-537), kSmoothDepth steps per launch
    const int n_tree = (1 << kSmoothDepth) - 1;
    // candidate paths, lengths and hit counts: the roadmap's goal-path
    // selection buffers, kept between calls
    SearchWs *W = search_ws(rm);
    if (!W)
        return PUMP_E_NOMEM;
    DevBuf &d_paths = W->sel_paths, &d_T = W->sel_T, &d_hits = W->sel_hits;
    TRY(d_paths.ensure(sizeof(float) * (size_t)n_tree * cap * 6));
    TRY(d_T.ensure(sizeof(int) * n_tree));
    TRY(d_hits.ensure(sizeof(int) * n_tree));
    std::vector<Candidate> tree(n_tree);
    std::vector<float> hpaths((size_t)n_tree * cap * 6);
    std::vector<int> hT(n_tree), hits(n_tree);
    StreamDrain drain{s};
    std::vector<float> last_xs = xs;
    float alpha = 1.0f, alpha_u = 1.0f, alpha_l = 0.0f;
    int k = 0;
    bool done = false;
    while (k < kSmoothItrs && !done) {
        const int depth = std::min(kSmoothDepth, kSmoothItrs - k);
        const int m = (1 << depth) - 1;
        // node q's children: 2q+1 (CP below the target), 2q+2 (at or above)
        tree[0].alpha = alpha;
        tree[0].lo = alpha_l;
        tree[0].hi = alpha_u;
        for (int q = 0; 2 * q + 2 < m; ++q) {
            const Candidate &c = tree[q];
            tree[2 * q + 1].lo = c.alpha;
            tree[2 * q + 1].hi = c.hi;
            tree[2 * q + 1].alpha = (c.alpha + c.hi) / 2;
            tree[2 * q + 2].lo = c.lo;
            tree[2 * q + 2].hi = c.alpha;
            tree[2 * q + 2].alpha = (c.lo + c.alpha) / 2;
        }
        for (int q = 0; q < m; ++q) {
            if (build_candidate(tree[q], xs, xs_opt, n, P.dt, cap) != PUMP_OK) {
                out->status = -5;
                return PUMP_OK;
            }
            std::memcpy(&hpaths[(size_t)q * cap * 6], tree[q].path.data(),
                        sizeof(float) * (size_t)tree[q].rows * 6);
            hT[q] = tree[q].rows;
        }
        TRYH(hipMemcpyAsync(d_paths.p, hpaths.data(), sizeof(float) * (size_t)m * cap * 6,
                            hipMemcpyHostToDevice, s));
        TRYH(hipMemcpyAsync(d_T.p, hT.data(), sizeof(int) * m, hipMemcpyHostToDevice, s));
        TRY(pump_collision_probability_many(rm->obstacles, rm->n_obs, rm->xcomb, rm->tmax_steps,
                                            rm->n_mc, P.offset_mult, (const float *)d_paths.p,
                                            cap, (const int *)d_T.p, m, (int *)d_hits.p, 1, s));
        TRYH(hipMemcpyAsync(hits.data(), d_hits.p, sizeof(int) * m, hipMemcpyDeviceToHost, s));
        TRYH(hipStreamSynchronize(s));
        out->mccp_launches++;
        // the sequential rule along the realised branch
        int q = 0;
        for (int step = 0; step < depth; ++step) {
            const Candidate &c = tree[q];
            const float cp = ((float)hits[q]) / ((float)rm->n_mc);
            out->trace[k][0] = c.alpha;
            out->trace[k][1] = cp;
            out->trace[k][2] = c.cost;
            last_xs = c.xs;
            if (cp < P.cp_target)
                alpha_l = alpha;
            if (cp >= P.cp_target)
                alpha_u = alpha;
            alpha = (alpha_l + alpha_u) / 2;
            const bool satisfies = cp < P.cp_target;
            const bool closer = cp > best_cp;
            const bool equal_lower = (cp == best_cp && c.cost < best_cost);
            if (satisfies && (closer || equal_lower)) {
                std::memcpy(best.data(), c.path.data(), sizeof(float) * (size_t)c.rows * 6);
                best_rows = c.rows;
                best_cost = c.cost;
                best_cp = cp;
            }
            ++k;
            if (std::fabs(P.cp_target - cp) < kEpsSmoothCP && P.cp_target - cp > 0) {
                done = true;
                break;
            }
            q = satisfies ? 2 * q + 1 : 2 * q + 2;
        }
    }
    out->itrs = k;
    out->cost = best_cost;
    out->cp = best_cp;
    out->path_rows = best_rows;
    if (path_host) {
        if (best_rows > path_cap)
            return PUMP_E_CAPACITY;
        std::memcpy(path_host, best.data(), sizeof(float) * (size_t)best_rows * 6);
    }
    if (xs_smooth_host)
        std::memcpy(xs_smooth_host, last_xs.data(), sizeof(float) * last_xs.size());

    // controller discretisation of the last iterate (PUMP.cu:543-568)
    int rows = 1;
    for (int i = 0; i < n - 1; ++i) {
        const float tau = pump_host_topt(&last_xs[(size_t)i * 6], &last_xs[(size_t)(i + 1) * 6],
                                         2.0f);
        rows += tau / kControlDt + 1;
    }
    out->discretized_rows = rows;
    if (disc_host) {
        if (rows > disc_cap)
            return PUMP_E_CAPACITY;
        std::memset(disc_host, 0, sizeof(float) * (size_t)rows * 6);
        int cur = 0;
        for (int i = 0; i < n - 1; ++i) {
            const float *x0 = &last_xs[(size_t)i * 6], *x1 = &last_xs[(size_t)(i + 1) * 6];
            const float tau = pump_host_topt(x0, x1, 2.0f);
            const int nw = tau / kControlDt + 1;
            if (cur + nw + 1 > rows) {
                out->status = -5; // the reference writes past its array here
                return PUMP_OK;
            }
            // findDiscretizedPath (2pBVP.cu:53-69): rows 0..nw of the edge
            const float topt = pump_host_topt(x0, x1, 5.0f);
            const TauPowers T = tau_powers(topt);
            for (int j = 0; j <= nw; ++j)
                waypoint(j, nw + 1, T, x0, x1, &disc_host[(size_t)(cur + j) * 6]);
            cur += nw;
        }
    }
    out->status = 0;
    out->ms_total =
        std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return PUMP_OK;
}

} // namespace

extern "C" int pump_smooth_run(pump_roadmap *rm, const pump_params *prm,
                               const pump_search_result *sres, pump_smooth_result *out,
                               float *path_host, int path_cap, float *xs_smooth_host,
                               float *discretized_host, int disc_cap)
{
    if (!rm || !prm || !sres || !out || path_cap < 1 || (discretized_host && disc_cap < 1))
        return PUMP_E_INVALID;
    std::memset(out, 0, sizeof(*out));
    if (!rm->xcomb || rm->n_mc < 1)
        return PUMP_E_INVALID;
    return smooth(rm, *prm, *sres, out, path_host, path_cap, xs_smooth_host, discretized_host,
                  disc_cap);
}
