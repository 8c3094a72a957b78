// batch.hip -- batched roadmap build over many independent problems (host
// C++; the kernels are in bvp.hip, edges.hip and hsmc.hip).
//
// Each instance is the reference's per-problem precompute
// (mainPUMP.cu:248-504) with its own samples, boxes and parameters.  The
// three device passes whose work grows with the instance -- all ordered
// pairs (2pBVP.cu:133-150), the fused discretize + freeEdges pass
// (2pBVP.cu:53-69, collisionCheck.cu:5-37) and cacheHalfspaces with the
// setupAsBs sentinels (hsmc.cu:10-99,201-211) -- run for every instance in one
// launch each.  The workgroups of a launch are split among the instances by a
// prefix table in device memory, so small instances share the chip instead of
// each leaving most CUs idle.  Each launch sequence is captured in a hipGraph
// and can be replayed.  The launch shape of the pair pass (resident
// workgroups per CU) and of the edge pass (edges per workgroup, boxes staged
// in LDS) is chosen by timing the candidates on the device, and kept for the
// process per configuration shape.  The host-synchronous stages between
// the passes (rn, edge numbering, neighbour tables, waypoint offsets) run per
// instance on the batch's one stream, as in the single-instance build.
#include <algorithm>
#include <array>
#include <chrono>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/pump_host.h"
#include "hsmc_wave.hpp"
#include "pump_device.hpp"
#include "pump_internal.hpp"
#include "roadmap_internal.hpp"

using namespace pump;
using namespace pump::rmb;

// One instance's MC step in the batch's third graph: its wavefront (free
// edges expanded from all-valid parents), parent and child validity bits,
// parent and child CPs, row sources and workspace.
struct BatchMc {
    int *edge = nullptr, *prev = nullptr, *row_src = nullptr;
    float *times = nullptr, *parent_cp = nullptr, *cp = nullptr;
    uint64_t *src = nullptr, *dst = nullptr;
    char *ws = nullptr;
    size_t ws_bytes = 0;
    int n_new = 0;
};

struct pump_batch {
    hipStream_t stream = nullptr;
    std::vector<pump_roadmap *> rms;
    float *offsets_time = nullptr, *xcomb = nullptr;
    uint16_t *offsets_split = nullptr;
    PairJob *pair_jobs = nullptr;
    int64_t *pair_prefix = nullptr;
    unsigned long long *counter = nullptr;
    EdgeJob *edge_jobs = nullptr;
    int64_t *edge_prefix = nullptr; // one table of n_inst + 1 per edges-per-block candidate
    HsJob *hs_jobs = nullptr;
    int64_t *hs_prefix = nullptr;
    hipGraph_t graph[3] = {nullptr, nullptr, nullptr};
    hipGraphExec_t exec[3] = {nullptr, nullptr, nullptr};
    pump_batch_info info{};
    std::vector<BatchMc> mc;
    int mc_parents = 0;
    // side streams (default priority) and fork/join events of the MC stage
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> ev;
    // the MC stage as one launch per stage for all problems (when every
    // problem's step takes the binned form): jobs, block maps, grids
    McJob *mc_jobs = nullptr;
    void *mc_maps = nullptr;
    McGrids mc_grids{};
    bool mc_many = false;
};

namespace {

constexpr int kEpb[3] = {64, 128, 256};    // edges per workgroup of the edge pass
constexpr int kBpc[3] = {4, 8, 16};        // resident workgroups per CU of the pair pass
constexpr int kObsTile[4] = {32, 64, 128, 256}; // boxes staged in LDS by the edge pass
constexpr int64_t kTunePairs = 1ll << 26;  // pairs timed per pair-pass candidate

// Tuned launch shapes, kept for the process.  Key: the stage, the instance
// count, its totals (pairs; or edges and waypoints) and the largest box
// count.  Value: the chosen candidate indices.
using TuneKey = std::array<int64_t, 5>;
std::mutex g_tune_mu;
std::map<TuneKey, std::array<int, 2>> g_tuned;

bool tuned_get(const TuneKey &k, std::array<int, 2> *v)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    const auto it = g_tuned.find(k);
    if (it == g_tuned.end())
        return false;
    *v = it->second;
    return true;
}

void tuned_put(const TuneKey &k, const std::array<int, 2> &v)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[k] = v;
}

template <typename T> int upload(T **d, const std::vector<T> &h, hipStream_t s)
{
    TRY(dalloc(d, h.size()));
    TRYH(hipMemcpyAsync(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
    return PUMP_OK;
}

// Captures what launch() enqueues on the batch stream as graph `stage`.
template <typename F> int capture(pump_batch *b, int stage, F launch)
{
    TRYH(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
    const int r = launch();
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (r != PUMP_OK || e != hipSuccess) {
        if (g)
            hipGraphDestroy(g);
        if (r != PUMP_OK)
            return r;
        TRYH(e);
    }
    b->graph[stage] = g;
    TRYH(hipGraphInstantiate(&b->exec[stage], g, nullptr, nullptr, 0));
    return PUMP_OK;
}

int launch_timed(pump_batch *b, int stage, float *ms)
{
    EventTimer tm(b->stream);
    tm.start();
    TRYH(hipGraphLaunch(b->exec[stage], b->stream));
    *ms = tm.stop();
    TRYH(hipStreamSynchronize(b->stream));
    return PUMP_OK;
}

int new_instance(pump_batch *b, const pump_batch_instance &I, int n_mc, int tmax_steps)
{
    pump_roadmap *rm = new pump_roadmap();
    b->rms.push_back(rm);
    rm->in_batch = true;
    rm->owns_stream = false;
    rm->owns_noise = false;
    rm->stream = b->stream;
    rm->num = I.num;
    rm->n_obs = I.n_obs;
    rm->n_mc = n_mc;
    rm->tmax_steps = tmax_steps;
    rm->prm = I.prm;
    rm->prm.fused_edges = 1;
    std::memcpy(rm->goal, I.goal, sizeof(rm->goal));
    rm->samples_h.assign(I.samples, I.samples + (size_t)I.num * 6);
    rm->offsets_time = b->offsets_time;
    rm->xcomb = b->xcomb;
    rm->offsets_split = b->offsets_split;
    return rm_upload(rm, I.samples, I.obstacles, nullptr, nullptr);
}

// all ordered pairs of every instance: one persistent launch
int pair_pass(pump_batch *b, int autotune)
{
    const int n = (int)b->rms.size();
    hipStream_t s = b->stream;
    std::vector<PairJob> jobs(n);
    std::vector<int64_t> prefix(n + 1, 0);
    for (int k = 0; k < n; ++k) {
        const pump_roadmap *rm = b->rms[k];
        const DivMagic dm = rm->n_pairs < (1ll << 32) ? div_magic((uint32_t)(rm->num - 1))
                                                       : DivMagic{0, 0, 0};
        jobs[k] = PairJob{rm->samples, rm->copts, rm->topts, rm->num,  rm->prm.tmax,
                          dm.m,        dm.s1,     dm.s2};
        prefix[k + 1] = prefix[k] + rm->n_pairs;
    }
    const int64_t total = prefix[n];
    TRY(upload(&b->pair_jobs, jobs, s));
    TRY(upload(&b->pair_prefix, prefix, s));
    TRY(dalloc(&b->counter, 1));
    int bpc = 8;
    const TuneKey key{0, n, total, 0, 0};
    std::array<int, 2> cached{};
    if (autotune && tuned_get(key, &cached)) {
        bpc = kBpc[cached[0]];
        b->info.autotune_cached |= 1;
    } else if (autotune) {
        // the first kTunePairs pairs (rewritten later by the captured pass)
        const int64_t sample = std::min(total, kTunePairs);
        float best = 0;
        int pick = 1;
        EventTimer tm(s);
        TRY(fill_pairs_many(b->pair_jobs, b->pair_prefix, n, sample, bpc, b->counter, s));
        for (int c = 0; c < 3; ++c) {
            tm.start();
            TRY(fill_pairs_many(b->pair_jobs, b->pair_prefix, n, sample, kBpc[c], b->counter, s));
            const float ms = tm.stop();
            b->info.ms_autotune += ms;
            if (best == 0 || ms < best) {
                best = ms;
                pick = c;
            }
        }
        bpc = kBpc[pick];
        tuned_put(key, {pick, 0});
    }
    b->info.pairs_blocks_per_cu = bpc;
    TRY(capture(b, 0, [&] {
        return fill_pairs_many(b->pair_jobs, b->pair_prefix, n, total, bpc, b->counter, s);
    }));
    return launch_timed(b, 0, &b->info.ms_pairs);
}

// the fused edge pass and the half-space cache of every instance: two launches
int edge_pass(pump_batch *b, int autotune)
{
    const int n = (int)b->rms.size();
    hipStream_t s = b->stream;
    std::vector<EdgeJob> ejobs(n);
    std::vector<HsJob> hjobs(n);
    std::vector<int64_t> eprefix(3 * (size_t)(n + 1), 0), hprefix(n + 1, 0);
    for (int k = 0; k < n; ++k) {
        const pump_roadmap *rm = b->rms[k];
        ejobs[k] = EdgeJob{rm->samples,  rm->edge_src,      rm->edge_dst, rm->topts_edge,
                           rm->n_edges,  rm->wp_offsets,    0,            rm->obstacles,
                           rm->n_obs,    rm->disc,          rm->is_free_edges};
        hjobs[k] = HsJob{rm->n_edges, rm->disc,   rm->wp_offsets, 0,      rm->is_free_edges,
                         rm->n_obs,   rm->max_hs, rm->obstacles,  rm->as, rm->bs,
                         rm->count,   rm->maxd2,  rm->edge_hs_mask};
        for (int c = 0; c < 3; ++c) {
            int64_t *p = &eprefix[(size_t)c * (n + 1)];
            p[k + 1] = p[k] + (rm->n_edges + kEpb[c] - 1) / kEpb[c];
        }
        hprefix[k + 1] = hprefix[k] + (rm->n_edges + kHsEdgesPerBlock - 1) / kHsEdgesPerBlock;
    }
    TRY(upload(&b->edge_jobs, ejobs, s));
    TRY(upload(&b->hs_jobs, hjobs, s));
    TRY(upload(&b->edge_prefix, eprefix, s));
    TRY(upload(&b->hs_prefix, hprefix, s));
    int max_obs = 0, max_hs = 0;
    for (const pump_roadmap *rm : b->rms) {
        max_obs = std::max(max_obs, rm->n_obs);
        max_hs = std::max(max_hs, rm->max_hs);
    }
    auto blocks = [&](int c) { return eprefix[(size_t)c * (n + 1) + n]; };
    auto prefix_of = [&](int c) { return b->edge_prefix + (size_t)c * (n + 1); };
    // obstacle tiles: those below the largest box count, and the first that
    // holds every box (the default)
    int n_tiles = 0, tile_all = 0;
    for (int t : kObsTile) {
        ++n_tiles;
        if (t >= max_obs)
            break;
    }
    tile_all = n_tiles - 1;
    int pick = 1, tile = tile_all;
    int64_t n_wp = 0, n_edges = 0;
    for (const pump_roadmap *rm : b->rms) {
        n_wp += rm->n_wp;
        n_edges += rm->n_edges;
    }
    const TuneKey key{1, n, n_edges, n_wp, max_obs};
    std::array<int, 2> cached{};
    if (autotune && tuned_get(key, &cached)) {
        pick = cached[0];
        tile = cached[1];
        b->info.autotune_cached |= 2;
    } else if (autotune) {
        // every candidate writes the same waypoints and flags
        float best = 0;
        EventTimer tm(s);
        TRY(edges_many(b->edge_jobs, prefix_of(1), n, blocks(1), kEpb[1], max_obs,
                       kObsTile[tile_all], s));
        for (int c = 0; c < 3; ++c)
            for (int ti = 0; ti < n_tiles; ++ti) {
                tm.start();
                TRY(edges_many(b->edge_jobs, prefix_of(c), n, blocks(c), kEpb[c], max_obs,
                               kObsTile[ti], s));
                const float ms = tm.stop();
                b->info.ms_autotune += ms;
                if (best == 0 || ms < best) {
                    best = ms;
                    pick = c;
                    tile = ti;
                }
            }
        tuned_put(key, {pick, tile});
    }
    b->info.edges_per_block = kEpb[pick];
    b->info.edge_obs_tile = std::min(kObsTile[tile], std::max(max_obs, 1));
    const int64_t hs_blocks = hprefix[n];
    TRY(capture(b, 1, [&] {
        int r = edges_many(b->edge_jobs, prefix_of(pick), n, blocks(pick), kEpb[pick], max_obs,
                           kObsTile[tile], s);
        if (r == PUMP_OK)
            r = cache_hs_many(b->hs_jobs, b->hs_prefix, n, hs_blocks, max_obs, max_hs, s);
        return r;
    }));
    return launch_timed(b, 1, &b->info.ms_edges);
}

// rn, edge numbering, neighbour tables and waypoint offsets of every
// instance.  Each instance's stages are host-synchronous (a stage reads back a
// count that sizes the next allocation), so instances run side by side on
// kTableThreads host threads, each with its own stream (default priority);
// an instance keeps the batch stream outside this call.
constexpr int kTableThreads = 8;

int tables_all(pump_batch *b)
{
    const int n = (int)b->rms.size();
    if (n == 0)
        return PUMP_OK;
    TRYH(hipStreamSynchronize(b->stream)); // the pair tables are complete
    const int T = std::min(kTableThreads, n);
    int dev = 0;
    TRYH(hipGetDevice(&dev));
    std::vector<hipStream_t> st(T, nullptr);
    int status = PUMP_OK;
    for (int j = 0; j < T && status == PUMP_OK; ++j) {
        const hipError_t e = hipStreamCreate(&st[j]);
        if (e != hipSuccess) {
            st[j] = nullptr;
            note_hip_error(e, __FILE__, __LINE__);
            status = PUMP_E_HIP;
        }
    }
    std::vector<int> rs(T, PUMP_OK);
    if (status == PUMP_OK) {
        std::vector<std::thread> th;
        for (int j = 0; j < T; ++j)
            th.emplace_back([&, j] {
                if (hipSetDevice(dev) != hipSuccess) {
                    rs[j] = PUMP_E_HIP;
                    return;
                }
                for (int k = j; k < n && rs[j] == PUMP_OK; k += T) {
                    pump_roadmap *rm = b->rms[k];
                    rm->stream = st[j];
                    rs[j] = rm_tables(rm);
                    if (rs[j] == PUMP_OK && hipStreamSynchronize(st[j]) != hipSuccess)
                        rs[j] = PUMP_E_HIP;
                    rm->stream = b->stream;
                }
            });
        for (std::thread &t : th)
            t.join();
        for (int j = 0; j < T; ++j)
            if (status == PUMP_OK && rs[j] != PUMP_OK)
                status = rs[j];
    }
    for (hipStream_t x : st)
        if (x)
            hipStreamDestroy(x);
    return status;
}

int build(pump_batch *b, int n, const pump_batch_instance *inst, const float *offsets_time_h,
          const float *xcomb_h, int n_mc, int tmax_steps, int autotune)
{
    const auto t_begin = std::chrono::steady_clock::now();
    hipStream_t s = b->stream;
    if (offsets_time_h && xcomb_h && n_mc > 0) {
        // one copy of the noise tables, shared by every instance
        const size_t nt = (size_t)tmax_steps * n_mc * 3;
        TRY(dalloc(&b->offsets_time, nt));
        TRY(dalloc(&b->xcomb, nt));
        TRYH(hipMemcpyAsync(b->offsets_time, offsets_time_h, sizeof(float) * nt,
                            hipMemcpyHostToDevice, s));
        TRYH(hipMemcpyAsync(b->xcomb, xcomb_h, sizeof(float) * nt, hipMemcpyHostToDevice, s));
        TRY(dalloc(&b->offsets_split,
                   pump_split_offsets_bytes(tmax_steps, n_mc) / sizeof(uint16_t)));
        TRY(pump_split_offsets(b->offsets_time, tmax_steps, n_mc, b->offsets_split, s));
    }
    for (int k = 0; k < n; ++k)
        TRY(new_instance(b, inst[k], n_mc, tmax_steps));
    TRY(pair_pass(b, autotune));
    TRY(tables_all(b));
    bool pairs_kept = true;
    for (pump_roadmap *rm : b->rms)
        pairs_kept = pairs_kept && rm->prm.keep_pairs;
    if (!pairs_kept) {
        // the captured pair pass would write freed buffers
        hipGraphExecDestroy(b->exec[0]);
        hipGraphDestroy(b->graph[0]);
        b->exec[0] = nullptr;
        b->graph[0] = nullptr;
    }
    for (pump_roadmap *rm : b->rms)
        TRY(rm_alloc_halfspaces(rm));
    TRY(edge_pass(b, autotune));
    pump_batch_info &I = b->info;
    I.n_inst = n;
    for (pump_roadmap *rm : b->rms) {
        TRY(rm_finish(rm, t_begin));
        I.n_pairs += rm->n_pairs;
        I.n_edges += rm->n_edges;
        I.n_waypoints += rm->n_wp;
    }
    I.ms_total = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() -
                                                          t_begin)
                     .count();
    return PUMP_OK;
}

bool valid_instance(const pump_batch_instance &I)
{
    return I.samples && I.num >= 2 && I.n_obs >= 0 && (I.n_obs == 0 || I.obstacles) && I.goal;
}

} // namespace

extern "C" void pump_batch_free(pump_batch *b)
{
    if (!b)
        return;
    if (b->stream)
        hipStreamSynchronize(b->stream);
    for (int i = 0; i < 3; ++i) {
        if (b->exec[i])
            hipGraphExecDestroy(b->exec[i]);
        if (b->graph[i])
            hipGraphDestroy(b->graph[i]);
    }
    for (hipStream_t st : b->side)
        hipStreamDestroy(st);
    for (hipEvent_t e : b->ev)
        hipEventDestroy(e);
    dfree(b->mc_jobs);
    if (b->mc_maps)
        hipFree(b->mc_maps);
    for (BatchMc &m : b->mc) {
        dfree(m.edge);
        dfree(m.prev);
        dfree(m.row_src);
        dfree(m.times);
        dfree(m.parent_cp);
        dfree(m.cp);
        dfree(m.src);
        dfree(m.dst);
        dfree(m.ws);
    }
    for (pump_roadmap *rm : b->rms)
        rm_release(rm);
    dfree(b->offsets_time);
    dfree(b->xcomb);
    dfree(b->offsets_split);
    dfree(b->pair_jobs);
    dfree(b->pair_prefix);
    dfree(b->counter);
    dfree(b->edge_jobs);
    dfree(b->edge_prefix);
    dfree(b->hs_jobs);
    dfree(b->hs_prefix);
    if (b->stream)
        hipStreamDestroy(b->stream);
    delete b;
}

extern "C" int pump_batch_build(int n_inst, const pump_batch_instance *inst,
                                const float *offsets_time_h, const float *xcomb_h, int n_mc,
                                int tmax_steps, int autotune, pump_batch **out)
{
    if (!out)
        return PUMP_E_INVALID;
    *out = nullptr;
    if (n_inst < 1 || !inst || n_mc < 0 || tmax_steps < 0)
        return PUMP_E_INVALID;
    for (int k = 0; k < n_inst; ++k)
        if (!valid_instance(inst[k]))
            return PUMP_E_INVALID;
    pump_batch *b = new pump_batch();
    const hipError_t ce = hipStreamCreate(&b->stream);
    if (ce != hipSuccess) {
        note_hip_error(ce, __FILE__, __LINE__);
        delete b;
        return PUMP_E_HIP;
    }
    const int r = build(b, n_inst, inst, offsets_time_h, xcomb_h, n_mc, tmax_steps, autotune);
    if (r != PUMP_OK) {
        pump_batch_free(b);
        return r;
    }
    *out = b;
    return PUMP_OK;
}

extern "C" int pump_batch_info_get(const pump_batch *b, pump_batch_info *info)
{
    if (!b || !info)
        return PUMP_E_INVALID;
    *info = b->info;
    return PUMP_OK;
}

extern "C" int pump_batch_roadmap(pump_batch *b, int k, pump_roadmap **rm)
{
    if (!b || !rm || k < 0 || k >= (int)b->rms.size())
        return PUMP_E_INVALID;
    *rm = b->rms[k];
    return PUMP_OK;
}

extern "C" int pump_batch_replay(pump_batch *b, int stage, int reps, float *ms)
{
    if (!b || stage < 0 || stage > 2 || reps < 1 || !b->exec[stage])
        return PUMP_E_INVALID;
    EventTimer tm(b->stream);
    tm.start();
    for (int i = 0; i < reps; ++i)
        TRYH(hipGraphLaunch(b->exec[stage], b->stream));
    const float t = tm.stop();
    TRYH(hipStreamSynchronize(b->stream));
    if (ms)
        *ms = t / reps;
    return PUMP_OK;
}

namespace {

// 64-bit splitmix step: the parent times of the batch's MC wavefronts
uint64_t splitmix(uint64_t &x)
{
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

constexpr int kMcStreams = 8; // side streams the MC stage spreads its instances over

// The MC step (copy-on-write, bit validity, FP32 tests) of every instance
// with a wavefront.  The steps are independent, so they fork from the batch
// stream onto the side streams (instance k on stream k mod kMcStreams) and
// join back: kernels of different instances overlap on the chip, and under
// capture the graph gets one branch per side stream.
int mc_launch(pump_batch *b)
{
    if (b->mc_many)
        return mc_cow_many(b->mc_jobs, b->mc_grids, b->stream);
    const int S = (int)b->side.size();
    TRYH(hipEventRecord(b->ev[0], b->stream));
    for (int j = 0; j < S; ++j)
        TRYH(hipStreamWaitEvent(b->side[j], b->ev[0], 0));
    for (size_t k = 0; k < b->rms.size(); ++k) {
        const BatchMc &m = b->mc[k];
        if (m.n_new == 0)
            continue;
        const pump_roadmap *rm = b->rms[k];
        TRY(pump_propagate_cp_cow(
            1, m.n_new, 0, rm->wp_offsets, m.edge, m.prev, rm->as, rm->bs, rm->edge_hs_mask,
            m.times, rm->topts_edge, rm->prm.dt, rm->max_hs, b->offsets_time, rm->tmax_steps,
            rm->prm.offset_mult, rm->n_mc, m.src, m.parent_cp, m.dst, 0, m.row_src, rm->prm.n_p,
            m.cp, 0, 0, nullptr, nullptr, nullptr, m.ws, m.ws_bytes,
            (pump_stream_t)b->side[k % (size_t)S]));
    }
    for (int j = 0; j < S; ++j) {
        TRYH(hipEventRecord(b->ev[1 + j], b->side[j]));
        TRYH(hipStreamWaitEvent(b->stream, b->ev[1 + j], 0));
    }
    return PUMP_OK;
}

} // namespace

extern "C" int pump_batch_mc_build(pump_batch *b, int n_parents, int max_paths, uint64_t seed,
                                   float *ms)
{
    if (!b || n_parents < 1 || max_paths < 1 || !b->offsets_time || !b->mc.empty())
        return PUMP_E_INVALID;
    hipStream_t s = b->stream;
    for (int j = 0; j < kMcStreams; ++j) {
        hipStream_t st = nullptr;
        TRYH(hipStreamCreate(&st));
        b->side.push_back(st);
    }
    for (int j = 0; j <= kMcStreams; ++j) {
        hipEvent_t e = nullptr;
        TRYH(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        b->ev.push_back(e);
    }
    b->mc.resize(b->rms.size());
    b->mc_parents = n_parents;
    std::vector<uint8_t> free_h;
    for (size_t k = 0; k < b->rms.size(); ++k) {
        const pump_roadmap *rm = b->rms[k];
        BatchMc &m = b->mc[k];
        if (!rm->as || !rm->is_free_edges || rm->prm.n_p <= 0 || rm->n_mc <= 0 ||
            rm->prm.n_p > rm->n_mc)
            return PUMP_E_INVALID;
        // the wavefront: the instance's free edges in edge order, child i
        // from parent i % n_parents
        free_h.resize((size_t)rm->n_edges);
        TRYH(hipMemcpyAsync(free_h.data(), rm->is_free_edges, free_h.size(), hipMemcpyDeviceToHost,
                            s));
        TRYH(hipStreamSynchronize(s));
        std::vector<int> edge, prev;
        for (int64_t e = 0; e < rm->n_edges && (int)edge.size() < max_paths; ++e)
            if (free_h[(size_t)e]) {
                prev.push_back((int)(edge.size() % (size_t)n_parents));
                edge.push_back((int)e);
            }
        std::vector<float> times((size_t)n_parents);
        uint64_t x = seed ^ (0xD1B54A32D192ED03ull * (k + 1));
        for (float &t : times)
            t = (float)(2.0 * (double)(splitmix(x) >> 11) * 0x1p-53);
        m.n_new = (int)edge.size();
        const int nw = (rm->prm.n_p + 63) / 64;
        TRY(upload(&m.times, times, s));
        TRY(dalloc(&m.parent_cp, (size_t)n_parents));
        TRYH(hipMemsetAsync(m.parent_cp, 0, sizeof(float) * n_parents, s));
        TRY(dalloc(&m.src, (size_t)n_parents * nw));
        TRY(pump_setup_valid_bits(m.src, rm->prm.n_p, n_parents, (pump_stream_t)s));
        if (m.n_new == 0) {
            TRYH(hipStreamSynchronize(s));
            continue;
        }
        TRY(upload(&m.edge, edge, s));
        TRY(upload(&m.prev, prev, s));
        TRY(dalloc(&m.row_src, (size_t)m.n_new));
        TRY(dalloc(&m.cp, (size_t)m.n_new));
        TRY(dalloc(&m.dst, (size_t)m.n_new * nw));
        m.ws_bytes = pump_propagate_workspace(m.n_new, rm->tmax_steps);
        TRY(dalloc(&m.ws, m.ws_bytes));
        // the host vectors end with this iteration: finish their uploads
        TRYH(hipStreamSynchronize(s));
    }
    TRYH(hipStreamSynchronize(s));
    // every problem's step in one launch per stage, where each takes the
    // binned form; else one step per problem on the side streams
    std::vector<McJob> jobs;
    bool many = true;
    for (size_t k = 0; k < b->rms.size() && many; ++k) {
        const BatchMc &m = b->mc[k];
        if (m.n_new == 0)
            continue;
        const pump_roadmap *rm = b->rms[k];
        McJob J;
        many = mc_cow_job(J, m.n_new, rm->wp_offsets, m.edge, m.prev, rm->as, rm->bs,
                          rm->edge_hs_mask, m.times, rm->topts_edge, rm->prm.dt, rm->max_hs,
                          b->offsets_time, rm->tmax_steps, rm->prm.offset_mult, rm->n_mc, m.src,
                          m.parent_cp, m.dst, m.row_src, rm->prm.n_p, m.cp, m.ws,
                          m.ws_bytes) == PUMP_OK;
        if (many)
            jobs.push_back(J);
    }
    if (many && !jobs.empty()) {
        // particles per wave of the one-launch form: the same for every
        // problem, the fastest of 64, 128, 256 and 512 for the whole step
        // (a warm-up pass, then two timed passes each)
        TRY(dalloc(&b->mc_jobs, jobs.size()));
        const auto setup = [&](int nq) -> int {
            for (McJob &J : jobs)
                J.B.nq = nq;
            if (b->mc_maps)
                hipFree(b->mc_maps);
            b->mc_maps = nullptr;
            TRY(mc_plan(jobs.data(), (int)jobs.size(), b->mc_grids, &b->mc_maps, s));
            TRYH(hipMemcpyAsync(b->mc_jobs, jobs.data(), jobs.size() * sizeof(McJob),
                                hipMemcpyHostToDevice, s));
            TRYH(hipStreamSynchronize(s));
            return PUMP_OK;
        };
        int best = 4;
        float best_ms = 0;
        for (int nq : {1, 2, 4, 8}) {
            TRY(setup(nq));
            TRY(mc_cow_many(b->mc_jobs, b->mc_grids, s));
            EventTimer tm(s);
            tm.start();
            for (int r = 0; r < 2; ++r)
                TRY(mc_cow_many(b->mc_jobs, b->mc_grids, s));
            const float t = tm.stop();
            if (nq == 1 || t < best_ms) {
                best = nq;
                best_ms = t;
            }
        }
        TRY(setup(best));
        b->mc_many = true;
        b->info.mc_one_launch = 1;
        b->info.mc_particles_per_wave = 64 * best;
    } else {
        // particles per wave, tuned once per MC configuration (particle
        // count, half-space slots) on the first instance that has it
        std::vector<std::array<int, 2>> tuned;
        for (size_t k = 0; k < b->rms.size(); ++k) {
            const BatchMc &m = b->mc[k];
            const pump_roadmap *rm = b->rms[k];
            const std::array<int, 2> key{rm->prm.n_p, rm->max_hs};
            if (m.n_new < 4096 || std::find(tuned.begin(), tuned.end(), key) != tuned.end())
                continue;
            tuned.push_back(key);
            int ppw = 0;
            TRY(pump_hsmc_autotune_cow(
                1, m.n_new, 0, rm->wp_offsets, m.edge, m.prev, rm->as, rm->bs, rm->edge_hs_mask,
                m.times, rm->topts_edge, rm->prm.dt, rm->max_hs, b->offsets_time, rm->tmax_steps,
                rm->prm.offset_mult, rm->n_mc, m.src, m.parent_cp, m.dst, 0, m.row_src,
                rm->prm.n_p, m.cp, 0, 0, nullptr, nullptr, nullptr, m.ws, m.ws_bytes,
                (pump_stream_t)s, &ppw));
        }
    }
    TRY(capture(b, 2, [&] { return mc_launch(b); }));
    float t = 0;
    TRY(launch_timed(b, 2, &t));
    if (ms)
        *ms = t;
    return PUMP_OK;
}

extern "C" int pump_batch_mc_eager(pump_batch *b, int reps, float *ms)
{
    if (!b || reps < 1 || b->mc.empty())
        return PUMP_E_INVALID;
    EventTimer tm(b->stream);
    tm.start();
    for (int i = 0; i < reps; ++i)
        TRY(mc_launch(b));
    const float t = tm.stop();
    TRYH(hipStreamSynchronize(b->stream));
    if (ms)
        *ms = t / reps;
    return PUMP_OK;
}

extern "C" int pump_batch_mc_get(pump_batch *b, int k, int cap, int *n_new, int *edge, int *prev,
                                 float *times, float *cp, int *row_src)
{
    if (!b || k < 0 || k >= (int)b->mc.size() || cap < 0 || !n_new)
        return PUMP_E_INVALID;
    const BatchMc &m = b->mc[(size_t)k];
    *n_new = m.n_new;
    const size_t n = (size_t)std::min(cap, m.n_new);
    hipStream_t s = b->stream;
    TRYH(hipStreamSynchronize(s));
    if (edge && n)
        TRYH(hipMemcpy(edge, m.edge, n * sizeof(int), hipMemcpyDeviceToHost));
    if (prev && n)
        TRYH(hipMemcpy(prev, m.prev, n * sizeof(int), hipMemcpyDeviceToHost));
    if (cp && n)
        TRYH(hipMemcpy(cp, m.cp, n * sizeof(float), hipMemcpyDeviceToHost));
    if (row_src && n)
        TRYH(hipMemcpy(row_src, m.row_src, n * sizeof(int), hipMemcpyDeviceToHost));
    if (times)
        TRYH(hipMemcpy(times, m.times, (size_t)std::min(cap, b->mc_parents) * sizeof(float),
                       hipMemcpyDeviceToHost));
    return PUMP_OK;
}
