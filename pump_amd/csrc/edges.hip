// edges.hip -- kernel (b): edge waypoint generation fused with the
// segment/AABB test (reference: findDiscretizedPath 2pBVP.cu:53-69 driven by
// the host loop mainPUMP.cu:327-346, then freeEdges collisionCheck.cu:5-37),
// plus the reference-shaped freeEdges and sampleFree (sampler.cu:61-88).
//
// Layout of the fused kernel: a workgroup owns a run of consecutive edges;
// its per-edge data (pathPoint constants, candidate boxes) and the obstacle
// boxes (structure of arrays) are staged once in LDS.  Each wave then walks
// tiles of 64 waypoints, one per lane, with no further workgroup barrier:
// the waypoints go to discMotions (consecutive lanes write consecutive
// 24-byte records), and segment (k, k+1) is tested by the lane holding
// waypoint k, with the far end taken from the next lane, against the boxes
// that are not separated from its edge's bounding box.
#include <algorithm>
#include <cstdlib>
#include <type_traits>

#include "pump_device.hpp"
#include "pump_internal.hpp"

#pragma clang fp contract(off)

namespace pump {
namespace {

constexpr int kEdgesPerBlock = 128;
constexpr int kTile = 256;

constexpr int kEdgeObs = 256;                  // boxes staged in LDS by the edge pass
constexpr int kEdgeMaskWords = kEdgeObs / 64;  // candidate-mask words per edge
constexpr int kMapCap = 2016;                  // block waypoints mapped to their edge in LDS

// Per-edge record of an edge-pass workgroup in LDS: everything a lane needs
// to evaluate any waypoint of the edge, read from one base address.
struct alignas(16) EdgeRec {
    EdgePoly E;         // pathPoint constants
    double t3, r2, r3;  // tau^3, RN(1/tau^2), RN(1/tau^3)
    float tau, dt;      // arrival time, waypoint spacing
    float x0[kDim];     // start state
    float x1[kDim];     // end state
};
static_assert(sizeof(EdgeRec) == 144, "edge record layout");

// A staged box: lower corner, upper corner (positions), each padded to 16
// bytes so that a box is two 16-byte LDS reads.
struct alignas(16) BoxRec {
    float lo[3], pad_lo;
    float hi[3], pad_hi;
};

// Per-edge record of the FP32-waypoint form: per position axis the cubic
// Hermite segment p0 + v0 t + c2 t^2 + c3 t^3 with FP32 coefficients, the
// margin m of its waypoints' distance from the reference's, and what an
// exact re-test of a segment needs (tau, spacing, end states).
struct alignas(16) EdgeRec32 {
    float c[3][4];   // p0, v0, c2, c3 per axis (p0, v0: the start state)
    float x1[kDim];  // end state
    float m[3];      // per axis margin: 1.1 x the waypoint error bound eps
    float tau, dt;   // arrival time, waypoint spacing
    float pad;
};
static_assert(sizeof(EdgeRec32) == 96, "FP32-form edge record layout");

// The record of the FP32 form.  In FP64, from r = RN(1/tau):
// c2 = 3(p1-p0) r^2 - (2v0+v1) r, c3 = 2(p0-p1) r^3 + (v0+v1) r^2, rounded
// to FP32 (the FP64 products are within 1e-15 of the exact coefficients, far
// inside one FP32 rounding).  With S = |p0| + |v0| tau + |c2| tau^2 +
// |c3| tau^3, the FMA Horner evaluation at t <= tau (three roundings) and
// the rounded coefficients are within 4.1 u S of the exact segment
// (u = 2^-24), and the reference's waypoint (its FP64 terms, 2pBVP.cu:117-131,
// rounded once to FP32) within 1.01 u S; eps = 12 u S covers both.  The
// stored margin m = 1.1 eps = eps + 1.2 u S also covers one FP32 rounding of
// a sum of a waypoint coordinate (|x| <= S (1 + 5u)) and the margin, which is
// what the separation and overlap tests below need.  A NaN or infinite tau
// gives a NaN or infinite margin, which leaves every segment of the edge
// undecided, for the exact re-test.
__device__ __forceinline__ void edge_rec32(EdgeRec32 &R, const float *x0, const float *x1,
                                           float tau, int W)
{
    const double td = (double)tau;
    const double r1 = 1.0 / td, r2 = r1 * r1, r3 = r2 * r1;
    for (int d = 0; d < 3; ++d) {
        const double p0 = x0[d], v0 = x0[d + 3], p1 = x1[d], v1 = x1[d + 3];
        const double c2 = 3.0 * (p1 - p0) * r2 - (2.0 * v0 + v1) * r1;
        const double c3 = 2.0 * (p0 - p1) * r3 + (v0 + v1) * r2;
        R.c[d][0] = x0[d];
        R.c[d][1] = x0[d + 3];
        R.c[d][2] = (float)c2;
        R.c[d][3] = (float)c3;
        const double S = fabs(p0) + fabs(v0) * td + fabs(c2) * td * td + fabs(c3) * td * td * td;
        // an FP32 coefficient that overflowed (tau far below any real
        // arrival time) makes the axis's waypoints meaningless: an infinite
        // margin leaves its segments undecided
        const bool finite = fabsf(R.c[d][2]) < INFINITY && fabsf(R.c[d][3]) < INFINITY;
        R.m[d] = finite ? (float)(1.1 * 12.0 * 0x1p-24 * S * 1.0001) : INFINITY;
    }
    for (int d = 0; d < kDim; ++d)
        R.x1[d] = x1[d];
    R.tau = tau;
    R.dt = tau / (float)(W - 1);
    R.pad = 0.0f;
}

// Waypoint k of an edge in the FP32 form: the end states exactly, interior
// points by Horner's rule at the reference's t = dt * k.  The start state is
// the record's p0 and v0 (no separate read); every lane reads the end
// state, so that no read sits in a divergent branch.
__device__ __forceinline__ void waypoint32(const EdgeRec32 &R, int k, int W, float *x)
{
    const float t = R.dt * (float)k;
    const bool first = k == 0, last = k == W - 1;
    float e1[kDim];
    for (int d = 0; d < kDim; ++d)
        e1[d] = R.x1[d];
    for (int d = 0; d < 3; ++d) {
        const float4 c = *reinterpret_cast<const float4 *>(R.c[d]);
        const float p = fmaf(t, fmaf(t, fmaf(t, c.w, c.z), c.y), c.x);
        const float v = fmaf(t, fmaf(t, 3.0f * c.w, 2.0f * c.z), c.y);
        x[d] = last ? e1[d] : first ? c.x : p;
        x[d + 3] = last ? e1[d + 3] : first ? c.y : v;
    }
}

// The bounding box of segment v -> w widened by the margin, for the
// reference's broad phase (collisionCheck.cu:104-112: separated iff
// bbMax <= lo or hi <= bbMin on some axis).  up = RN(max + m) and
// dn = RN(min - m): the reference's segment, whose ends lie within eps of
// v and w per axis, has bbMax <= max + eps <= up and bbMin >= dn.
struct SegBox {
    float up[3], dn[3];
};

__device__ __forceinline__ SegBox seg_box32(const float *v, const float *w, const float *m)
{
    SegBox s;
    for (int d = 0; d < 3; ++d) {
        const bool gt = v[d] > w[d];
        s.up[d] = (gt ? v[d] : w[d]) + m[d];
        s.dn[d] = (gt ? w[d] : v[d]) - m[d];
    }
    return s;
}

// the reference's broad phase rejects the box for every segment within the
// margin of v -> w.  A NaN coordinate comes only from a non-finite
// coefficient of its axis, whose margin is then infinite: its bounds are
// NaN or infinite and reject nothing.  Branch-free: up <= lo iff
// lo - up >= 0 (IEEE subtraction keeps the sign of the exact difference,
// f32 denormals are kept), and fmaxf passes over a NaN term, so an axis
// whose bounds are NaN neither rejects nor hides the other axes' verdicts.
__device__ __forceinline__ bool surely_separated(const SegBox &s, const float *lo,
                                                 const float *hi)
{
    const float g0 = fmaxf(lo[0] - s.up[0], s.dn[0] - hi[0]);
    const float g1 = fmaxf(lo[1] - s.up[1], s.dn[1] - hi[1]);
    const float g2 = fmaxf(lo[2] - s.up[2], s.dn[2] - hi[2]);
    return fmaxf(fmaxf(g0, g1), g2) >= 0.0f;
}

constexpr int kMiss = 0, kHit = 1, kUndecided = 2;

// The reference's segment/box test (collisionCheck.cu:104-143) decided for
// every segment v' -> w' with |v' - v|, |w' - w| <= eps per axis, evaluated
// in FP32 without contraction as the reference does: hit iff the broad phase
// overlaps and, for some axis d, the point of the line on the plane
// P = (v'_d < lo_d ? lo_d : hi_d), at lambda' = (P - v'_d) / (w'_d - v'_d),
// lies in [lo_o, hi_o] on both other axes o.  Returns kHit or kMiss when
// every such segment gives that answer, else kUndecided.
//   * Plane: v_d + m < lo_d gives v'_d < lo_d, v_d - m >= lo_d gives
//     v'_d >= lo_d (m covers eps and the rounding of the sum).
//   * lambda: with A = P - v_d, B = w_d - v_d and |B| > 4.1 m_d, the exact
//     quotient for v', w' is within D1 = 2.1 m_d (1 + 2.01 L) / |B| of
//     Lambda = A / B (L = |lambda| computed here), and the reference's
//     rounded lambda' and the one computed here (sub, rcp within 1 ulp, mul)
//     within DL = 1.01 D1 + 9 u L of each other.
//   * Projection on axis o: |proj' - proj| <= m_o (1 + 2 L') + G_o DL +
//     8 u (|v_o| + (G_o + 2 m_o) L') with G_o = |w_o - v_o|, L' = L + DL
//     (both the reference's and this rounding of the sub, mul and add);
//     the bound is scaled by 1.05 to cover its own FP32 evaluation and the
//     rounding of the comparisons' differences.
// Any NaN or infinity fails every comparison that decides, so such a
// segment is undecided.
__device__ __forceinline__ int classify32(const float *v, const float *w, const float *m,
                                          const float *lo, const float *hi)
{
    bool overlap = true; // the reference's broad phase surely overlaps
    for (int d = 0; d < 3; ++d) {
        const float mx = fmaxf(v[d], w[d]), mn = fminf(v[d], w[d]);
        overlap = overlap && (lo[d] < mx - m[d]) && (mn + m[d] < hi[d]);
    }
    bool any_in = false, all_out = true;
    for (int d = 0; d < 3; ++d) {
        const bool below = v[d] + m[d] < lo[d];
        const bool above = v[d] - m[d] >= lo[d];
        const float P = below ? lo[d] : hi[d];
        const float B = w[d] - v[d];
        const float lam = (P - v[d]) * __builtin_amdgcn_rcpf(B);
        const float L = fabsf(lam);
        const float D1 = 2.1f * m[d] * (1.0f + 2.01f * L) / fabsf(B);
        const float DL = 1.01f * D1 + 9.0f * 0x1p-24f * L;
        const float Lp = L + DL;
        const bool ok = (below || above) && (fabsf(B) > 4.1f * m[d]);
        bool in = ok, out = false;
        for (int o = 0; o < 3; ++o) {
            if (o == d)
                continue;
            const float G = fabsf(w[o] - v[o]);
            const float proj = v[o] + (w[o] - v[o]) * lam;
            const float delta =
                1.05f * (m[o] * (1.0f + 2.0f * Lp) + G * DL +
                         8.0f * 0x1p-24f * (fabsf(v[o]) + (G + 2.0f * m[o]) * Lp));
            in = in && (proj - lo[o] > delta) && (hi[o] - proj > delta);
            out = out || (lo[o] - proj > delta) || (proj - hi[o] > delta);
        }
        any_in = any_in || in;
        all_out = all_out && ok && out;
    }
    if (overlap && any_in)
        return kHit;
    return all_out ? kMiss : kUndecided;
}

// LDS of an edge-pass workgroup, sized at launch: EPB edge records,
// candidate masks of words_cap words per edge and obs_cap staged boxes (so a
// workload with few boxes keeps more workgroups resident).  At 128 edges and
// a few dozen boxes seven workgroups fit the 160 KiB of a CU.
struct EdgeLds {
    int recs, emask, loff, ok, obs, wmap, rlist, rcnt, bytes;
};

// segments of a workgroup set aside for the exact re-test of the FP32 form
constexpr int kRetestCap = 512;

__host__ __device__ inline EdgeLds edge_lds(int epb, int obs_cap, int words_cap,
                                            int rec_bytes = (int)sizeof(EdgeRec),
                                            int retest_cap = 0)
{
    EdgeLds L{};
    int o = 0;
    auto take = [&o](int bytes, int align) {
        o = (o + align - 1) / align * align;
        const int at = o;
        o += bytes;
        return at;
    };
    L.recs = take(epb * rec_bytes, 16);
    L.obs = take(obs_cap * (int)sizeof(BoxRec), 16);
    L.emask = take(epb * words_cap * 8, 8);
    L.loff = take((epb + 1) * 4, 4);
    L.ok = take(epb, 4);
    L.wmap = take(kMapCap, 4);
    L.rlist = take(retest_cap * 4, 4);
    L.rcnt = take(4, 4);
    L.bytes = (o + 15) / 16 * 16;
    return L;
}

__host__ __device__ inline int edge_obs_cap(int n_obs) { return n_obs < kEdgeObs ? n_obs : kEdgeObs; }
__host__ __device__ inline int edge_words(int obs_cap) { return (obs_cap + 63) >> 6; }

// index q of the edge with loff[q] <= g < loff[q + 1] (local waypoint index)
__device__ __forceinline__ int find_edge_local(const int *loff, int ne, int g)
{
    int lo = 0, hi = ne;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (loff[mid] <= g)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Candidate boxes of one edge per lane (the mask part of an edge-pass
// workgroup): bit b of word wd set when staged box wd*64+b is not separated
// from the edge's waypoint bounding box (the reference's <= test,
// collisionCheck.cu:104-112).  Called by whole waves; own: the lane has an
// edge.
__device__ __forceinline__ void candidate_masks(bool own, const float *x0, const float *x1,
                                                float tau, const float *__restrict__ obstacles,
                                                int n_masked, int n_words,
                                                unsigned long long *__restrict__ emask_row)
{
    const int lane = threadIdx.x & 63;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (own)
        edge_bbox(x0, x1, tau, lo, hi);
    // The wave's union of its edges' boxes: a box separated from the
    // union (the same <= test) is separated from each of them, so only
    // the boxes that meet the union are tested edge by edge, with their
    // bounds read from the lane that loaded them.
    float ulo[3], uhi[3];
    for (int d = 0; d < 3; ++d) {
        ulo[d] = wave_min(lo[d]);
        uhi[d] = wave_max(hi[d]);
    }
    for (int wd = 0; wd < n_words; ++wd) {
        const int bl = wd * 64 + lane;
        float bx[6] = {0, 0, 0, 0, 0, 0}; // lo x, y, z, hi x, y, z of box bl
        bool meets = false;
        if (bl < n_masked) {
            for (int d = 0; d < 3; ++d) {
                bx[d] = obstacles[bl * 12 + d];
                bx[3 + d] = obstacles[bl * 12 + 6 + d];
            }
            meets = !((uhi[0] <= bx[0]) | (bx[3] <= ulo[0]) | (uhi[1] <= bx[1]) |
                      (bx[4] <= ulo[1]) | (uhi[2] <= bx[2]) | (bx[5] <= ulo[2]));
        }
        unsigned long long cand = __ballot(meets);
        unsigned long long m = 0ull;
        // four candidates per step (the last step repeats its first
        // index where fewer remain), so that their cross-lane reads and
        // tests overlap
        while (cand) {
            int idx[4];
            for (int u = 0; u < 4; ++u) {
                idx[u] = cand ? __builtin_ctzll(cand) : idx[0];
                cand &= cand - 1;
            }
            for (int u = 0; u < 4; ++u) {
                float c[6];
                for (int d = 0; d < 6; ++d)
                    c[d] = read_lane(bx[d], idx[u]);
                const bool sep = (hi[0] <= c[0]) | (c[3] <= lo[0]) | (hi[1] <= c[1]) |
                                 (c[4] <= lo[1]) | (hi[2] <= c[2]) | (c[5] <= lo[2]);
                m |= (unsigned long long)(!sep) << idx[u];
            }
        }
        if (own)
            emask_row[wd] = m;
    }
}

// Phase 1 of the FP32 form's decision for segment v -> w of an edge: true
// when some candidate box is not separated from it by the margin-widened
// broad phase (a few segments in a hundred).  Only those are decided further.
__device__ __forceinline__ bool seg_near32(const float *v, const float *w, const EdgeRec32 &R,
                                           const unsigned long long *emask_row, int n_words,
                                           const BoxRec *boxes, const float *__restrict__ obstacles,
                                           int n_masked, int n_obs)
{
    const SegBox sb = seg_box32(v, w, R.m);
    bool near = false;
    for (int wd = 0; wd < n_words && !near; ++wd) {
        unsigned long long mk = emask_row[wd];
        while (mk && !near) {
            const BoxRec B = boxes[wd * 64 + __builtin_ctzll(mk)];
            mk &= mk - 1;
            near = !surely_separated(sb, B.lo, B.hi);
        }
    }
    // boxes past the staged ones
    for (int b = n_masked; b < n_obs && !near; ++b) {
        const float *ob = obstacles + (int64_t)b * 12;
        const float lo[3] = {ob[0], ob[1], ob[2]};
        const float hi[3] = {ob[6], ob[7], ob[8]};
        near = !surely_separated(sb, lo, hi);
    }
    return near;
}

// Phase 2: kHit when the reference's test surely hits one of the edge's
// candidate boxes, kMiss when it surely misses all of them, else kUndecided
// (left to the exact re-test).  Boxes separated by the widened broad phase
// are misses; the others go through classify32.
__device__ __forceinline__ int seg_state32(const float *v, const float *w, const EdgeRec32 &R,
                                           const unsigned long long *emask_row, int n_words,
                                           const BoxRec *boxes, const float *__restrict__ obstacles,
                                           int n_masked, int n_obs)
{
    const SegBox sb = seg_box32(v, w, R.m);
    int state = kMiss;
    for (int wd = 0; wd < n_words && state != kHit; ++wd) {
        unsigned long long mk = emask_row[wd];
        while (mk && state != kHit) {
            const BoxRec B = boxes[wd * 64 + __builtin_ctzll(mk)];
            mk &= mk - 1;
            if (!surely_separated(sb, B.lo, B.hi)) {
                const int c = classify32(v, w, R.m, B.lo, B.hi);
                state = c == kMiss ? state : c;
            }
        }
    }
    for (int b = n_masked; b < n_obs && state != kHit; ++b) {
        const float *ob = obstacles + (int64_t)b * 12;
        const float lo[3] = {ob[0], ob[1], ob[2]};
        const float hi[3] = {ob[6], ob[7], ob[8]};
        if (!surely_separated(sb, lo, hi)) {
            const int c = classify32(v, w, R.m, lo, hi);
            state = c == kMiss ? state : c;
        }
    }
    return state;
}

// The reference's own test on segment (k, k+1) of block-local edge le: its
// FP64 waypoints k and k+1 (pathPoint, 2pBVP.cu:117-131, from the edge's
// EdgePoly) against the edge's candidate boxes (collisionCheck.cu:65-143).
// A hit clears the edge's flag (only zeros are written: concurrent writes
// agree).  Out of line: it runs a few times per pass, and its FP64
// registers, and any spills, stay off the rest of the kernel.
__device__ __noinline__ void exact32(const EdgeRec32 *recs, const int *loff,
                                     const unsigned long long *emask_row, int n_words,
                                     const BoxRec *boxes, const float *__restrict__ obstacles,
                                     int n_masked, int n_obs, uint8_t *ok, int le, int k)
{
    const EdgeRec32 &R = recs[le];
    const int Wq = loff[le + 1] - loff[le];
    const float x0[kDim] = {R.c[0][0], R.c[1][0], R.c[2][0], R.c[0][1], R.c[1][1], R.c[2][1]};
    const TauPowers P = tau_powers(R.tau);
    const EdgePoly E = edge_poly(x0, R.x1, P);
    float ve[kDim], we[kDim], bmin[3], bmax[3];
    waypoint_dt(k, Wq, P, E, R.dt, x0, R.x1, ve);
    waypoint_dt(k + 1, Wq, P, E, R.dt, x0, R.x1, we);
    seg_bbox(ve, we, bmin, bmax);
    bool hit = false;
    for (int wd = 0; wd < n_words && !hit; ++wd) {
        unsigned long long m = emask_row[wd];
        while (m && !hit) {
            const BoxRec B = boxes[wd * 64 + __builtin_ctzll(m)];
            m &= m - 1;
            hit = segment_hits_box(ve, we, bmin, bmax, B.lo, B.hi);
        }
    }
    for (int b = n_masked; b < n_obs && !hit; ++b) {
        const float *ob = obstacles + (int64_t)b * 12;
        const float lo[3] = {ob[0], ob[1], ob[2]};
        const float hi[3] = {ob[6], ob[7], ob[8]};
        hit = segment_hits_box(ve, we, bmin, bmax, lo, hi);
    }
    if (hit)
        ok[le] = 0;
}

// A listed segment, decided after the tiles: from its FP32 waypoints by
// seg_state32 (a sure hit clears the edge's flag), by exact32 when
// undecided.  Returns 1 when the exact re-test ran.
__device__ __forceinline__ int decide32(const EdgeRec32 *recs, const int *loff,
                                        const unsigned long long *emask, int n_words,
                                        const BoxRec *boxes, const float *__restrict__ obstacles,
                                        int n_masked, int n_obs, uint8_t *ok, int le, int k)
{
    const EdgeRec32 &R = recs[le];
    const int Wq = loff[le + 1] - loff[le];
    const unsigned long long *emask_row = emask + le * n_words;
    float v[kDim], w[kDim];
    waypoint32(R, k, Wq, v);
    waypoint32(R, k + 1, Wq, w);
    const int st = seg_state32(v, w, R, emask_row, n_words, boxes, obstacles, n_masked, n_obs);
    if (st == kHit)
        ok[le] = 0;
    if (st != kUndecided)
        return 0;
    exact32(recs, loff, emask_row, n_words, boxes, obstacles, n_masked, n_obs, ok, le, k);
    return 1;
}

// A workgroup owns a run of EPB consecutive edges.  Set-up, per edge: the
// pathPoint constants, and a candidate mask over the boxes -- a box
// separated from the edge's waypoint bounding box (with the reference's <=
// test, collisionCheck.cu:104-112) is rejected by the broad phase for every
// segment of the edge, so leaving it out changes no result.  Each wave then
// walks its own tiles of 64 waypoints (consecutive tiles overlap by one
// waypoint, so a tile tests 63 segments): a lane finds its edge in a byte
// map of the block's waypoints, evaluates and writes its waypoint, takes the
// far end of its segment from the next lane (DPP), and tests the segment
// against its edge's candidate boxes only.  An edge already known to be
// blocked skips its remaining segments, and a lane stops at its first hit.
template <bool kCheck, bool kWrite, int EPB, bool kF32 = false>
__device__ __forceinline__ void edge_block(const EdgeJob &J, int64_t e0, char *smem, int obs_cap,
                                           int words_cap,
                                           unsigned long long *__restrict__ retests = nullptr)
{
    // the FP32-waypoint form keeps its own per-edge record
    using Rec = typename std::conditional<kF32, EdgeRec32, EdgeRec>::type;
    static_assert(EPB <= 256, "edge index of the waypoint map is a byte");
    const float *__restrict__ samples = J.samples;
    const int *__restrict__ edge_src = J.edge_src;
    const int *__restrict__ edge_dst = J.edge_dst;
    const float *__restrict__ topts_edge = J.topts_edge;
    const int64_t *__restrict__ wp_offsets = J.wp_offsets;
    const float *__restrict__ obstacles = J.obstacles;
    const int n_obs = J.n_obs;
    uint8_t *__restrict__ is_free = J.is_free;
    constexpr int kWaves = kTile / 64;
    const EdgeLds L = edge_lds(EPB, obs_cap, words_cap, (int)sizeof(Rec), kF32 ? kRetestCap : 0);
    Rec *recs = reinterpret_cast<Rec *>(smem + L.recs);           // per-edge records
    BoxRec *boxes = reinterpret_cast<BoxRec *>(smem + L.obs);     // staged boxes
    unsigned long long *emask = reinterpret_cast<unsigned long long *>(smem + L.emask);
    int *loff = reinterpret_cast<int *>(smem + L.loff);           // first waypoint (block-local)
    uint8_t *ok = reinterpret_cast<uint8_t *>(smem + L.ok);       // flag
    uint8_t *wmap = reinterpret_cast<uint8_t *>(smem + L.wmap);   // edge of each local waypoint
    uint32_t *rlist = reinterpret_cast<uint32_t *>(smem + L.rlist); // FP32 form: (edge, k) to re-test
    int *rcnt = reinterpret_cast<int *>(smem + L.rcnt);

    const int ne = (int)((J.n_edges - e0) < EPB ? (J.n_edges - e0) : EPB);
    const int t = threadIdx.x, lane = t & 63;
    const int wid = __builtin_amdgcn_readfirstlane(t >> 6);
    const int n_masked = n_obs < obs_cap ? n_obs : obs_cap; // staged boxes with candidate masks
    const int n_words = edge_words(n_masked);
    const int64_t g_begin = wp_offsets ? wp_offsets[e0] : e0 * (int64_t)J.w_fixed;
    float *__restrict__ disc = J.disc ? J.disc + g_begin * kDim : nullptr;
    // kTile / EPB lanes per edge: lane t serves edge q = t % EPB as part
    // t / EPB.  Part 0 forms the edge's record (constants, end states), its
    // flag and its entries of the waypoint map; the mask part (part 1, or
    // the only one) its candidate boxes.  Both work from their own global
    // reads, so the whole set-up needs a single barrier.
    constexpr int kParts = kTile / EPB;
    constexpr int kMaskPart = kParts == 1 ? 0 : 1;
    const int q = t % EPB;
    const int part = __builtin_amdgcn_readfirstlane(t / EPB); // uniform within a wave
    const bool own = q < ne;
    const int64_t W_fixed = J.w_fixed;
    const int n_loc =
        (int)((wp_offsets ? wp_offsets[e0 + ne] : (e0 + ne) * W_fixed) - g_begin);
    const bool mapped = n_loc <= kMapCap;
    if (kCheck)
        for (int i = t; i < n_masked * 3; i += kTile) {
            const int b = i / 3, d = i - 3 * (i / 3);
            boxes[b].lo[d] = obstacles[b * 12 + d];
            boxes[b].hi[d] = obstacles[b * 12 + 6 + d];
            if (d == 0)
                boxes[b].pad_lo = boxes[b].pad_hi = 0.0f;
        }
    for (int i = t; i <= ne; i += kTile)
        loff[i] = (int)((wp_offsets ? wp_offsets[e0 + i] : (e0 + i) * W_fixed) - g_begin);
    if (kF32 && t == 0)
        *rcnt = 0;
    const bool works = part == 0 || (kCheck && part == kMaskPart);
    float tau = 1.0f;
    float x0[kDim] = {0, 0, 0, 0, 0, 0}, x1[kDim] = {0, 0, 0, 0, 0, 0};
    int a = 0, W = 0;
    if (own && works) {
        const int64_t e = e0 + q;
        tau = topts_edge[e];
        const int src = edge_src[e], dst = edge_dst[e];
        const int64_t ga = wp_offsets ? wp_offsets[e] : e * W_fixed;
        const int64_t gb = wp_offsets ? wp_offsets[e + 1] : (e + 1) * W_fixed;
        a = (int)(ga - g_begin);
        W = (int)(gb - ga);
        for (int d = 0; d < kDim; ++d) {
            x0[d] = samples[(int64_t)src * kDim + d];
            x1[d] = samples[(int64_t)dst * kDim + d];
        }
    }
    if (own && part == 0) {
        if constexpr (kF32) {
            edge_rec32(recs[q], x0, x1, tau, W);
        } else {
            const TauPowers P = tau_powers(tau);
            EdgeRec &R = recs[q];
            R.E = edge_poly(x0, x1, P);
            R.t3 = P.t3;
            R.r2 = P.r2;
            R.r3 = P.r3;
            R.tau = tau;
            R.dt = tau / (float)(W - 1);
            for (int d = 0; d < kDim; ++d) {
                R.x0[d] = x0[d];
                R.x1[d] = x1[d];
            }
        }
        ok[q] = 1;
        if (mapped)
            for (int j = 0; j < W; ++j)
                wmap[a + j] = (uint8_t)q;
    }
    if (kCheck && n_words > 0 && part == kMaskPart)
        candidate_masks(own, x0, x1, tau, obstacles, n_masked, n_words, emask + q * n_words);
    __syncthreads();

    // tiles start every 63 waypoints; tile j belongs to wave j % kWaves
    for (int base = wid * 63; base < n_loc; base += kWaves * 63) {
        const int g = base + lane;
        const bool have = g < n_loc;
        int le = 0, k = 0, Wq = 0;
        float x[kDim] = {0, 0, 0, 0, 0, 0};
        if (have) {
            le = mapped ? (int)wmap[g] : find_edge_local(loff, ne, g);
            const int a0 = loff[le];
            k = g - a0;
            Wq = loff[le + 1] - a0;
            if constexpr (kF32) {
                waypoint32(recs[le], k, Wq, x);
            } else {
                const EdgeRec &R = recs[le];
                const TauPowers P{R.tau, R.t3, R.r2, R.r3};
                waypoint_dt(k, Wq, P, R.E, R.dt, R.x0, R.x1, x);
            }
            // lane 63's waypoint is lane 0's of the next tile
            if (kWrite && (lane < 63 || g + 1 >= n_loc))
                for (int d = 0; d < kDim; ++d)
                    disc[(int64_t)g * kDim + d] = x[d];
        }
        if (!kCheck)
            continue;
        // segment (k, k+1): far end from the next lane
        float w[3];
        for (int d = 0; d < 3; ++d)
            w[d] = dpp_move<0x130>(x[d]); // from lane + 1
        const bool test = have && lane < 63 && k < Wq - 1 && ok[le];
        if (kF32 && test) {
            // FP32 waypoints: a segment that some candidate box does not
            // surely clear is listed, and decided after the tiles by the
            // whole workgroup at once
            const float v[3] = {x[0], x[1], x[2]};
            if (seg_near32(v, w, reinterpret_cast<const EdgeRec32 *>(recs)[le],
                           emask + le * n_words, n_words, boxes, obstacles, n_masked, n_obs)) {
                const int i = atomicAdd(rcnt, 1);
                if (i < kRetestCap)
                    rlist[i] = ((uint32_t)le << 24) | (uint32_t)k;
            }
        }
        if (!kF32 && test) {
            const float v[3] = {x[0], x[1], x[2]};
            float bmin[3], bmax[3];
            seg_bbox(v, w, bmin, bmax);
            bool hit = false;
            for (int wd = 0; wd < n_words && !hit; ++wd) {
                unsigned long long m = emask[le * n_words + wd];
                while (m && !hit) {
                    const int b = wd * 64 + __builtin_ctzll(m);
                    m &= m - 1;
                    const BoxRec B = boxes[b];
                    hit = segment_hits_box(v, w, bmin, bmax, B.lo, B.hi);
                }
            }
            // boxes past the staged ones: a uniform loop, so that their
            // bounds come through scalar loads and no vector load is left
            // pending across the waypoint stores
            for (int b = n_masked; b < n_obs; ++b) {
                const float *ob = obstacles + __builtin_amdgcn_readfirstlane(b) * 12;
                const float lo[3] = {ob[0], ob[1], ob[2]};
                const float hi[3] = {ob[6], ob[7], ob[8]};
                hit = hit || segment_hits_box(v, w, bmin, bmax, lo, hi);
            }
            if (hit)
                ok[le] = 0; // only zeros are written: concurrent writes agree
        }
    }
    if (kF32) {
        // the listed segments, decided by the whole workgroup (decide32)
        __syncthreads();
        const int n_set = *rcnt;
        int n_exact = 0;
        if (n_set <= kRetestCap) {
            for (int i = t; i < n_set; i += kTile) {
                const uint32_t r = rlist[i];
                const int le = (int)(r >> 24), k = (int)(r & 0xffffffu);
                if (ok[le])
                    n_exact += decide32(reinterpret_cast<const EdgeRec32 *>(recs), loff, emask,
                                        n_words, boxes, obstacles, n_masked, n_obs, ok, le,
                                        k);
            }
        } else {
            // more than the list holds: every segment of the block's edges
            // is decided again, and the undecided ones re-tested
            for (int g = t; g < n_loc; g += kTile) {
                const int le = mapped ? (int)wmap[g] : find_edge_local(loff, ne, g);
                const int k = g - loff[le];
                const int Wq = loff[le + 1] - loff[le];
                if (k >= Wq - 1 || !ok[le])
                    continue;
                const EdgeRec32 &R = reinterpret_cast<const EdgeRec32 *>(recs)[le];
                float v[kDim], w[kDim];
                waypoint32(R, k, Wq, v);
                waypoint32(R, k + 1, Wq, w);
                if (seg_near32(v, w, R, emask + le * n_words, n_words, boxes, obstacles, n_masked,
                               n_obs))
                    n_exact += decide32(reinterpret_cast<const EdgeRec32 *>(recs), loff, emask,
                                        n_words, boxes, obstacles, n_masked, n_obs, ok, le, k);
            }
        }
        // exact re-tests (rare: a few per pass)
        if (retests && n_exact > 0)
            atomicAdd(retests, (unsigned long long)n_exact);
    }
    if (kCheck) {
        __syncthreads();
        for (int i = t; i < ne; i += kTile)
            is_free[e0 + i] = ok[i];
    }
}

template <bool kCheck, bool kWrite, int EPB>
__global__ __launch_bounds__(kTile) void k_edges(
    const float *__restrict__ samples, const int *__restrict__ edge_src,
    const int *__restrict__ edge_dst, const float *__restrict__ topts_edge, int64_t n_edges,
    const int64_t *__restrict__ wp_offsets, int w_fixed, const float *__restrict__ obstacles,
    int n_obs, float *__restrict__ disc, uint8_t *__restrict__ is_free)
{
    extern __shared__ __align__(16) char smem[];
    const EdgeJob J{samples, edge_src, edge_dst, topts_edge, n_edges, wp_offsets,
                    w_fixed, obstacles, n_obs, disc, is_free};
    const int cap = kCheck ? edge_obs_cap(n_obs) : 0;
    edge_block<kCheck, kWrite, EPB>(J, (int64_t)blockIdx.x * EPB, smem, cap, edge_words(cap));
}

// the fused edge pass with FP32 waypoints and exact flags (chosen form)
template <bool kWrite, int EPB>
__global__ __launch_bounds__(kTile, 8) void k_edges32(
    const float *__restrict__ samples, const int *__restrict__ edge_src,
    const int *__restrict__ edge_dst, const float *__restrict__ topts_edge, int64_t n_edges,
    const int64_t *__restrict__ wp_offsets, int w_fixed, const float *__restrict__ obstacles,
    int n_obs, float *__restrict__ disc, uint8_t *__restrict__ is_free,
    unsigned long long *__restrict__ retests)
{
    extern __shared__ __align__(16) char smem[];
    const EdgeJob J{samples, edge_src, edge_dst, topts_edge, n_edges, wp_offsets,
                    w_fixed, obstacles, n_obs, disc, is_free};
    const int cap = edge_obs_cap(n_obs);
    edge_block<true, kWrite, EPB, true>(J, (int64_t)blockIdx.x * EPB, smem, cap, edge_words(cap),
                                        retests);
}

// the fused edge pass of many instances in one launch; LDS holds obs_cap
// boxes (the obstacle tile: at most the largest count of the instances and
// kEdgeObs), the boxes of an instance past the tile are tested from global
// memory
template <int EPB>
__global__ __launch_bounds__(kTile) void k_edges_many(const EdgeJob *__restrict__ jobs,
                                                      const int64_t *__restrict__ prefix,
                                                      int n_inst, int obs_cap)
{
    extern __shared__ __align__(16) char smem[];
    const int k = instance_of(prefix, n_inst, blockIdx.x);
    const EdgeJob J = global_job(jobs[k]);
    edge_block<true, true, EPB>(J, ((int64_t)blockIdx.x - prefix[k]) * EPB, smem, obs_cap,
                                edge_words(obs_cap));
}

// freeEdges over precomputed discMotions: one lane per segment; the edge's
// flag lives in LDS for the workgroup that owns the edge.
__global__ __launch_bounds__(kTile) void k_free_edges(const float *__restrict__ obstacles,
                                                      int n_obs, int w_fixed,
                                                      const int64_t *__restrict__ wp_offsets,
                                                      const float *__restrict__ disc,
                                                      uint8_t *__restrict__ is_free,
                                                      int64_t n_edges)
{
    __shared__ ObsLds o;
    __shared__ int64_t off[kEdgesPerBlock + 1];
    __shared__ int ok[kEdgesPerBlock];
    const int64_t e0 = (int64_t)blockIdx.x * kEdgesPerBlock;
    const int ne = (int)((n_edges - e0) < kEdgesPerBlock ? (n_edges - e0) : kEdgesPerBlock);
    const int t = threadIdx.x;
    stage_obstacles(o, obstacles, n_obs);
    for (int q = t; q <= ne; q += kTile)
        off[q] = wp_offsets ? wp_offsets[e0 + q] : (e0 + q) * (int64_t)w_fixed;
    for (int q = t; q < ne; q += kTile)
        ok[q] = 1;
    __syncthreads();
    const int64_t g_begin = off[0], g_end = off[ne];
    for (int64_t g = g_begin + t; g < g_end; g += kTile) {
        int le = find_edge(off, ne, g);
        int k = (int)(g - off[le]);
        int W = (int)(off[le + 1] - off[le]);
        if (k < W - 1 && ok[le]) {
            float v[3], w[3];
            for (int d = 0; d < 3; ++d) {
                v[d] = disc[g * kDim + d];
                w[d] = disc[(g + 1) * kDim + d];
            }
            if (segment_blocked(v, w, o, obstacles, n_obs))
                ok[le] = 0;
        }
    }
    __syncthreads();
    for (int q = t; q < ne; q += kTile)
        is_free[e0 + q] = (uint8_t)ok[q];
}

__global__ __launch_bounds__(256) void k_sample_free(const float *__restrict__ obstacles, int n_obs,
                                                     const float *__restrict__ samples, int num,
                                                     uint8_t *__restrict__ is_free)
{
    __shared__ ObsLds o;
    stage_obstacles(o, obstacles, n_obs);
    __syncthreads();
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= num)
        return;
    float p[3] = {samples[(int64_t)n * kDim], samples[(int64_t)n * kDim + 1],
                  samples[(int64_t)n * kDim + 2]};
    bool free_ = true;
    for (int b = 0; b < n_obs && free_; ++b) {
        float lo[3], hi[3];
        box_bounds(o, obstacles, b, lo, hi);
        if (p[0] > lo[0] && p[0] < hi[0] && p[1] > lo[1] && p[1] < hi[1] && p[2] > lo[2] &&
            p[2] < hi[2])
            free_ = false;
    }
    is_free[n] = free_ ? 1 : 0;
}

} // namespace
} // namespace pump

using namespace pump;

static int edge_blocks(int64_t n_edges, unsigned &blocks)
{
    int64_t b = (n_edges + kEdgesPerBlock - 1) / kEdgesPerBlock;
    if (b > 0x7fffffff)
        return PUMP_E_RANGE;
    blocks = (unsigned)(b > 0 ? b : 1);
    return PUMP_OK;
}

// dynamic LDS above 64 KiB has to be allowed per kernel
static int allow_lds(const void *fn, size_t bytes)
{
    if (bytes > 65536)
        PUMP_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return PUMP_OK;
}

// Edges per workgroup of the single-instance edge pass: 64, 128 (default) or
// 256, from PUMP_EDGES_PER_BLOCK.
static int edges_per_block()
{
    static int v = 0;
    if (v == 0) {
        const char *e = std::getenv("PUMP_EDGES_PER_BLOCK");
        const int x = e ? std::atoi(e) : 128;
        v = (x == 64 || x == 256) ? x : 128;
    }
    return v;
}

template <bool kCheck, bool kWrite, int EPB>
static int launch_edges(const float *d_samples, const int *d_edge_src, const int *d_edge_dst,
                        const float *d_topts_edge, int64_t n_edges, const int64_t *d_wp_offsets,
                        int waypoints_count, const float *d_obstacles, int n_obs, float *d_disc,
                        uint8_t *d_is_free, hipStream_t s)
{
    const int64_t nb = (n_edges + EPB - 1) / EPB;
    if (nb > 0x7fffffff)
        return PUMP_E_RANGE;
    const int cap = kCheck ? edge_obs_cap(n_obs) : 0;
    const size_t lds = (size_t)edge_lds(EPB, cap, edge_words(cap)).bytes;
    const void *fn = reinterpret_cast<const void *>(&k_edges<kCheck, kWrite, EPB>);
    const int r = allow_lds(fn, lds);
    if (r != PUMP_OK)
        return r;
    hipLaunchKernelGGL((k_edges<kCheck, kWrite, EPB>), dim3((unsigned)nb), dim3(kTile), lds, s,
                       d_samples, d_edge_src, d_edge_dst, d_topts_edge, n_edges, d_wp_offsets,
                       waypoints_count, d_obstacles, n_obs, d_disc, d_is_free);
    PUMP_LAUNCHED();
    return PUMP_OK;
}

template <bool kCheck, bool kWrite>
static int launch_edges_any(const float *d_samples, const int *d_edge_src, const int *d_edge_dst,
                            const float *d_topts_edge, int64_t n_edges,
                            const int64_t *d_wp_offsets, int waypoints_count,
                            const float *d_obstacles, int n_obs, float *d_disc,
                            uint8_t *d_is_free, hipStream_t s)
{
    switch (edges_per_block()) {
    case 64:
        return launch_edges<kCheck, kWrite, 64>(d_samples, d_edge_src, d_edge_dst, d_topts_edge,
                                                n_edges, d_wp_offsets, waypoints_count,
                                                d_obstacles, n_obs, d_disc, d_is_free, s);
    case 256:
        return launch_edges<kCheck, kWrite, 256>(d_samples, d_edge_src, d_edge_dst, d_topts_edge,
                                                 n_edges, d_wp_offsets, waypoints_count,
                                                 d_obstacles, n_obs, d_disc, d_is_free, s);
    default:
        return launch_edges<kCheck, kWrite, 128>(d_samples, d_edge_src, d_edge_dst, d_topts_edge,
                                                 n_edges, d_wp_offsets, waypoints_count,
                                                 d_obstacles, n_obs, d_disc, d_is_free, s);
    }
}

extern "C" int pump_discretize_edges(const float *d_samples, const int *d_edge_src,
                                     const int *d_edge_dst, const float *d_topts_edge,
                                     int64_t n_edges, const int64_t *d_wp_offsets,
                                     int waypoints_count, float *d_disc_motions,
                                     pump_stream_t stream)
{
    PUMP_REQUIRE(n_edges >= 0);
    if (n_edges == 0)
        return PUMP_OK;
    PUMP_REQUIRE(d_samples && d_edge_src && d_edge_dst && d_topts_edge && d_disc_motions);
    PUMP_REQUIRE(d_wp_offsets || waypoints_count >= 2);
    PUMP_REQUIRE(waypoints_count < kMaxWaypointsPerEdge);
    return launch_edges_any<false, true>(d_samples, d_edge_src, d_edge_dst, d_topts_edge, n_edges,
                                         d_wp_offsets, waypoints_count, nullptr, 0,
                                         d_disc_motions, nullptr, as_stream(stream));
}

extern "C" int pump_edges_discretize_check(const float *d_samples, const int *d_edge_src,
                                           const int *d_edge_dst, const float *d_topts_edge,
                                           int64_t n_edges, const int64_t *d_wp_offsets,
                                           int waypoints_count, const float *d_obstacles,
                                           int n_obs, float *d_disc_motions,
                                           uint8_t *d_is_free_edges, pump_stream_t stream)
{
    PUMP_REQUIRE(n_edges >= 0 && n_obs >= 0);
    if (n_edges == 0)
        return PUMP_OK;
    PUMP_REQUIRE(d_samples && d_edge_src && d_edge_dst && d_topts_edge && d_is_free_edges);
    PUMP_REQUIRE(n_obs == 0 || d_obstacles);
    PUMP_REQUIRE(d_wp_offsets || waypoints_count >= 2);
    PUMP_REQUIRE(waypoints_count < kMaxWaypointsPerEdge);
    if (d_disc_motions)
        return launch_edges_any<true, true>(d_samples, d_edge_src, d_edge_dst, d_topts_edge,
                                            n_edges, d_wp_offsets, waypoints_count, d_obstacles,
                                            n_obs, d_disc_motions, d_is_free_edges,
                                            as_stream(stream));
    return launch_edges_any<true, false>(d_samples, d_edge_src, d_edge_dst, d_topts_edge, n_edges,
                                         d_wp_offsets, waypoints_count, d_obstacles, n_obs,
                                         nullptr, d_is_free_edges, as_stream(stream));
}

extern "C" int pump_edges_discretize_check_fp32(
    const float *d_samples, const int *d_edge_src, const int *d_edge_dst,
    const float *d_topts_edge, int64_t n_edges, const int64_t *d_wp_offsets, int waypoints_count,
    const float *d_obstacles, int n_obs, float *d_disc_motions, uint8_t *d_is_free_edges,
    unsigned long long *d_retests, pump_stream_t stream)
{
    PUMP_REQUIRE(n_edges >= 0 && n_obs >= 0);
    if (n_edges == 0)
        return PUMP_OK;
    PUMP_REQUIRE(d_samples && d_edge_src && d_edge_dst && d_topts_edge && d_is_free_edges);
    PUMP_REQUIRE(n_obs == 0 || d_obstacles);
    PUMP_REQUIRE(d_wp_offsets || waypoints_count >= 2);
    PUMP_REQUIRE(waypoints_count < kMaxWaypointsPerEdge);
    constexpr int EPB = 128;
    const int64_t nb = (n_edges + EPB - 1) / EPB;
    if (nb > 0x7fffffff)
        return PUMP_E_RANGE;
    const int cap = edge_obs_cap(n_obs);
    const size_t lds =
        (size_t)edge_lds(EPB, cap, edge_words(cap), (int)sizeof(EdgeRec32), kRetestCap).bytes;
    hipStream_t s = as_stream(stream);
    if (d_disc_motions) {
        const void *fn = reinterpret_cast<const void *>(&k_edges32<true, EPB>);
        int r = allow_lds(fn, lds);
        if (r != PUMP_OK)
            return r;
        hipLaunchKernelGGL((k_edges32<true, EPB>), dim3((unsigned)nb), dim3(kTile), lds, s,
                           d_samples, d_edge_src, d_edge_dst, d_topts_edge, n_edges, d_wp_offsets,
                           waypoints_count, d_obstacles, n_obs, d_disc_motions, d_is_free_edges,
                           d_retests);
    } else {
        const void *fn = reinterpret_cast<const void *>(&k_edges32<false, EPB>);
        int r = allow_lds(fn, lds);
        if (r != PUMP_OK)
            return r;
        hipLaunchKernelGGL((k_edges32<false, EPB>), dim3((unsigned)nb), dim3(kTile), lds, s,
                           d_samples, d_edge_src, d_edge_dst, d_topts_edge, n_edges, d_wp_offsets,
                           waypoints_count, d_obstacles, n_obs, (float *)nullptr,
                           d_is_free_edges, d_retests);
    }
    PUMP_LAUNCHED();
    return PUMP_OK;
}

namespace {

int free_edges_launch(const float *d_obstacles, int n_obs, int waypoints_count,
                      const int64_t *d_wp_offsets, const float *d_disc_motions,
                      uint8_t *d_is_free_edges, int64_t n_edges, pump_stream_t stream)
{
    unsigned blocks;
    int r = edge_blocks(n_edges, blocks);
    if (r)
        return r;
    hipLaunchKernelGGL(k_free_edges, dim3(blocks), dim3(kTile), 0, as_stream(stream), d_obstacles,
                       n_obs, waypoints_count, d_wp_offsets, d_disc_motions, d_is_free_edges,
                       n_edges);
    PUMP_LAUNCHED();
    return PUMP_OK;
}

} // namespace

extern "C" int pump_free_edges(const float *d_obstacles, int n_obs, const float *d_samples,
                               const uint8_t *d_is_free_samples, int num_disc,
                               const float *d_disc_motions, uint8_t *d_is_free_edges,
                               int64_t n_edges, pump_stream_t stream)
{
    (void)d_samples;          // passed but unused, as in the reference
    (void)d_is_free_samples;  // idem (collisionCheck.cu:5-37)
    PUMP_REQUIRE(n_edges >= 0 && n_obs >= 0 && num_disc >= 0);
    PUMP_REQUIRE(num_disc < kMaxWaypointsPerEdge);
    if (n_edges == 0)
        return PUMP_OK;
    PUMP_REQUIRE(d_disc_motions && d_is_free_edges && (n_obs == 0 || d_obstacles));
    // numDisc segments per edge, its waypoints (numDisc + 1 rows) contiguous
    return free_edges_launch(d_obstacles, n_obs, num_disc + 1, nullptr, d_disc_motions,
                             d_is_free_edges, n_edges, stream);
}

extern "C" int pump_free_edges_var(const float *d_obstacles, int n_obs,
                                   const int64_t *d_wp_offsets, const float *d_disc_motions,
                                   uint8_t *d_is_free_edges, int64_t n_edges, pump_stream_t stream)
{
    PUMP_REQUIRE(n_edges >= 0 && n_obs >= 0);
    if (n_edges == 0)
        return PUMP_OK;
    PUMP_REQUIRE(d_wp_offsets && d_disc_motions && d_is_free_edges && (n_obs == 0 || d_obstacles));
    return free_edges_launch(d_obstacles, n_obs, 0, d_wp_offsets, d_disc_motions, d_is_free_edges,
                             n_edges, stream);
}

extern "C" int pump_sample_free(const float *d_obstacles, int n_obs, const float *d_samples,
                                int num, uint8_t *d_is_free_samples, pump_stream_t stream)
{
    PUMP_REQUIRE(num >= 0 && n_obs >= 0);
    if (num == 0)
        return PUMP_OK;
    PUMP_REQUIRE(d_samples && d_is_free_samples && (n_obs == 0 || d_obstacles));
    hipLaunchKernelGGL(k_sample_free, dim3((num + 255) / 256), dim3(256), 0, as_stream(stream),
                       d_obstacles, n_obs, d_samples, num, d_is_free_samples);
    PUMP_LAUNCHED();
    return PUMP_OK;
}

int pump::edges_many(const EdgeJob *d_jobs, const int64_t *d_block_prefix, int n_inst,
                     int64_t n_blocks, int edges_per_block, int max_obs, int obs_tile,
                     hipStream_t s)
{
    PUMP_REQUIRE(d_jobs && d_block_prefix && n_inst > 0 && n_blocks >= 0 &&
                 n_blocks <= 0x7fffffff && max_obs >= 0 && obs_tile > 0);
    if (n_blocks == 0)
        return PUMP_OK;
    const int cap = std::min(edge_obs_cap(max_obs), obs_tile);
    const size_t lds = (size_t)edge_lds(edges_per_block, cap, edge_words(cap)).bytes;
    const void *fn;
    if (edges_per_block == 64)
        fn = reinterpret_cast<const void *>(&k_edges_many<64>);
    else if (edges_per_block == 128)
        fn = reinterpret_cast<const void *>(&k_edges_many<128>);
    else if (edges_per_block == 256)
        fn = reinterpret_cast<const void *>(&k_edges_many<256>);
    else
        return PUMP_E_INVALID;
    const int r = allow_lds(fn, lds);
    if (r != PUMP_OK)
        return r;
    if (edges_per_block == 64)
        hipLaunchKernelGGL(k_edges_many<64>, dim3((unsigned)n_blocks), dim3(kTile), lds, s, d_jobs,
                           d_block_prefix, n_inst, cap);
    else if (edges_per_block == 128)
        hipLaunchKernelGGL(k_edges_many<128>, dim3((unsigned)n_blocks), dim3(kTile), lds, s,
                           d_jobs, d_block_prefix, n_inst, cap);
    else
        hipLaunchKernelGGL(k_edges_many<256>, dim3((unsigned)n_blocks), dim3(kTile), lds, s,
                           d_jobs, d_block_prefix, n_inst, cap);
    PUMP_LAUNCHED();
    return PUMP_OK;
}
