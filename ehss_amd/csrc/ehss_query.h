// ehss_query.h -- nearest-hit and any-hit queries over a device scene.
//
// nearest_plain:  ehss_cpu.cpp:59-72 (FP64 box tests, reference traversal and counters).
// nearest_fast:   the same query through an FP32 filter with an exact FP64 re-test of
//                 every sphere the filter cannot exclude (identical result).
// nearest_packet: nearest_fast with a wave-uniform node sequence and scalar loads.
// nearest_packet_ordered: nearest_packet visiting children front to back.
// pa_plain:       pa_spatial.py:114-130 over the PA tree (FP64, reference counters).
// pa_fast:        any-hit through the FP32 filter over the static tree, FP64 re-test
//                 of undecided spheres (identical hit flag).
//
// Every query takes `tol`, the current grazing tolerance (a distance).  The
// filter never excludes a sphere whose exact FP64 outcome lies within tol of a
// decision threshold, so its grazing flag is conservative; the plain kernel
// keeps the reference traversal and flags boxes it culls within tol of best.
#pragma once
#include "ehss_internal.h"

namespace ehss {

struct Query {
    int32_t atom;
    double best;
    double second;  // second-smallest finite entry distance (nearest-atom tie test)
    double graze;   // smallest along-ray distance of a near-threshold sphere test
};

__device__ __forceinline__ void query_init(Query& q) {
    q.atom = -1;
    q.best = INFINITY;
    q.second = INFINITY;
    q.graze = INFINITY;
}

__device__ __forceinline__ void consider(Query& q, double t, int32_t wanted) {
    if (better(t, wanted, q.best, q.atom)) {
        q.second = q.best;
        q.best = t;
        q.atom = wanted;
    } else if (t < q.second) {
        q.second = t;
    }
}

// A query is grazing when a near-threshold sphere test lies at or before the
// winning distance, or when the runner-up is within tol of the winner.
__device__ __forceinline__ bool grazing(const Query& q, double tol) {
    return (q.graze < INFINITY && q.graze <= q.best + tol) || q.second - q.best <= tol;
}

template <class C = uint64_t>
__device__ __forceinline__ void nearest_plain(const SceneView& S, const V3& o, const V3& d, int32_t last, double tol,
                                              Query& q, C& boxes, C& spheres) {
    query_init(q);
    int32_t node = 0;
    const int32_t nn = S.n_nodes;
    while (node < nn) {
        const Node64& n = S.nodes[node];
        ++boxes;
        double near;
        if (!box64(o, d, n.lo, n.hi, q.best, &near)) {
            // A box culled by the reference rule near > best may still hold a
            // sphere within tol of best: count it as a potential near-tie.
            if (near <= q.best + tol) q.graze = fmin(q.graze, near);
            node = n.end;
            continue;
        }
        const int32_t g = n.leaf;
        if (g >= 0) {
            const int32_t k1 = S.offsets[g + 1];
            for (int32_t k = S.offsets[g]; k < k1; ++k) {
                const int32_t wanted = S.ids[k];
                if (wanted == last) continue;
                ++spheres;
                const double4 c = S.sph[wanted];
                const double t = entry64(o, d, c.x, c.y, c.z, c.w, tol, &q.graze);
                consider(q, t, wanted);
            }
        }
        ++node;
    }
}

// FP32 copy of a ray relative to the filter origin, with a safe reciprocal.
struct Ray32 {
    float ox, oy, oz, dx, dy, dz, ix, iy, iz, oix, oiy, oiz;
};

__device__ __forceinline__ Ray32 ray32(const SceneView& S, const V3& o, const V3& d) {
    Ray32 r;
    r.ox = (float)(o.x - S.cx);
    r.oy = (float)(o.y - S.cy);
    r.oz = (float)(o.z - S.cz);
    r.dx = (float)d.x;
    r.dy = (float)d.y;
    r.dz = (float)d.z;
    const float sx = fabsf(r.dx) < 1e-20f ? copysignf(1e-20f, r.dx) : r.dx;
    const float sy = fabsf(r.dy) < 1e-20f ? copysignf(1e-20f, r.dy) : r.dy;
    const float sz = fabsf(r.dz) < 1e-20f ? copysignf(1e-20f, r.dz) : r.dz;
    r.ix = 1.0f / sx;
    r.iy = 1.0f / sy;
    r.iz = 1.0f / sz;
    r.oix = r.ox * r.ix;
    r.oiy = r.oy * r.iy;
    r.oiz = r.oz * r.iz;
    return r;
}

// Inflated-box slab test: true unless the ray segment [0, bestf] provably misses.
// `near` receives the slab entry distance (used to choose a visiting order).
__device__ __forceinline__ bool box32_near(const Ray32& r, const Node32& n, float bestf, float& near) {
    const float tx0 = __builtin_fmaf(n.lo[0], r.ix, -r.oix), tx1 = __builtin_fmaf(n.hi[0], r.ix, -r.oix);
    const float ty0 = __builtin_fmaf(n.lo[1], r.iy, -r.oiy), ty1 = __builtin_fmaf(n.hi[1], r.iy, -r.oiy);
    const float tz0 = __builtin_fmaf(n.lo[2], r.iz, -r.oiz), tz1 = __builtin_fmaf(n.hi[2], r.iz, -r.oiz);
    near = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.f));
    const float far = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
    // near <= far && near <= bestf as one comparison (the same for all
    // non-NaN values; ray32 keeps every reciprocal finite, so no slab distance
    // is NaN, and neither is bestf).
    return near <= fminf(far, bestf);
}

__device__ __forceinline__ bool box32(const Ray32& r, const Node32& n, float bestf) {
    float near;
    return box32_near(r, n, bestf, near);
}

// FP32 upper bound of best + tol used for box pruning.
__device__ __forceinline__ float box_bound(const SceneView& S, double best, double tol) {
    return (float)(best + tol) * (1.0f + 4.0f * kU32) + S.m_box;
}

// FP64 any-hit sphere test (inside_out_pa.py:93-101) with a flag for decisions an
// input perturbation of size tol could flip: hit = (disc >= 0) and (exit > 0),
// critical when disc is within 2 r tol of 0 while the exit condition (-b > 0 at
// disc = 0) nearly holds, or when exit is within tol of 0 while disc nearly holds.
__device__ __forceinline__ bool sphere_hit64_graze(const V3& o, const V3& d, double cx, double cy, double cz, double r,
                                                   double tol, bool* critical) {
    double x = o.x - cx, y = o.y - cy, z = o.z - cz;
    double b = (x * d.x + y * d.y) + z * d.z;
    double cc = ((x * x + y * y) + z * z) - r * r;
    double disc = b * b - cc;
    double exit = -b + sqrt(smax(0., disc));
    const double tdisc = 2 * r * tol;
    *critical = (fabs(disc) <= tdisc && -b > -tol) || (fabs(exit) <= tol && disc >= -tdisc);
    return disc >= 0 && exit > 0;
}

// Any-hit PA decided during the first nearest query of an external ray
// (origin outside every sphere).  A sphere whose FP32 test is a robust hit is a
// robust PA hit; every sphere the query re-tests in FP64 also gets the FP64
// any-hit test.  When the query ends without a robust PA hit but with a winner,
// the caller runs the full any-hit query.
struct PAState {
    bool decided;   // a robust hit was found
    bool hit;       // some FP64 any-hit test returned true
    bool critical;  // some FP64 any-hit test was near a threshold
};

// Exact FP64 entry of one sphere (and its any-hit test when PA is pending).
__device__ __forceinline__ void eval64(const V3& o, const V3& d, const double4& c, int32_t wanted, double tol,
                                       Query& q, PAState* pa) {
    const double t = entry64(o, d, c.x, c.y, c.z, c.w, tol, &q.graze);
    consider(q, t, wanted);
    if (pa && !pa->decided) {
        bool crit;
        const bool h = sphere_hit64_graze(o, d, c.x, c.y, c.z, c.w, tol, &crit);
        pa->critical |= crit;
        if (h) {
            pa->hit = true;
            if (!crit) pa->decided = true;
        }
    }
}

// FP32 classification of one sphere against the running FP64 best.  Returns true
// when the exact FP64 entry must be evaluated: false only when the FP64 entry is
// provably infinite or larger than best + tol, and no FP64 threshold test of this
// sphere is within tol.
__device__ __forceinline__ bool needs_exact(const SceneView& S, const Ray32& r, const float4& s, double best,
                                            double tol, PAState* pa) {
    const float x0 = r.ox - s.x, x1 = r.oy - s.y, x2 = r.oz - s.z;
    const float p = x0 * r.dx + x1 * r.dy + x2 * r.dz;
    const float ftol = (float)tol * 1.01f;
    if (p >= S.m_p + ftol) return false;  // projection64 > tol: behind
    const float p0 = x0 - p * r.dx, p1 = x1 - p * r.dy, p2 = x2 - p * r.dz;
    const float disc = s.w * s.w - (p0 * p0 + p1 * p1 + p2 * p2);
    const float mdisc = S.m_disc + 2.01f * s.w * ftol;
    if (disc < -mdisc) return false;  // disc64 < -2 r tol: clear miss
    if (pa && disc > mdisc && p <= -(S.m_p + ftol)) pa->decided = pa->hit = true;  // robust hit in both tests
    // Upper bound of the square root from the hardware one (v_sqrt_f32, within
    // 1 ulp): the factor 1 + 2^-21, less the product's own rounding, still
    // exceeds 3 ulp (an ulp is at most 2^-23 of the value).  Two instructions
    // instead of the correctly rounded sequence.
    const float root = __builtin_amdgcn_sqrtf(fmaxf(disc, 0.f) + S.m_disc) * (1.0f + 0x1p-21f);
    const float tlo = (-p - S.m_p) - root;
    return !((double)tlo > best + tol);  // entry64 > best + tol: cannot win, tie or graze
}

template <class C = uint64_t>
__device__ __forceinline__ void nearest_fast(const SceneView& S, const V3& o, const V3& d, int32_t last, double tol,
                                             Query& q, C& boxes, C& spheres, C& retests,
                                             PAState* pa = nullptr) {
    query_init(q);
    const Ray32 r = ray32(S, o, d);
    float bestf = INFINITY;
    int32_t node = 0;
    const int32_t nn = S.n_nodes;
    while (node < nn) {
        const Node32 n = S.nodes32[node];
        ++boxes;
        if (!box32(r, n, bestf)) {
            node = n.end;
            continue;
        }
        const int32_t g = n.leaf;
        if (g >= 0) {
            const int32_t k1 = S.offsets[g + 1];
            for (int32_t k = S.offsets[g]; k < k1; ++k) {
                const int32_t wanted = S.ids[k];
                if (wanted == last) continue;
                ++spheres;
                if (!needs_exact(S, r, S.sph32[k], q.best, tol, pa)) continue;
                ++retests;
                const int32_t before = q.atom;
                eval64(o, d, S.sphs[k], wanted, tol, q, pa);
                if (q.atom != before) bestf = box_bound(S, q.best, tol);
            }
        }
        ++node;
    }
}

// Wave-uniform (packet) traversal: the 64 lanes walk one node sequence; a node is
// entered when any live lane's inflated box test passes.  Node, leaf and sphere
// records are read through the constant address space, so the compiler issues
// scalar loads (one fetch per wave) for them.  Per-lane state and the exact
// FP64 re-tests are those of nearest_fast, so every lane gets the same answer.
typedef const __attribute__((address_space(4))) float* CFloat;
typedef const __attribute__((address_space(4))) int32_t* CInt32;

__device__ __forceinline__ Node32 load_node(const Node32* base, int32_t node) {
    const CFloat f = (CFloat)(base + node);
    const CInt32 w = (CInt32)(base + node);
    Node32 n;
    n.lo[0] = f[0];
    n.lo[1] = f[1];
    n.lo[2] = f[2];
    n.hi[0] = f[3];
    n.hi[1] = f[4];
    n.hi[2] = f[5];
    n.end = w[6];
    n.leaf = w[7];
    return n;
}

__device__ __forceinline__ float4 load_sphere32(const float4* base, int32_t k) {
    const CFloat f = (CFloat)(base + k);
    return make_float4(f[0], f[1], f[2], f[3]);
}

typedef const __attribute__((address_space(4))) double* CDouble;

__device__ __forceinline__ double4 load_sphere64(const double4* base, int32_t k) {
    const CDouble f = (CDouble)(base + k);
    return make_double4(f[0], f[1], f[2], f[3]);
}

// kLds: the view's arrays that the packet queries read point into LDS (a
// block's staged copy); every lane then reads the same LDS address (a
// broadcast) instead of issuing a scalar load.
typedef const __attribute__((address_space(3))) float* LFloat;
typedef const __attribute__((address_space(3))) int32_t* LInt32;

// The links are made wave-uniform (readfirstlane) so that the traversal's
// control flow stays scalar, as with constant-address-space loads.
__device__ __forceinline__ Node32 load_node_lds(const Node32* base, int32_t node) {
    const LFloat f = (LFloat)(base + node);
    const LInt32 w = (LInt32)(base + node);
    Node32 n;
    n.lo[0] = f[0];
    n.lo[1] = f[1];
    n.lo[2] = f[2];
    n.hi[0] = f[3];
    n.hi[1] = f[4];
    n.hi[2] = f[5];
    n.end = __builtin_amdgcn_readfirstlane(w[6]);
    n.leaf = __builtin_amdgcn_readfirstlane(w[7]);
    return n;
}

__device__ __forceinline__ float4 load_sphere32_lds(const float4* base, int32_t k) {
    const LFloat f = (LFloat)(base + k);
    return make_float4(f[0], f[1], f[2], f[3]);
}

// Wave-uniform reads of the scene for the packet queries (the packet tree, its
// leaf offsets, slot ids, FP32 spheres): scalar loads, or LDS broadcasts when
// the block has staged the scene (kLds).
template <int kLds>
struct PacketMem {
    static __device__ __forceinline__ Node32 node(const SceneView& S, int32_t i) {
        if constexpr (kLds) return load_node_lds(S.pk_nodes32, i);
        else return load_node(S.pk_nodes32, i);
    }
    static __device__ __forceinline__ int32_t offset(const SceneView& S, int32_t g) {
        if constexpr (kLds) return __builtin_amdgcn_readfirstlane(((LInt32)S.pk_offsets)[g]);
        else return ((CInt32)S.pk_offsets)[g];
    }
    static __device__ __forceinline__ int32_t id(const SceneView& S, int32_t k) {
        if constexpr (kLds) return __builtin_amdgcn_readfirstlane(((LInt32)S.ids)[k]);
        else return ((CInt32)S.ids)[k];
    }
    static __device__ __forceinline__ float4 sphere(const SceneView& S, int32_t k) {
        if constexpr (kLds) return load_sphere32_lds(S.sph32, k);
        else return load_sphere32(S.sph32, k);
    }
};

template <int kLds = 0, class C = uint64_t>
__device__ __forceinline__ void nearest_packet(const SceneView& S, const V3& o, const V3& d, int32_t last,
                                               double tol, bool live, Query& q, C& boxes,
                                               C& spheres, C& retests, PAState* pa = nullptr) {
    query_init(q);
    const Ray32 r = ray32(S, o, d);
    typedef PacketMem<kLds> M;
    auto node_at = [&](int32_t i) { return M::node(S, i); };
    auto offset_at = [&](int32_t g) { return M::offset(S, g); };
    auto id_at = [&](int32_t k) { return M::id(S, k); };
    auto sphere_at = [&](int32_t k) { return M::sphere(S, k); };
    float bestf = INFINITY;
    int32_t node = 0;
    const int32_t nn = S.pk_nodes;
    while (node < nn) {
        const Node32 n = node_at(node);
        boxes += live ? 1 : 0;
        const bool hit = live && box32(r, n, bestf);
        if (__ballot(hit) == 0) {
            node = n.end;
            continue;
        }
        const int32_t g = n.leaf;
        if (g >= 0) {
            const int32_t k1 = offset_at(g + 1);
            for (int32_t k = offset_at(g); k < k1; ++k) {
                const int32_t wanted = id_at(k);
                const float4 s = sphere_at(k);
                if (!hit || wanted == last) continue;
                ++spheres;
                if (!needs_exact(S, r, s, q.best, tol, pa)) continue;
                ++retests;
                const int32_t before = q.atom;
                eval64(o, d, load_sphere64(S.sphs, k), wanted, tol, q, pa);
                if (q.atom != before) bestf = box_bound(S, q.best, tol);
            }
        }
        ++node;
    }
}

// Adds per_lane to the lane's work counter when `on`.  (Wave-uniform sums kept
// in scalar registers from ballots were measured slower: the scalar unit is as
// loaded as the vector unit in these loops.)
template <class C>
__device__ __forceinline__ void count_lanes(C& c, uint32_t per_lane, bool on) {
    c += on ? per_lane : 0;
}

// Front-to-back packet walk: the 64 lanes walk one node sequence of the packet
// tree.  At an internal node both children's boxes are tested; when lanes enter
// both, the packet descends first into the child that the first lane entering
// both reaches first, and keeps the other on a wave stack held across the lanes
// of one register (entry k in lane k).  The group tree halves its range at
// every level, so at most 31 entries are ever pushed.  A stacked node is
// re-tested against the lanes' bounds when it is popped.  Which lanes enter a
// node is kept as a wave-uniform lane mask (ballot of the box test, and of the
// parent's mask), so the walk's decisions stay in scalar registers.
// `leaf(g, mask)` is called (by every lane) for each leaf g some lane enters;
// `mask` holds the lanes that enter it.  `bestf` is each lane's current pruning
// bound; the leaf callback may lower it.
template <int kLds, class C, class LeafFn>
__device__ __forceinline__ void packet_walk(const SceneView& S, const Ray32& r, bool live, const float& bestf,
                                            C& boxes, LeafFn&& leaf) {
    typedef PacketMem<kLds> M;
    const uint64_t lanes = __ballot(live);
    int32_t stack = 0;
    int32_t sp = 0;
    int32_t node = 0;
    Node32 n = M::node(S, 0);
    count_lanes(boxes, 1, live);
    uint64_t hm = __ballot(box32(r, n, bestf)) & lanes;
    if (hm == 0) return;
    for (;;) {
        bool descend = false;
        if (n.leaf >= 0) {
            leaf(n.leaf, hm);
        } else {
            const int32_t a = node + 1, b = right_child(n.leaf);
            const Node32 na = M::node(S, a);
            const Node32 nb = M::node(S, b);
            count_lanes(boxes, 2, live);
            // Both tests on every lane, then masked: no branch, so both
            // children's records are fetched together.
            float neara, nearb;
            const uint64_t ma = __ballot(box32_near(r, na, bestf, neara)) & hm;
            const uint64_t mb = __ballot(box32_near(r, nb, bestf, nearb)) & hm;
            if (ma && mb) {
                // Order by the first lane that enters both children; when no lane
                // does, visit first the child more lanes enter.
                const uint64_t both = ma & mb;
                bool a_first;
                if (both) {
                    const int lb = __builtin_amdgcn_readfirstlane(__builtin_ctzll(both));
                    const float ra = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(neara), lb));
                    const float rb = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(nearb), lb));
                    a_first = !(rb < ra);
                } else {
                    a_first = __popcll(ma) >= __popcll(mb);
                }
                const int32_t far_child = a_first ? b : a;
                stack = (int32_t)(threadIdx.x & 63) == sp ? far_child : stack;
                ++sp;
                node = a_first ? a : b;
                n = a_first ? na : nb;
                hm = a_first ? ma : mb;
                descend = true;
            } else if (ma) {
                node = a;
                n = na;
                hm = ma;
                descend = true;
            } else if (mb) {
                node = b;
                n = nb;
                hm = mb;
                descend = true;
            }
        }
        if (descend) continue;
        // Pop until a stacked node is entered by some lane.
        bool found = false;
        while (sp > 0) {
            --sp;
            node = __builtin_amdgcn_readlane(stack, sp);
            n = M::node(S, node);
            count_lanes(boxes, 1, live);
            hm = __ballot(box32(r, n, bestf)) & lanes;
            if (hm) {
                found = true;
                break;
            }
        }
        if (!found) break;
    }
}

// Whether this lane's bit is set in a wave-uniform lane mask.
__device__ __forceinline__ bool lane_in(uint64_t mask) { return (mask >> (threadIdx.x & 63)) & 1u; }

// Packet traversal in front-to-back order (packet_walk) with the exact FP64
// re-test of every sphere the FP32 filter cannot exclude run as the walk meets
// it.  Reaching the near hits first shrinks every lane's bound sooner, so fewer
// far subtrees are entered; every sphere that can win is still evaluated in
// FP64 on every lane that needs it, so the answer is nearest_packet's.
template <int kLds = 0, class C = uint64_t>
__device__ __forceinline__ void nearest_packet_ordered(const SceneView& S, const V3& o, const V3& d, int32_t last,
                                                       double tol, bool live, Query& q, C& boxes, C& spheres,
                                                       C& retests, PAState* pa = nullptr) {
    typedef PacketMem<kLds> M;
    query_init(q);
    if (__ballot(live) == 0) return;
    const Ray32 r = ray32(S, o, d);
    float bestf = INFINITY;
    packet_walk<kLds>(S, r, live, bestf, boxes, [&](int32_t g, uint64_t mask) {
        const bool hit = lane_in(mask);
        const int32_t k1 = M::offset(S, g + 1);
        for (int32_t k = M::offset(S, g); k < k1; ++k) {
            const int32_t wanted = M::id(S, k);
            const float4 s = M::sphere(S, k);
            if (!hit || wanted == last) continue;
            ++spheres;
            if (!needs_exact(S, r, s, q.best, tol, pa)) continue;
            ++retests;
            const int32_t before = q.atom;
            eval64(o, d, load_sphere64(S.sphs, k), wanted, tol, q, pa);
            if (q.atom != before) bestf = box_bound(S, q.best, tol);
        }
    });
}

// Candidate lists of the two-pass query: kCand (slot, tlo) pairs per thread of a
// block of at most 256 threads, pair j of thread t at [j][t] (consecutive lanes
// read consecutive 8-byte words).  Row kCand takes the writes of lanes that have
// no candidate to add or whose list is full, so that appends need no branch.
#ifndef EHSS_PACKET_CANDIDATES
#define EHSS_PACKET_CANDIDATES 8
#endif
constexpr int kCand = EHSS_PACKET_CANDIDATES;
struct CandLists {
    int2 e[kCand + 1][256];  // (slot, bits of the FP32 lower bound)
};
static __shared__ CandLists g_cand;

// Static LDS of the two-pass packet kernels; a staged scene (dynamic LDS) must
// fit beside it in a workgroup's 64 KB.
constexpr size_t kPacketStaticLds = sizeof(CandLists);

// Drops the thread's candidates whose lower bound exceeds ut; returns the count kept.
__device__ __forceinline__ int32_t compact_candidates(int tid, int32_t nc, float ut) {
    int32_t w = 0;
    for (int32_t j = 0; j < nc; ++j) {
        const int2 c = g_cand.e[j][tid];
        if (!(__int_as_float(c.y) > ut)) {
            g_cand.e[w][tid] = c;
            ++w;
        }
    }
    return w;
}

// An FP32 value >= u + ftol (u, ftol >= 0): the sum rounded to nearest, then
// scaled by 1 + 2^-22 (more than the two roundings can lose); FLT_MAX in place
// of +inf, so that an excluded sphere's key (+inf) never compares below it.
__device__ __forceinline__ float upper_sum(float u, float ftol) {
    return fminf((u + ftol) * (1.0f + 0x1p-22f), 3.40282347e+38f);
}

// Slot id read by one lane (its own slot; an LDS read when the scene is staged).
template <int kLds>
__device__ __forceinline__ int32_t lane_slot_id(const SceneView& S, int32_t k) {
    if constexpr (kLds) return ((LInt32)S.ids)[k];
    else return S.ids[k];
}

// Two-pass packet query, same answer as nearest_packet_ordered.
//
// Pass 1 walks the packet tree (packet_walk) in FP32 only.  For each sphere of
// an entered leaf every lane computes the projection p of x = o - c on the ray,
// the perpendicular part x - p d and disc = r^2 - |x - p d|^2 (FMA forms: each
// fused operation rounds once, within the rounding the FP32 margins m_p and
// m_disc allow for the separate operations).  A sphere is excluded on a lane
// for every FP64 outcome within tol when it is the last atom, behind, or a
// clear miss (needs_exact's tests); when every lane excludes it, one uniform
// branch skips it.  Otherwise each lane computes
//   - tlo, a lower bound of the FP64 entry (needs_exact's bound, with the square
//     root taken as x * rsq(x) scaled up by 1 + 2^-20: the hardware rsq is within
//     1 ulp and the two products round by 2^-24 each), as key, or +inf where
//     the sphere is excluded;
//   - thi, an upper bound of the FP64 entry when the sphere is a robust hit in
//     front (FP64 disc surely > 0 and tlo > 0, so the FP64 entry is
//     -projection - sqrt(disc / aa), finite), else +inf.  The square root's
//     lower bound is (disc - m) / sqrt(disc + m) <= sqrt(disc - m) with the same
//     rsq, scaled down by 1 - 2^-20; a slack of 2^-21 of the projection term
//     covers the FP32 roundings of the two differences.
// U, the smallest thi so far, bounds the FP64 best from above.  A sphere with
// key <= U + tol joins the lane's candidate list; boxes are pruned at U + tol.
//
// Pass 2 runs the exact FP64 re-tests in lock step, each lane on its own j-th
// candidate, skipping those whose tlo exceeds the final U + tol.
//
// Every sphere whose FP64 entry lies within best + tol has tlo <= U + tol at its
// visit and at the end, so it is evaluated; so is every near-threshold sphere
// that the grazing test and the fused PA test can see.  The winner, the tie
// rule, `second`, the grazing flag and the PA flags are therefore those of the
// one-pass query.  Compared with that query, the FP64 block runs once per
// candidate rank instead of once per sphere some lane needs, the walk carries
// only FP32 state, and control flow stays uniform except for the rare
// compaction of a full list.  A lane whose list overflows (kCand candidates
// left after dropping those a smaller U excludes) sets `overflowed`: its q is
// then not the answer, and the caller must have the query done by another path
// (the ray's state is unchanged, so any query path gives the same answer); its
// PA state is still exact when `decided`.
template <int kLds = 0, class C = uint64_t>
__device__ __forceinline__ void nearest_packet_2p(const SceneView& S, const V3& o, const V3& d, int32_t last,
                                                  double tol, bool live, Query& q, C& boxes, C& spheres,
                                                  C& retests, bool& overflowed, PAState* pa = nullptr) {
    typedef PacketMem<kLds> M;
    query_init(q);
    overflowed = false;
    if (__ballot(live) == 0) return;
    const int tid = (int)threadIdx.x;
    const float ftol = (float)tol * 1.01f;
    const float plim = S.m_p + ftol;
    const float k2 = 2.01f * ftol;
    int32_t nc = 0;
    int32_t over = 0;
    int32_t decided = 0;
    float U = INFINITY;          // FP32 upper bound of the FP64 winning entry
    float ut = 3.40282347e+38f;  // >= U + tol (FLT_MAX while U is +inf)
    {
        const Ray32 r = ray32(S, o, d);
        float bestf = INFINITY;
        packet_walk<kLds>(S, r, live, bestf, boxes, [&](int32_t g, uint64_t mask) {
            if (!lane_in(mask)) return;
            const int32_t k0 = M::offset(S, g), k1 = M::offset(S, g + 1);
            // Every sphere of the leaf is tested on the lane (the last atom's test
            // runs in lock step with the others and is then discarded).
            spheres += (uint32_t)(k1 - k0);
            for (int32_t k = k0; k < k1; ++k) {
                const int32_t wanted = M::id(S, k);
                const float4 s = M::sphere(S, k);
                const float x0 = r.ox - s.x, x1 = r.oy - s.y, x2 = r.oz - s.z;
                const float p = __builtin_fmaf(x2, r.dz, __builtin_fmaf(x1, r.dy, x0 * r.dx));
                const float q0 = __builtin_fmaf(-p, r.dx, x0), q1 = __builtin_fmaf(-p, r.dy, x1),
                            q2 = __builtin_fmaf(-p, r.dz, x2);
                const float disc = __builtin_fmaf(s.w, s.w, -__builtin_fmaf(q2, q2, __builtin_fmaf(q1, q1, q0 * q0)));
                const float mdisc = __builtin_fmaf(s.w, k2, S.m_disc);
                const bool other = wanted != last;
                // Excluded on every lane (the common case): one uniform branch.
                const bool keep = other && p < plim && disc >= -mdisc;
                if (__ballot(keep) == 0) continue;
                if (pa) decided |= (keep && disc > mdisc && p <= -plim) ? 1 : 0;  // robust hit in both tests
                const float xh = fmaxf(disc, 0.f) + S.m_disc;
                const float ri = __builtin_amdgcn_rsqf(xh);
                const float tlo = (-p - S.m_p) - xh * ri * (1.0f + 0x1p-20f);
                const float a = S.m_p - p;
                const float thi0 = (a - (disc - S.m_disc) * ri * (1.0f - 0x1p-20f)) + a * 0x1p-21f;
                const float key = keep ? tlo : INFINITY;
                float thi = disc > S.m_disc ? thi0 : INFINITY;
                thi = tlo > 0.f ? thi : INFINITY;
                thi = keep ? thi : INFINITY;
                const bool ok = key <= ut;
                if (__ballot(ok && nc == kCand)) {
                    if (ok && nc == kCand) nc = compact_candidates(tid, nc, ut);
                }
                const bool room = nc < kCand;
                g_cand.e[ok && room ? nc : kCand][tid] = make_int2(k, __float_as_int(tlo));
                nc += ok && room ? 1 : 0;
                over |= ok && !room ? 1 : 0;
                U = fminf(U, thi);
                ut = upper_sum(U, ftol);
                bestf = ut * (1.0f + 4.0f * kU32) + S.m_box;
            }
        });
    }
    if (pa && decided) pa->decided = pa->hit = true;
    for (int32_t j = 0;; ++j) {
        const bool act = j < nc && !over;
        if (__ballot(act) == 0) break;
        const int2 c = g_cand.e[act ? j : kCand][tid];
        const bool run = act && !(__int_as_float(c.y) > ut);
        count_lanes(retests, 1, run);
        if (run) eval64(o, d, S.sphs[c.x], lane_slot_id<kLds>(S, c.x), tol, q, pa);
    }
    overflowed = over != 0;
}

// Cooperative nearest query: one ray, a group of G lanes (G a power of two <= 64,
// groups aligned within the wave).  The ray's state must be the same on every
// lane of the group.  Up to kBruteGroups groups, lane `sub` scans the leaf boxes
// sub, sub + G, ... with four independent loads in flight; larger scenes have
// lane `sub` walk the subtrees rooted at frontier entries sub, sub + G, ...
// Either way each lane follows the nearest_fast rules with its own running
// best, and the group then merges its per-lane queries.  Every sphere that can
// win (exact entry within best + tol) is
// evaluated in FP64 on some lane, and the winner is the lexicographic minimum of
// (entry, atom id), so the answer equals nearest_fast's; `second` and `graze`
// stay conservative.
__device__ __forceinline__ void group_merge_query(Query& q, int G) {
    for (int off = G >> 1; off > 0; off >>= 1) {
        const double ob = __shfl_xor(q.best, off, kWave);
        const int32_t oa = __shfl_xor(q.atom, off, kWave);
        const double os = __shfl_xor(q.second, off, kWave);
        const double og = __shfl_xor(q.graze, off, kWave);
        q.second = fmin(fmin(q.second, os), fmax(q.best, ob));
        if (better(ob, oa, q.best, q.atom)) {
            q.best = ob;
            q.atom = oa;
        }
        q.graze = fmin(q.graze, og);
    }
}

// Spheres of one leaf (slots first .. first + count - 1) against the lane's query.
template <class C = uint64_t>
__device__ __forceinline__ void scan_leaf(const SceneView& S, const Ray32& r, const V3& o, const V3& d,
                                          int32_t first, int32_t count, int32_t last, double tol, Query& q,
                                          float& bestf, C& spheres, C& retests) {
    for (int32_t k = first; k < first + count; ++k) {
        const int32_t wanted = S.ids[k];
        if (wanted == last) continue;
        ++spheres;
        if (!needs_exact(S, r, S.sph32[k], q.best, tol, nullptr)) continue;
        ++retests;
        const int32_t before = q.atom;
        eval64(o, d, S.sphs[k], wanted, tol, q, nullptr);
        if (q.atom != before) bestf = box_bound(S, q.best, tol);
    }
}

template <class C = uint64_t>
__device__ __forceinline__ void nearest_group(const SceneView& S, const V3& o, const V3& d, int32_t last, double tol,
                                              int G, int sub, Query& q, C& boxes, C& spheres,
                                              C& retests) {
    query_init(q);
    const Ray32 r = ray32(S, o, d);
    float bestf = INFINITY;
    const int32_t ng = S.n_groups;
    if (ng <= kBruteGroups) {
        // Every leaf box, four independent loads in flight per lane.
        for (int32_t g0 = sub; g0 < ng; g0 += 4 * G) {
            Node32 leaf[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (g0 + u * G < ng) leaf[u] = S.leaves32[g0 + u * G];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (g0 + u * G >= ng) break;
                ++boxes;
                if (box32(r, leaf[u], bestf))
                    scan_leaf(S, r, o, d, leaf[u].end, leaf[u].leaf, last, tol, q, bestf, spheres, retests);
            }
        }
    } else {
        for (int32_t e = sub; e < S.n_frontier; e += G) {
            int32_t node = S.frontier[e];
            const int32_t stop = S.nodes32[node].end;
            while (node < stop) {
                const Node32 n = S.nodes32[node];
                ++boxes;
                if (!box32(r, n, bestf)) {
                    node = n.end;
                    continue;
                }
                const int32_t g = n.leaf;
                if (g >= 0) scan_leaf(S, r, o, d, S.offsets[g], S.offsets[g + 1] - S.offsets[g], last, tol, q, bestf,
                                      spheres, retests);
                ++node;
            }
        }
    }
    group_merge_query(q, G);
}

// Two-pass form of nearest_group (drain kernels, at most 256 threads per
// block).  The group's lanes split the leaf boxes as nearest_group does, but a
// sphere of an entered leaf gets only the FP32 test of nearest_packet_2p: the
// lane keeps (slot, tlo) when tlo <= U + tol, U being the smallest upper bound
// thi of a robust hit so far, shared over the group after every round of
// 4 G leaf boxes.  The FP64 re-tests then run in lock step, each lane on its
// own j-th candidate, skipping those above the group's final U + tol; the
// group merges as before.  nearest_group instead runs each lane's FP64 re-tests
// inside its own scan, so the wave issues the FP64 block at nearly every sphere
// position some lane reaches.  A lane whose kGroupCand rows are full after
// dropping what its bound excludes re-tests them at once and starts over (a
// superset of the needed re-tests, so the answer is unchanged).  Every sphere
// whose FP64 entry lies within best + tol is re-tested on some lane (the
// argument of nearest_packet_2p), so the answer is nearest_group's.
#ifndef EHSS_GROUP_CANDIDATES
#define EHSS_GROUP_CANDIDATES 4
#endif
constexpr int kGroupCand = EHSS_GROUP_CANDIDATES;
struct GroupCandLists {
    int2 e[kGroupCand][256];
};
static __shared__ GroupCandLists g_gcand;

// Minimum over the G lanes of an aligned group, on every lane of the group.
__device__ __forceinline__ float group_min(float v, int G) {
    for (int off = G >> 1; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, kWave));
    return v;
}

template <class C = uint64_t>
__device__ __forceinline__ void nearest_group_2p(const SceneView& S, const V3& o, const V3& d, int32_t last,
                                                 double tol, int G, int sub, Query& q, C& boxes, C& spheres,
                                                 C& retests) {
    query_init(q);
    const Ray32 r = ray32(S, o, d);
    const int tid = (int)threadIdx.x;
    const float ftol = (float)tol * 1.01f;
    const float plim = S.m_p + ftol;
    const float k2 = 2.01f * ftol;
    float U = INFINITY;          // FP32 upper bound of the FP64 winning entry
    float ut = 3.40282347e+38f;  // >= U + tol (FLT_MAX while U is +inf)
    float bestf = INFINITY;      // box pruning bound
    int32_t nc = 0;
    // The lane's candidates with tlo above ut: dropped; if the rows are still
    // full, re-tested now.
    auto make_room = [&]() {
        int32_t w = 0;
        for (int32_t j = 0; j < nc; ++j) {
            const int2 c = g_gcand.e[j][tid];
            if (!(__int_as_float(c.y) > ut)) g_gcand.e[w++][tid] = c;
        }
        nc = w;
        if (nc < kGroupCand) return;
        for (int32_t j = 0; j < nc; ++j) {
            const int2 c = g_gcand.e[j][tid];
            ++retests;
            eval64(o, d, S.sphs[c.x], S.ids[c.x], tol, q, nullptr);
        }
        nc = 0;
    };
    auto filter_leaf = [&](int32_t first, int32_t count) {
        for (int32_t k = first; k < first + count; ++k) {
            const int32_t wanted = S.ids[k];
            if (wanted == last) continue;
            ++spheres;
            const float4 s = S.sph32[k];
            const float x0 = r.ox - s.x, x1 = r.oy - s.y, x2 = r.oz - s.z;
            const float p = __builtin_fmaf(x2, r.dz, __builtin_fmaf(x1, r.dy, x0 * r.dx));
            const float q0 = __builtin_fmaf(-p, r.dx, x0), q1 = __builtin_fmaf(-p, r.dy, x1),
                        q2 = __builtin_fmaf(-p, r.dz, x2);
            const float disc = __builtin_fmaf(s.w, s.w, -__builtin_fmaf(q2, q2, __builtin_fmaf(q1, q1, q0 * q0)));
            const float mdisc = __builtin_fmaf(s.w, k2, S.m_disc);
            if (!(p < plim && disc >= -mdisc)) continue;  // behind or a clear miss
            const float xh = fmaxf(disc, 0.f) + S.m_disc;
            const float ri = __builtin_amdgcn_rsqf(xh);
            const float tlo = (-p - S.m_p) - xh * ri * (1.0f + 0x1p-20f);
            const float a = S.m_p - p;
            const float thi0 = (a - (disc - S.m_disc) * ri * (1.0f - 0x1p-20f)) + a * 0x1p-21f;
            float thi = disc > S.m_disc ? thi0 : INFINITY;
            thi = tlo > 0.f ? thi : INFINITY;
            if (tlo <= ut) {
                if (nc == kGroupCand) make_room();
                g_gcand.e[nc++][tid] = make_int2(k, __float_as_int(tlo));
            }
            U = fminf(U, thi);
            ut = upper_sum(U, ftol);
            bestf = ut * (1.0f + 4.0f * kU32) + S.m_box;
        }
    };
    const int32_t ng = S.n_groups;
    if (ng <= kBruteGroups) {
        // Rounds of 4 G leaf boxes (the same count on every lane), four loads in
        // flight per lane; the bound is shared over the group after each round.
        for (int32_t gb = 0; gb < ng; gb += 4 * G) {
            const int32_t g0 = gb + sub;
            Node32 leaf[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (g0 + u * G < ng) leaf[u] = S.leaves32[g0 + u * G];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (g0 + u * G >= ng) break;
                ++boxes;
                if (box32(r, leaf[u], bestf)) filter_leaf(leaf[u].end, leaf[u].leaf);
            }
            U = group_min(U, G);
            ut = upper_sum(U, ftol);
            bestf = ut * (1.0f + 4.0f * kU32) + S.m_box;
        }
    } else {
        for (int32_t e = sub; e < S.n_frontier; e += G) {
            int32_t node = S.frontier[e];
            const int32_t stop = S.nodes32[node].end;
            while (node < stop) {
                const Node32 n = S.nodes32[node];
                ++boxes;
                if (!box32(r, n, bestf)) {
                    node = n.end;
                    continue;
                }
                const int32_t g = n.leaf;
                if (g >= 0) filter_leaf(S.offsets[g], S.offsets[g + 1] - S.offsets[g]);
                ++node;
            }
        }
        U = group_min(U, G);
        ut = upper_sum(U, ftol);
    }
    for (int32_t j = 0;; ++j) {
        const bool act = j < nc;
        if (__ballot(act) == 0) break;
        const int2 c = g_gcand.e[act ? j : 0][tid];
        if (act && !(__int_as_float(c.y) > ut)) {
            ++retests;
            eval64(o, d, S.sphs[c.x], S.ids[c.x], tol, q, nullptr);
        }
    }
    group_merge_query(q, G);
}

// The whole wave on one ray.
template <class C = uint64_t>
__device__ __forceinline__ void nearest_coop(const SceneView& S, const V3& o, const V3& d, int32_t last, double tol,
                                             Query& q, C& boxes, C& spheres, C& retests) {
    nearest_group(S, o, d, last, tol, kWave, threadIdx.x & (kWave - 1), q, boxes, spheres, retests);
}

// pa_spatial.py:114-130.  Counters follow the reference (stop at the first hit);
// after a near-threshold first hit the scan continues, uncounted, looking for a
// robust hit so that the grazing flag does not depend on traversal order.
template <class C = uint64_t>
__device__ __forceinline__ bool pa_plain(const SceneView& S, const V3& o, const V3& d, int32_t owner, double tol,
                                         C& boxes, C& tests, bool& graze) {
    int32_t node = 0;
    const int32_t stop = S.pa[0].end;
    bool hit = false, counting = true, critical_seen = false;
    while (node < stop) {
        const NodePA& n = S.pa[node];
        if (counting) ++boxes;
        if (!box_hit64(o, d, n.lo, n.hi)) {
            node = n.end;
            continue;
        }
        for (int32_t k = n.first; k < n.first + n.count; ++k) {
            const int32_t j = S.pa_ids[k];
            if (j == owner) continue;
            if (counting) ++tests;
            const double4 c = S.sph[j];
            bool critical;
            const bool h = sphere_hit64_graze(o, d, c.x, c.y, c.z, c.w, tol, &critical);
            critical_seen |= critical;
            if (h) {
                hit = true;
                counting = false;
                if (!critical) {
                    graze = false;
                    return true;
                }
            }
        }
        ++node;
    }
    graze = critical_seen;
    return hit;
}

// Any-hit through the FP32 filter on the static tree (any conservative tree gives
// the same flag).  With x = o - c, b = x.d, cc = |x|^2 - r^2, disc = b^2 - cc:
// clear miss when disc < -m; robust hit when disc > m and b <= -m_p (then
// -b + sqrt(disc) > 0); clear miss when b >= m_p and cc > m (origin outside,
// centre behind); otherwise the FP64 test decides.  Margins include tol.
template <class C = uint64_t>
__device__ __forceinline__ bool pa_fast(const SceneView& S, const V3& o, const V3& d, int32_t owner, double tol,
                                        C& boxes, C& tests, C& retests, bool& graze) {
    const Ray32 r = ray32(S, o, d);
    const float ftol = (float)tol * 1.01f;
    const float m_p = S.m_p + ftol;
    int32_t node = 0;
    const int32_t nn = S.n_nodes;
    bool hit = false, critical_seen = false;
    while (node < nn) {
        const Node32 n = S.nodes32[node];
        ++boxes;
        if (!box32(r, n, INFINITY)) {
            node = n.end;
            continue;
        }
        const int32_t g = n.leaf;
        if (g >= 0) {
            const int32_t k1 = S.offsets[g + 1];
            for (int32_t k = S.offsets[g]; k < k1; ++k) {
                const int32_t j = S.ids[k];
                if (j == owner) continue;
                ++tests;
                const float4 s = S.sph32[k];
                const float m = S.m_disc + 2.01f * s.w * ftol;
                const float x0 = r.ox - s.x, x1 = r.oy - s.y, x2 = r.oz - s.z;
                const float b = x0 * r.dx + x1 * r.dy + x2 * r.dz;
                const float cc = (x0 * x0 + x1 * x1 + x2 * x2) - s.w * s.w;
                const float disc = b * b - cc;
                if (disc < -m) continue;
                if (disc > m && b <= -m_p) {
                    graze = false;
                    return true;
                }
                if (b >= m_p && cc > m + 4.0f * (float)S.scale * ftol) continue;
                ++retests;
                const double4 c = S.sphs[k];
                bool critical;
                const bool h = sphere_hit64_graze(o, d, c.x, c.y, c.z, c.w, tol, &critical);
                critical_seen |= critical;
                if (h) {
                    hit = true;
                    if (!critical) {
                        graze = false;
                        return true;
                    }
                }
            }
        }
        ++node;
    }
    graze = critical_seen;
    return hit;
}

}  // namespace ehss
