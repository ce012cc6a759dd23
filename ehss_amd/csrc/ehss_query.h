// ehss_query.h -- per-lane nearest-hit and any-hit queries over a device scene.
//
// nearest_plain: ehss_cpu.cpp:59-72 (FP64 box tests, reference traversal and counters).
// nearest_fast:  the same query through an FP32 filter with an exact FP64 re-test of
//                every sphere the filter cannot exclude (identical result).
// pa_plain:      pa_spatial.py:114-130 over the PA tree (FP64, reference counters).
// pa_fast:       any-hit through the FP32 filter over the static tree, FP64 re-test
//                of undecided spheres (identical hit flag).
#pragma once
#include "ehss_internal.h"

namespace ehss {

struct Query {
    int32_t atom;
    double best;
    double second;  // second-smallest finite entry distance (nearest-atom tie test)
    double graze;   // smallest along-ray distance of a near-threshold sphere test
};

__device__ __forceinline__ void consider(Query& q, double t, int32_t wanted) {
    if (better(t, wanted, q.best, q.atom)) {
        q.second = q.best;
        q.best = t;
        q.atom = wanted;
    } else if (t < q.second) {
        q.second = t;
    }
}

// A ray is grazing when a near-threshold sphere test lies at or before the
// winning distance, or when the runner-up is within tolerance of the winner.
__device__ __forceinline__ bool grazing(const Query& q, double tol) {
    return q.graze <= q.best + tol || q.second - q.best <= tol;
}

__device__ __forceinline__ void nearest_plain(const SceneView& S, const V3& o, const V3& d, int32_t last, Query& q,
                                              uint64_t& boxes, uint64_t& spheres) {
    q.atom = -1;
    q.best = INFINITY;
    q.second = INFINITY;
    q.graze = INFINITY;
    int32_t node = 0;
    const int32_t nn = S.n_nodes;
    while (node < nn) {
        const Node64& n = S.nodes[node];
        ++boxes;
        if (!box64(o, d, n.lo, n.hi, q.best)) {
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
                const double t = entry64(o, d, c.x, c.y, c.z, c.w, &q.graze);
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
__device__ __forceinline__ bool box32(const Ray32& r, const Node32& n, float bestf) {
    const float tx0 = __builtin_fmaf(n.lo[0], r.ix, -r.oix), tx1 = __builtin_fmaf(n.hi[0], r.ix, -r.oix);
    const float ty0 = __builtin_fmaf(n.lo[1], r.iy, -r.oiy), ty1 = __builtin_fmaf(n.hi[1], r.iy, -r.oiy);
    const float tz0 = __builtin_fmaf(n.lo[2], r.iz, -r.oiz), tz1 = __builtin_fmaf(n.hi[2], r.iz, -r.oiz);
    const float near = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.f));
    const float far = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
    return near <= far && near <= bestf;
}

__device__ __forceinline__ void nearest_fast(const SceneView& S, const V3& o, const V3& d, int32_t last, Query& q,
                                             uint64_t& boxes, uint64_t& spheres, uint64_t& retests) {
    q.atom = -1;
    q.best = INFINITY;
    q.second = INFINITY;
    q.graze = INFINITY;
    const Ray32 r = ray32(S, o, d);
    const float m_p = S.m_p, m_disc = S.m_disc, m_box = S.m_box;
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
                const float4 s = S.sph32[k];
                const float x0 = r.ox - s.x, x1 = r.oy - s.y, x2 = r.oz - s.z;
                const float p = x0 * r.dx + x1 * r.dy + x2 * r.dz;
                if (p >= m_p) continue;  // projection64 > 0: miss
                const float p0 = x0 - p * r.dx, p1 = x1 - p * r.dy, p2 = x2 - p * r.dz;
                const float disc = s.w * s.w - (p0 * p0 + p1 * p1 + p2 * p2);
                if (disc < -m_disc) continue;  // disc64 < 0: miss
                const float tlo = (-p - m_p) - sqrtf(fmaxf(disc, 0.f) + m_disc);
                if ((double)tlo > q.best) continue;  // entry64 > best: cannot win or tie
                ++retests;
                const double4 c = S.sph[wanted];
                const double t = entry64(o, d, c.x, c.y, c.z, c.w, &q.graze);
                const int32_t before = q.atom;
                consider(q, t, wanted);
                if (q.atom != before) bestf = (float)q.best * (1.0f + 4.0f * kU32) + m_box;
            }
        }
        ++node;
    }
}

// FP64 any-hit sphere test (inside_out_pa.py:93-101) with a flag for decisions a
// small perturbation could flip: hit = (disc >= 0) and (exit > 0), so it is
// critical when disc is within tol r^2 of 0 while the exit condition (-b > 0 at
// disc = 0) nearly holds, or when exit is within tol r of 0 while disc nearly
// holds.
__device__ __forceinline__ bool sphere_hit64_graze(const V3& o, const V3& d, double cx, double cy, double cz, double r,
                                                   bool* critical) {
    double x = o.x - cx, y = o.y - cy, z = o.z - cz;
    double b = (x * d.x + y * d.y) + z * d.z;
    double cc = ((x * x + y * y) + z * z) - r * r;
    double disc = b * b - cc;
    double exit = -b + sqrt(smax(0., disc));
    const double tr = kGrazeTol * r, trr = kGrazeTol * r * r;
    *critical = (fabs(disc) <= trr && -b > -tr) || (fabs(exit) <= tr && disc >= -trr);
    return disc >= 0 && exit > 0;
}

// pa_spatial.py:114-130.  Counters follow the reference (stop at the first hit);
// after a near-threshold first hit the scan continues, uncounted, looking for a
// robust hit so that the grazing flag does not depend on traversal order.
__device__ __forceinline__ bool pa_plain(const SceneView& S, const V3& o, const V3& d, int32_t owner,
                                         uint64_t& boxes, uint64_t& tests, bool& graze) {
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
            const bool h = sphere_hit64_graze(o, d, c.x, c.y, c.z, c.w, &critical);
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
// definite miss when disc < -m; definite hit when disc > m and b <= -m_p
// (then -b + sqrt(disc) > 0); definite miss when b >= m_p and cc > m (origin
// outside, centre behind); otherwise the FP64 test decides.
__device__ __forceinline__ bool pa_fast(const SceneView& S, const V3& o, const V3& d, int32_t owner,
                                        uint64_t& boxes, uint64_t& tests, uint64_t& retests, bool& graze) {
    const Ray32 r = ray32(S, o, d);
    const float m_p = S.m_p, m = S.m_disc;
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
                const float x0 = r.ox - s.x, x1 = r.oy - s.y, x2 = r.oz - s.z;
                const float b = x0 * r.dx + x1 * r.dy + x2 * r.dz;
                const float cc = (x0 * x0 + x1 * x1 + x2 * x2) - s.w * s.w;
                const float disc = b * b - cc;
                if (disc < -m) continue;
                if (disc > m && b <= -m_p) {
                    graze = false;
                    return true;
                }
                if (b >= m_p && cc > m) continue;
                ++retests;
                const double4 c = S.sph[j];
                bool critical;
                const bool h = sphere_hit64_graze(o, d, c.x, c.y, c.z, c.w, &critical);
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
