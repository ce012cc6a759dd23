// ehss_device.h -- device arithmetic shared by the gfx950 kernels.
//
// The FP64 functions restate the reference scalar kernel operation for
// operation (src/native/ehss_cpu.cpp:7-40).  The whole library is compiled with
// -ffp-contract=off, and IEEE division and square root are used, so these
// return the same bits as the reference on the same inputs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace ehss {

constexpr double kEps64 = 2.220446049250313e-16;  // DBL_EPSILON
constexpr float kU32 = 5.9604645e-08f;            // 2^-24, FP32 unit roundoff
constexpr int kInfo = 0;
constexpr int kWave = 64;                         // wavefront width on gfx950
constexpr int kFrontier = 8 * kWave;              // cooperative-query subtree roots, at most
// Cooperative queries scan every leaf box up to kBruteGroups groups: that beats
// walking the frontier subtrees (a chain of dependent loads per lane) for every
// scene measured on MI355X, up to 8000 groups (24000 atoms).
constexpr int32_t kBruteGroups = 16384;

// Leaves of the static group tree (the halving tree over k groups: a node over
// groups [a, b) has children over [a, a + (b - a) / 2) and the rest) when every
// subtree of at most `cut` groups is made one leaf.  The sizes at one depth are
// q and q + 1; their counts follow the halving in closed form.
__host__ __device__ inline int32_t cut_leaves(int32_t k, int32_t cut) {
    int64_t q = k, cq = 1, cq1 = 0, leaves = 0;
    for (;;) {
        if (q <= cut) {
            leaves += cq;
            cq = 0;
        }
        if (q + 1 <= cut) {
            leaves += cq1;
            cq1 = 0;
        }
        if (cq + cq1 == 0) break;
        const int64_t h = q / 2;  // q = 2h -> (h, h), 2h + 1 -> (h, h + 1), 2h + 2 -> (h + 1, h + 1)
        const int64_t nq = (q % 2 == 0) ? 2 * cq + cq1 : cq;
        const int64_t nq1 = (q % 2 == 0) ? cq1 : cq + 2 * cq1;
        q = h;
        cq = nq;
        cq1 = nq1;
    }
    return (int32_t)leaves;
}

// Nodes of that cut tree (a full binary tree).
__host__ __device__ inline int32_t cut_nodes(int32_t k, int32_t cut) { return 2 * cut_leaves(k, cut) - 1; }

// Groups per leaf of the packet tree, the static tree cut as above, that the
// packet (wave-uniform) queries walk: at most kPacketLeafAtoms / (atoms per
// group) groups per leaf.  Any tree whose boxes enclose their spheres gives the
// same answers under the FP32 filter.  Measured at 16 x 2^20 rays, cap 64: on
// the 650-atom globule static trees of leaf size 2 / 4 / 8 / 16 run 10.72 /
// 10.10 / 9.69 / 10.28 ms per launch, and the leaf-4 tree cut at 4 / 6 / 10
// atoms runs 10.07 / 9.71 / 9.62 (10.06 uncut); at 10 atoms C60, the
// 2000-atom ellipsoid and globules of 272, 829 and 4241 atoms run 4.19, 13.17,
// 5.65, 8.78 and 24.67 ms against 4.28, 13.24, 5.73, 8.76 and 24.89 uncut.
#ifndef EHSS_PACKET_LEAF_ATOMS
#define EHSS_PACKET_LEAF_ATOMS 10
#endif
constexpr int32_t kPacketLeafAtoms = EHSS_PACKET_LEAF_ATOMS;
__host__ __device__ inline int32_t packet_cut(int32_t n_atoms, int32_t n_groups) {
    if (n_groups <= 1 || n_atoms <= 0) return 1;
    const int64_t c = (int64_t)kPacketLeafAtoms * n_groups / n_atoms;
    return c < 1 ? 1 : (c > 16 ? 16 : (int32_t)c);
}

// std::max / std::min semantics: the first argument wins unless strictly beaten.
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }

struct V3 {
    double x, y, z;
};

// Static EHSS tree node (FP64), reference layout (ehss_static.py:84-90):
// padded box, preorder skip pointer end[node], leaf group (or -1).
struct alignas(64) Node64 {
    double lo[3];
    double hi[3];
    int32_t end;
    int32_t leaf;
    int32_t pad[2];
};

// FP32 filter node: outward-rounded box inflated by the scene's filter margin.
// `leaf` is the leaf's group, or for an internal node -2 - (its right child), so
// that a traversal can fetch both children at once (the left child is node + 1).
struct alignas(32) Node32 {
    float lo[3];
    float hi[3];
    int32_t end;
    int32_t leaf;
};

// The encoding is its own inverse: right_child(right_child(r)) == r.
__host__ __device__ constexpr int32_t right_child(int32_t internal_leaf) { return -2 - internal_leaf; }

// PA any-hit tree node (pa_spatial.py:17-28): unpadded box, skip pointer,
// first id slot and count (count 0 for internal nodes).
struct alignas(64) NodePA {
    double lo[3];
    double hi[3];
    int32_t end;
    int32_t first;
    int32_t count;
    int32_t pad;
};

// Grazing bookkeeping.  A decision is "grazing" when an input perturbation of
// size tol (a distance) could flip it: tol starts at kGrazeTol times the scene
// scale and grows along the path with a first-order bound on how far rounding
// differences in the source could have moved the ray (see PathError).
constexpr double kGrazeTol = 1e-9;

// src/native/ehss_cpu.cpp:7-17.  Also lowers *graze_t to the along-ray distance of
// a decision within tol of its threshold: tangency (|disc| <= 2 r tol, sphere not
// behind) or the -64 eps r snap (|near - snap| <= tol).  The sign of the
// projection can only flip the outcome together with a near-zero disc.
__device__ __forceinline__ double entry64(const V3& o, const V3& d, double cx, double cy, double cz, double r,
                                          double tol, double* graze_t) {
    double x0 = o.x - cx, x1 = o.y - cy, x2 = o.z - cz;
    double aa = (d.x * d.x + d.y * d.y) + d.z * d.z;
    double projection = ((x0 * d.x + x1 * d.y) + x2 * d.z) / aa;
    double p0 = x0 - projection * d.x, p1 = x1 - projection * d.y, p2 = x2 - projection * d.z;
    double disc = r * r - ((p0 * p0 + p1 * p1) + p2 * p2);
    const double tdisc = 2 * r * tol;
    const bool tangent = fabs(disc) <= tdisc;
    if (disc < 0. || projection >= 0.) {
        if (tangent && projection < tol) *graze_t = fmin(*graze_t, fmax(-projection, 0.));
        return INFINITY;
    }
    double near = -projection - sqrt(smax(0., disc) / aa);
    double snap = -64 * kEps64 * r;
    if (tangent) *graze_t = fmin(*graze_t, -projection);
    if (fabs(near - snap) <= tol) *graze_t = 0.;
    if (near < snap) return INFINITY;
    return smax(0., near);
}

// First-order bound on the displacement of a ray (position ep, direction ed)
// caused by rounding-level differences in its source, propagated through free
// flights and specular reflections: a flight of length t adds t*ed to ep; a
// reflection at incidence cosine c off radius r maps ep -> ep / c and adds
// 2 ep / r to ed.  Used only to widen the grazing tolerance.
struct PathError {
    double ep, ed;
};

__device__ __forceinline__ PathError path_error_start(double scale) {
    return PathError{16 * kEps64 * scale, 16 * kEps64};
}

__device__ __forceinline__ double path_tol(const PathError& e, double scale) {
    return fmax(kGrazeTol * scale, 8.0 * (e.ep + 2.0 * scale * e.ed));
}

__device__ __forceinline__ void path_error_reflect(PathError& e, double t, double cos_inc, double r, double scale) {
    const double c = fmax(fabs(cos_inc), 1e-6);
    const double ep = (e.ep + t * e.ed) / c + 4 * kEps64 * scale;
    e.ed = e.ed + 2 * ep / r + 4 * kEps64;
    e.ep = ep;
}

// src/native/ehss_cpu.cpp:19-29.  *near_out receives the slab entry distance.
__device__ __forceinline__ bool box64(const V3& o, const V3& d, const double* lo, const double* hi,
                                      double best, double* near_out = nullptr) {
    double near = 0., far = INFINITY;
    const double oo[3] = {o.x, o.y, o.z};
    const double dd[3] = {d.x, d.y, d.z};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (dd[c] == 0.) {
            if (oo[c] < lo[c] || oo[c] > hi[c]) return false;
        } else {
            double a = (lo[c] - oo[c]) / dd[c], b = (hi[c] - oo[c]) / dd[c];
            near = smax(near, smin(a, b));
            far = smin(far, smax(a, b));
        }
    }
    if (near_out) *near_out = (far >= near && far >= 0.) ? near : INFINITY;
    return far >= near && near <= best && far >= 0.;
}

// src/native/ehss_cpu.cpp:31-40 (in place on o, d).
__device__ __forceinline__ double reflect64(V3& o, V3& d, double cx, double cy, double cz, double r,
                                            double distance) {
    double n0 = o.x + distance * d.x - cx, n1 = o.y + distance * d.y - cy, n2 = o.z + distance * d.z - cz;
    double norm = sqrt((n0 * n0 + n1 * n1) + n2 * n2);
    n0 /= norm;
    n1 /= norm;
    n2 /= norm;
    o.x = cx + r * n0;
    o.y = cy + r * n1;
    o.z = cz + r * n2;
    double dot = (d.x * n0 + d.y * n1) + d.z * n2;
    d.x = d.x - 2 * dot * n0;
    d.y = d.y - 2 * dot * n1;
    d.z = d.z - 2 * dot * n2;
    norm = sqrt((d.x * d.x + d.y * d.y) + d.z * d.z);
    d.x /= norm;
    d.y /= norm;
    d.z /= norm;
    return dot;  // incidence cosine (unit incoming direction)
}

// Nearest-atom update rule of src/native/ehss_cpu.cpp:67 (tie: lower id wins).
__device__ __forceinline__ bool better(double distance, int32_t wanted, double best, int32_t atom) {
    return distance < best || (distance == best && isfinite(distance) && (atom < 0 || wanted < atom));
}

// inside_out_pa.py:93-101
__device__ __forceinline__ bool sphere_hit64(const V3& o, const V3& d, double cx, double cy, double cz, double r) {
    double x = o.x - cx, y = o.y - cy, z = o.z - cz;
    double b = (x * d.x + y * d.y) + z * d.z;
    double cc = ((x * x + y * y) + z * z) - r * r;
    double disc = b * b - cc;
    return disc >= 0 && -b + sqrt(smax(0., disc)) > 0;
}

// pa_spatial.py:98-111
__device__ __forceinline__ bool box_hit64(const V3& o, const V3& d, const double* lo, const double* hi) {
    double near = 0., far = INFINITY;
    const double oo[3] = {o.x, o.y, o.z};
    const double dd[3] = {d.x, d.y, d.z};
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {
        if (dd[axis] == 0) {
            if (oo[axis] < lo[axis] || oo[axis] > hi[axis]) return false;
        } else {
            double a = (lo[axis] - oo[axis]) / dd[axis], b = (hi[axis] - oo[axis]) / dd[axis];
            near = smax(near, smin(a, b));
            far = smin(far, smax(a, b));
            if (far < near) return false;
        }
    }
    return true;
}

// Wave-level helpers (wave64).
__device__ __forceinline__ int64_t wave_sum_i64(int64_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace ehss
