/*
 * ehss_oracle.c -- CPU restatement of the reference EHSS/PA hot path, in plain C.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: it is compiled into
 * oracle/liboracle.so and may be loaded only by tests/, by the smoke check in
 * __graft_entry__.py and by the cpu_baseline leg of bench.py.  The product
 * library (ehss_amd/libehss_hip.so) never links or calls it.
 *
 * Every function restates one reference function; the citation names the file
 * and lines (relative to the reference checkout) whose arithmetic it follows.
 * Operation order and parenthesisation follow the reference exactly, and the
 * file is compiled with -ffp-contract=off and without fast-math so that each
 * FP64 result is bit-identical to the reference's.
 *
 * Pinning (how each function is checked; see tests/test_oracle_*.py):
 *   oracle_trace        == oracle/_ref build of src/native/ehss_cpu.cpp (bitwise),
 *                          closed-form reflection, trapped pair, tangency grid.
 *   oracle_trace_flat   == oracle_trace (traversal independence, bitwise).
 *   oracle_trace_allatom within 1e-10 of oracle_trace (independent twin).
 *   oracle_groups       == PA-tree leaf order from oracle_pa_build (two builders).
 *   oracle_static_tree  closed-form topology algebra, box nesting/containment.
 *   oracle_pa_query     == brute-force any-hit over every sphere.
 *   oracle_solve_cpu    == the NumPy-path oracle on the same u at C1 2^16 and C2 2^20
 *                          rays: omega within 1e-12, hit and unresolved counts and
 *                          counters exact (tests/test_oracle_source.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <omp.h>

/* std::max / std::min semantics (first argument wins unless strictly beaten). */
static inline double smax(double a, double b) { return (a < b) ? b : a; }
static inline double smin(double a, double b) { return (b < a) ? b : a; }

/* ---------------------------------------------------------------------------
 * Scalar geometry.
 * ------------------------------------------------------------------------- */

/* src/native/ehss_cpu.cpp:7-17 (== ehss_cpu.py:12-24, ehss_reference.py:8-23) */
static double entry(const double* o, const double* d, const double* c, double r) {
    double x0 = o[0] - c[0], x1 = o[1] - c[1], x2 = o[2] - c[2];
    double aa = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
    double projection = ((x0 * d[0] + x1 * d[1]) + x2 * d[2]) / aa;
    double p0 = x0 - projection * d[0], p1 = x1 - projection * d[1], p2 = x2 - projection * d[2];
    double disc = r * r - ((p0 * p0 + p1 * p1) + p2 * p2);
    if (disc < 0. || projection >= 0.) return INFINITY;
    double near = -projection - sqrt(smax(0., disc) / aa);
    if (near < -64 * DBL_EPSILON * r) return INFINITY;
    return smax(0., near);
}

/* src/native/ehss_cpu.cpp:19-29 (== ehss_support_tree.py:19-31 used at ehss_cpu.py:48-49) */
static int box(const double* o, const double* d, const double* lo, const double* hi, double best) {
    double near = 0., far = INFINITY;
    for (int c = 0; c < 3; ++c) {
        if (d[c] == 0.) {
            if (o[c] < lo[c] || o[c] > hi[c]) return 0;
        } else {
            double a = (lo[c] - o[c]) / d[c], b = (hi[c] - o[c]) / d[c];
            near = smax(near, smin(a, b));
            far = smin(far, smax(a, b));
        }
    }
    return far >= near && near <= best && far >= 0.;
}

/* src/native/ehss_cpu.cpp:31-40 (== ehss_cpu.py:27-40) */
static void reflect(double* o, double* d, const double* c, double r, double distance) {
    double n0 = o[0] + distance * d[0] - c[0], n1 = o[1] + distance * d[1] - c[1],
           n2 = o[2] + distance * d[2] - c[2];
    double norm = sqrt((n0 * n0 + n1 * n1) + n2 * n2);
    n0 /= norm; n1 /= norm; n2 /= norm;
    o[0] = c[0] + r * n0; o[1] = c[1] + r * n1; o[2] = c[2] + r * n2;
    double dot = (d[0] * n0 + d[1] * n1) + d[2] * n2;
    d[0] = d[0] - 2 * dot * n0; d[1] = d[1] - 2 * dot * n1; d[2] = d[2] - 2 * dot * n2;
    norm = sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
    d[0] /= norm; d[1] /= norm; d[2] /= norm;
}

double oracle_entry(const double* o, const double* d, const double* c, double r) { return entry(o, d, c, r); }
int oracle_box(const double* o, const double* d, const double* lo, const double* hi, double best) {
    return box(o, d, lo, hi, best);
}
void oracle_reflect(double* o, double* d, const double* c, double r, double t) { reflect(o, d, c, r, t); }

/* ---------------------------------------------------------------------------
 * Trace loops.
 * ------------------------------------------------------------------------- */

/* Nearest hit over the stackless preorder tree: src/native/ehss_cpu.cpp:59-72
 * (== ehss_cpu.py:43-63).  Returns the atom (-1 for none) and writes best. */
static int64_t nearest_tree(const double* o, const double* v, int64_t last, int64_t nn,
                            const double* lo, const double* hi, const int64_t* end, const int64_t* leaf,
                            const int64_t* offsets, const int64_t* ids, const double* world,
                            const double* radii, double* best_out, int64_t* boxes, int64_t* spheres) {
    int64_t atom = -1, node = 0;
    double best = INFINITY;
    while (node < nn) {
        ++*boxes;
        if (!box(o, v, lo + 3 * node, hi + 3 * node, best)) { node = end[node]; continue; }
        int64_t g = leaf[node];
        if (g >= 0)
            for (int64_t k = offsets[g]; k < offsets[g + 1]; ++k) {
                int64_t wanted = ids[k];
                if (wanted == last) continue;
                ++*spheres;
                double distance = entry(o, v, world + 3 * wanted, radii[wanted]);
                if (distance < best || (distance == best && isfinite(distance) && (atom < 0 || wanted < atom))) {
                    atom = wanted; best = distance;
                }
            }
        ++node;
    }
    *best_out = best;
    return atom;
}

/* Flat (all-atom) search with the same tie rule: ehss_native_ablation.py:35-47.
 * Traversal-independence twin of nearest_tree. */
static int64_t nearest_flat(const double* o, const double* v, int64_t last, int64_t natoms,
                            const double* world, const double* radii, double* best_out, int64_t* spheres) {
    int64_t atom = -1;
    double best = INFINITY;
    for (int64_t wanted = 0; wanted < natoms; ++wanted) {
        if (wanted == last) continue;
        ++*spheres;
        double distance = entry(o, v, world + 3 * wanted, radii[wanted]);
        if (distance < best || (distance == best && isfinite(distance) && (atom < 0 || wanted < atom))) {
            atom = wanted; best = distance;
        }
    }
    *best_out = best;
    return atom;
}

/* src/native/ehss_cpu.cpp:42-82, the array contract of ehss_trace (== ehss_cpu.py:66-100).
 * mode 0: BVH search (nn nodes); mode 1: flat search over natoms spheres. */
static void trace_impl(int mode, int64_t nr, int64_t nn, int64_t natoms, int cap, int threads, int record,
                       double* p, double* d, const int64_t* lasts, int64_t* counts, uint8_t* escaped,
                       uint8_t* unresolved, int32_t* history, const double* lo, const double* hi,
                       const int64_t* end, const int64_t* leaf, const int64_t* offsets, const int64_t* ids,
                       const double* world, const double* radii, int64_t* counters) {
    int64_t boxes = 0, spheres = 0;
    if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(static) reduction(+ : boxes, spheres)
    for (int64_t ray = 0; ray < nr; ++ray) {
        double* o = p + 3 * ray;
        double* v = d + 3 * ray;
        int64_t last = lasts[ray];
        if (record && counts[ray] == 1) history[ray * cap] = (int32_t)last;
        for (;;) {
            double best;
            int64_t atom = mode == 0
                ? nearest_tree(o, v, last, nn, lo, hi, end, leaf, offsets, ids, world, radii, &best, &boxes, &spheres)
                : nearest_flat(o, v, last, natoms, world, radii, &best, &spheres);
            if (atom < 0) { escaped[ray] = 1; break; }
            if (counts[ray] >= cap) { unresolved[ray] = 1; break; }
            reflect(o, v, world + 3 * atom, radii[atom], best);
            if (record) history[ray * cap + counts[ray]] = (int32_t)atom;
            ++counts[ray];
            last = atom;
        }
    }
    counters[0] = boxes;
    counters[1] = spheres;
}

void oracle_trace(int64_t nr, int64_t nn, int cap, int threads, int record, double* p, double* d,
                  const int64_t* lasts, int64_t* counts, uint8_t* escaped, uint8_t* unresolved, int32_t* history,
                  const double* lo, const double* hi, const int64_t* end, const int64_t* leaf,
                  const int64_t* offsets, const int64_t* ids, const double* world, const double* radii,
                  int64_t* counters) {
    trace_impl(0, nr, nn, 0, cap, threads, record, p, d, lasts, counts, escaped, unresolved, history, lo, hi, end,
               leaf, offsets, ids, world, radii, counters);
}

void oracle_trace_flat(int64_t nr, int64_t natoms, int cap, int threads, int record, double* p, double* d,
                       const int64_t* lasts, int64_t* counts, uint8_t* escaped, uint8_t* unresolved,
                       int32_t* history, const double* world, const double* radii, int64_t* counters) {
    trace_impl(1, nr, 0, natoms, cap, threads, record, p, d, lasts, counts, escaped, unresolved, history, NULL,
               NULL, NULL, NULL, NULL, NULL, world, radii, counters);
}

/* Independent all-atom reference: ehss_reference.py:26-64.  First strict minimum
 * in index order; the reflection moves o first (ehss_reference.py:57-62). */
void oracle_trace_allatom(int64_t nr, int64_t natoms, int cap, double* p, double* d, const int64_t* lasts,
                          int64_t* counts, uint8_t* escaped, uint8_t* unresolved, int32_t* history,
                          const double* world, const double* radii) {
    for (int64_t ray = 0; ray < nr; ++ray) {
        double* o = p + 3 * ray;
        double* v = d + 3 * ray;
        int64_t last = lasts[ray];
        if (counts[ray] == 1) history[ray * cap] = (int32_t)last;
        for (;;) {
            double best = INFINITY;
            int64_t atom = -1;
            for (int64_t j = 0; j < natoms; ++j) {
                if (j == last) continue;
                double t = entry(o, v, world + 3 * j, radii[j]);
                if (t < best) { best = t; atom = j; }
            }
            if (atom < 0) { escaped[ray] = 1; break; }
            if (counts[ray] >= cap) { unresolved[ray] = 1; break; }
            const double* c = world + 3 * atom;
            for (int k = 0; k < 3; ++k) o[k] = o[k] + best * v[k];
            double n[3] = {o[0] - c[0], o[1] - c[1], o[2] - c[2]};
            double nrm = sqrt((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]);
            for (int k = 0; k < 3; ++k) n[k] /= nrm;
            for (int k = 0; k < 3; ++k) o[k] = c[k] + radii[atom] * n[k];
            double dot = (v[0] * n[0] + v[1] * n[1]) + v[2] * n[2];
            for (int k = 0; k < 3; ++k) v[k] = v[k] - 2 * dot * n[k];
            nrm = sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
            for (int k = 0; k < 3; ++k) v[k] /= nrm;
            history[ray * cap + counts[ray]] = (int32_t)atom;
            ++counts[ray];
            last = atom;
        }
    }
}

/* ---------------------------------------------------------------------------
 * Static grouping and the flattened EHSS tree.
 * ------------------------------------------------------------------------- */

/* Stable merge sort of idx[0..n) by key[idx*3+axis] (numpy kind='mergesort'). */
static void stable_sort_by(int64_t* idx, int64_t n, const double* key, int axis, int64_t* tmp) {
    if (n < 2) return;
    for (int64_t width = 1; width < n; width *= 2) {
        for (int64_t lo = 0; lo < n; lo += 2 * width) {
            int64_t mid = lo + width < n ? lo + width : n;
            int64_t hi = lo + 2 * width < n ? lo + 2 * width : n;
            int64_t i = lo, j = mid, k = lo;
            while (i < mid && j < hi) {
                /* take the right element only when strictly smaller: stability */
                if (key[idx[j] * 3 + axis] < key[idx[i] * 3 + axis]) tmp[k++] = idx[j++];
                else tmp[k++] = idx[i++];
            }
            while (i < mid) tmp[k++] = idx[i++];
            while (j < hi) tmp[k++] = idx[j++];
        }
        memcpy(idx, tmp, (size_t)n * sizeof(int64_t));
    }
}

/* ehss_static.py:53-59: keys = ((c - r) + (c + r)) / 2, elementwise. */
void oracle_group_keys(int64_t n, const double* centers, const double* radii, double* keys) {
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            double lo = centers[3 * i + c] - radii[i], hi = centers[3 * i + c] + radii[i];
            keys[3 * i + c] = (lo + hi) / 2;
        }
}

/* ehss_static.py:16-50 (_median_groups).  Returns the number of groups, or -1. */
int64_t oracle_groups(int64_t n, const double* keys, int64_t leaf_size, int64_t* groups) {
    int64_t* ids = malloc((size_t)n * sizeof(int64_t));
    int64_t* tmp = malloc((size_t)n * sizeof(int64_t));
    if (!ids || !tmp) { free(ids); free(tmp); return -1; }
    for (int64_t i = 0; i < n; ++i) ids[i] = i;
    int64_t starts[64], stops[64], depths[64];
    starts[0] = 0; stops[0] = n; depths[0] = 0;
    int pending = 1;
    int64_t group = 0;
    while (pending) {
        --pending;
        int64_t a = starts[pending], b = stops[pending], depth = depths[pending];
        double span[3];
        for (int c = 0; c < 3; ++c) {
            double low = INFINITY, high = -INFINITY;
            for (int64_t k = a; k < b; ++k) {
                double v = keys[ids[k] * 3 + c];
                low = smin(low, v);   /* Python min(low, v) */
                high = smax(high, v); /* Python max(high, v) */
            }
            span[c] = high - low;
        }
        int axis = 0; /* np.argmax: first maximum */
        if (span[1] > span[axis]) axis = 1;
        if (span[2] > span[axis]) axis = 2;
        if (b - a <= leaf_size || depth >= 24 || span[axis] == 0) {
            for (int64_t k = a; k < b; ++k) groups[ids[k]] = group;
            ++group;
            continue;
        }
        stable_sort_by(ids + a, b - a, keys, axis, tmp);
        int64_t mid = a + (b - a) / 2;
        if (pending + 2 > 64) { free(ids); free(tmp); return -1; }
        starts[pending] = mid; stops[pending] = b; depths[pending] = depth + 1;
        starts[pending + 1] = a; stops[pending + 1] = mid; depths[pending + 1] = depth + 1;
        pending += 2;
    }
    free(ids);
    free(tmp);
    return group;
}

/* ehss_cpu_geometry.py:19-38 (topology) for ng groups; arrays of 2*ng-1 nodes. */
void oracle_topology(int64_t ng, int64_t* left, int64_t* right, int64_t* end, int64_t* parent, int64_t* leaf,
                     int64_t* leaf_node, int64_t* depth) {
    int64_t nodes = 2 * ng - 1;
    int64_t* begin = calloc((size_t)nodes, sizeof(int64_t));
    int64_t* stop = calloc((size_t)nodes, sizeof(int64_t));
    for (int64_t k = 0; k < nodes; ++k) { left[k] = right[k] = parent[k] = leaf[k] = -1; depth[k] = 0; }
    stop[0] = ng;
    for (int64_t node = 0; node < nodes; ++node) {
        int64_t a = begin[node], b = stop[node];
        end[node] = node + 2 * (b - a) - 1;
        if (b - a == 1) {
            leaf[node] = a;
            leaf_node[a] = node;
        } else {
            int64_t mid = a + (b - a) / 2;
            int64_t l = node + 1, r = node + 2 * (mid - a);
            left[node] = l; right[node] = r;
            begin[l] = a; stop[l] = mid; begin[r] = mid; stop[r] = b;
            parent[l] = node; parent[r] = node;
            depth[l] = depth[node] + 1; depth[r] = depth[node] + 1;
        }
    }
    free(begin);
    free(stop);
}

/* ehss_static.py:80-88 + ehss_cpu_geometry.py:41-71: offsets, ids, padded leaf
 * boxes and parent unions for the static tree over ng groups. */
void oracle_static_tree(int64_t n, int64_t ng, const int64_t* groups, const double* world, const double* radii,
                        double* lo, double* hi, int64_t* end, int64_t* leaf, int64_t* offsets, int64_t* ids) {
    int64_t nodes = 2 * ng - 1;
    int64_t* left = malloc((size_t)nodes * sizeof(int64_t));
    int64_t* right = malloc((size_t)nodes * sizeof(int64_t));
    int64_t* parent = malloc((size_t)nodes * sizeof(int64_t));
    int64_t* depth = malloc((size_t)nodes * sizeof(int64_t));
    int64_t* leaf_node = malloc((size_t)ng * sizeof(int64_t));
    oracle_topology(ng, left, right, end, parent, leaf, leaf_node, depth);
    /* offsets = r_[0, cumsum(bincount(groups))]; ids = argsort(groups, stable) */
    for (int64_t g = 0; g <= ng; ++g) offsets[g] = 0;
    for (int64_t i = 0; i < n; ++i) offsets[groups[i] + 1]++;
    for (int64_t g = 0; g < ng; ++g) offsets[g + 1] += offsets[g];
    int64_t* fill = malloc((size_t)ng * sizeof(int64_t));
    for (int64_t g = 0; g < ng; ++g) fill[g] = offsets[g];
    for (int64_t i = 0; i < n; ++i) ids[fill[groups[i]]++] = i;
    double rmax = radii[0];
    for (int64_t i = 1; i < n; ++i) rmax = radii[i] > rmax ? radii[i] : rmax;
    double pad = 1e-10 * smax(1., rmax);
    for (int64_t g = 0; g < ng; ++g) {
        int64_t node = leaf_node[g];
        for (int c = 0; c < 3; ++c) {
            double low = INFINITY, high = -INFINITY;
            for (int64_t k = offsets[g]; k < offsets[g + 1]; ++k) {
                int64_t atom = ids[k];
                low = smin(low, world[3 * atom + c] - radii[atom]);
                high = smax(high, world[3 * atom + c] + radii[atom]);
            }
            lo[3 * node + c] = low - pad;
            hi[3 * node + c] = high + pad;
        }
    }
    for (int64_t node = nodes - 1; node >= 0; --node)
        if (left[node] >= 0)
            for (int c = 0; c < 3; ++c) {
                lo[3 * node + c] = smin(lo[3 * left[node] + c], lo[3 * right[node] + c]);
                hi[3 * node + c] = smax(hi[3 * left[node] + c], hi[3 * right[node] + c]);
            }
    free(left); free(right); free(parent); free(depth); free(leaf_node); free(fill);
}

/* ---------------------------------------------------------------------------
 * PA any-hit tree: pa_spatial.py:38-75 (build_tree, kind='bvh').
 * ------------------------------------------------------------------------- */

typedef struct {
    const double *plo, *phi, *cen;
    int64_t leaf_size;
    double *lo, *hi;
    int64_t *end, *first, *count, *ids;
    int64_t nodes, nids;
    int64_t* tmp;
} pa_builder;

static void pa_visit(pa_builder* B, int64_t* idx, int64_t len, int depth) {
    int64_t index = B->nodes++;
    for (int c = 0; c < 3; ++c) {
        double l = INFINITY, h = -INFINITY;
        for (int64_t k = 0; k < len; ++k) {
            double a = B->plo[3 * idx[k] + c], b = B->phi[3 * idx[k] + c];
            if (a < l) l = a;
            if (b > h) h = b;
        }
        B->lo[3 * index + c] = l;
        B->hi[3 * index + c] = h;
    }
    B->end[index] = 0;
    B->first[index] = B->nids;
    B->count[index] = 0;
    double span[3];
    for (int c = 0; c < 3; ++c) {
        double l = INFINITY, h = -INFINITY;
        for (int64_t k = 0; k < len; ++k) {
            double v = B->cen[3 * idx[k] + c];
            if (v < l) l = v;
            if (v > h) h = v;
        }
        span[c] = h - l;
    }
    if (len <= B->leaf_size || depth >= 24 || (span[0] == 0 && span[1] == 0 && span[2] == 0)) {
        B->count[index] = len;
        for (int64_t k = 0; k < len; ++k) B->ids[B->nids++] = idx[k];
    } else {
        int axis = 0;
        if (span[1] > span[axis]) axis = 1;
        if (span[2] > span[axis]) axis = 2;
        stable_sort_by(idx, len, B->cen, axis, B->tmp);
        int64_t middle = len / 2;
        pa_visit(B, idx, middle, depth + 1);
        pa_visit(B, idx + middle, len - middle, depth + 1);
    }
    B->end[index] = B->nodes;
}

/* Builds the PA tree over spheres (centers, radii).  Arrays need room for 2n
 * nodes and n ids.  Returns the node count. */
int64_t oracle_pa_build(int64_t n, const double* centers, const double* radii, int64_t leaf_size, double* lo,
                        double* hi, int64_t* end, int64_t* first, int64_t* count, int64_t* ids) {
    double* plo = malloc((size_t)n * 3 * sizeof(double));
    double* phi = malloc((size_t)n * 3 * sizeof(double));
    double* cen = malloc((size_t)n * 3 * sizeof(double));
    int64_t* idx = malloc((size_t)n * sizeof(int64_t));
    int64_t* tmp = malloc((size_t)n * sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) {
        idx[i] = i;
        for (int c = 0; c < 3; ++c) {
            plo[3 * i + c] = centers[3 * i + c] - radii[i];
            phi[3 * i + c] = centers[3 * i + c] + radii[i];
            cen[3 * i + c] = (plo[3 * i + c] + phi[3 * i + c]) / 2;
        }
    }
    pa_builder B = {plo, phi, cen, leaf_size, lo, hi, end, first, count, ids, 0, 0, tmp};
    pa_visit(&B, idx, n, 0);
    free(plo); free(phi); free(cen); free(idx); free(tmp);
    return B.nodes;
}

/* inside_out_pa.py:93-101 */
static int sphere_hit(const double* o, const double* d, const double* c, double r) {
    double x = o[0] - c[0], y = o[1] - c[1], z = o[2] - c[2];
    double b = (x * d[0] + y * d[1]) + z * d[2];
    double cc = ((x * x + y * y) + z * z) - r * r;
    double disc = b * b - cc;
    return disc >= 0 && -b + sqrt(smax(0., disc)) > 0;
}

/* pa_spatial.py:98-111 */
static int box_hit(const double* o, const double* d, const double* lo, const double* hi) {
    double near = 0., far = INFINITY;
    for (int axis = 0; axis < 3; ++axis) {
        if (d[axis] == 0) {
            if (o[axis] < lo[axis] || o[axis] > hi[axis]) return 0;
        } else {
            double a = (lo[axis] - o[axis]) / d[axis], b = (hi[axis] - o[axis]) / d[axis];
            near = smax(near, smin(a, b));
            far = smin(far, smax(a, b));
            if (far < near) return 0;
        }
    }
    return 1;
}

int oracle_sphere_hit(const double* o, const double* d, const double* c, double r) { return sphere_hit(o, d, c, r); }

/* pa_spatial.py:114-142 (_one + tree_blocked, root 0, no excluded group). */
void oracle_pa_query(int64_t nr, const double* origins, const double* directions, const int64_t* owners,
                     const double* centers, const double* radii, const double* lo, const double* hi,
                     const int64_t* end, const int64_t* first, const int64_t* count, const int64_t* ids,
                     int threads, uint8_t* hits, int64_t* counters) {
    int64_t boxes = 0, tests = 0;
    if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(static) reduction(+ : boxes, tests)
    for (int64_t i = 0; i < nr; ++i) {
        const double* o = origins + 3 * i;
        const double* d = directions + 3 * i;
        int64_t node = 0, stop = end[0];
        uint8_t hit = 0;
        while (node < stop) {
            ++boxes;
            if (!box_hit(o, d, lo + 3 * node, hi + 3 * node)) { node = end[node]; continue; }
            for (int64_t k = first[node]; k < first[node] + count[node]; ++k) {
                int64_t j = ids[k];
                if (j != owners[i]) {
                    ++tests;
                    if (sphere_hit(o, d, centers + 3 * j, radii[j])) { hit = 1; break; }
                }
            }
            if (hit) break;
            ++node;
        }
        hits[i] = hit;
    }
    counters[0] = boxes;
    counters[1] = tests;
}

/* Brute-force any-hit over every sphere (no tree): the pin for oracle_pa_query. */
void oracle_pa_brute(int64_t nr, const double* origins, const double* directions, const int64_t* owners,
                     int64_t natoms, const double* centers, const double* radii, uint8_t* hits) {
    for (int64_t i = 0; i < nr; ++i) {
        uint8_t hit = 0;
        for (int64_t j = 0; j < natoms && !hit; ++j)
            if (j != owners[i] && sphere_hit(origins + 3 * i, directions + 3 * i, centers + 3 * j, radii[j])) hit = 1;
        hits[i] = hit;
    }
}

/* ---------------------------------------------------------------------------
 * Whole-path CPU solve, for the CPU baseline timing: Sobol -> source -> PA ->
 * trace -> readout, FP64, OpenMP over rays.
 * ------------------------------------------------------------------------- */

/* Scrambled Sobol point i in Gray-code order (SciPy qmc.Sobol semantics, 30 bits):
 * q = shift ^ XOR_{bit j of gray(i)} sv[j]; u = q * 2^-30. */
static void sobol_point(const uint32_t* sv /*4x30*/, const uint32_t* shift, uint64_t i, double* u) {
    uint64_t g = i ^ (i >> 1);
    for (int dim = 0; dim < 4; ++dim) {
        uint32_t q = shift[dim];
        for (int j = 0; j < 30; ++j)
            if ((g >> j) & 1u) q ^= sv[dim * 30 + j];
        u[dim] = (double)q * (1.0 / 1073741824.0);
    }
}

/* NumPy evaluates cos and sin separately.  Called directly, the compiler merges
 * the pair into one sincos(), whose last bits differ from the separate calls on
 * rare angles; calling through volatile pointers keeps them separate. */
static double (*volatile libm_cos)(double) = cos;
static double (*volatile libm_sin)(double) = sin;

/* ehss_state.py:42-52 with inside_out_pa.py:45-58, per ray, in NumPy's operation order. */
static void source_ray(const double* u, const double* center, double radius, double* o, double* d) {
    const double two_pi = 2 * M_PI;
    double z = 1 - 2 * u[0];
    double phi = two_pi * u[1];
    double rho = sqrt(smax(0., 1 - z * z));
    d[0] = rho * libm_cos(phi); d[1] = rho * libm_sin(phi); d[2] = z;
    double ax[3] = {0., 0., 1.};
    if (fabs(d[2]) > .9) { ax[0] = 1.; ax[1] = 0.; ax[2] = 0.; }
    double t[3] = {ax[1] * d[2] - ax[2] * d[1], ax[2] * d[0] - ax[0] * d[2], ax[0] * d[1] - ax[1] * d[0]};
    double tn = sqrt((t[0] * t[0] + t[1] * t[1]) + t[2] * t[2]);
    t[0] /= tn; t[1] /= tn; t[2] /= tn;
    double b[3] = {d[1] * t[2] - d[2] * t[1], d[2] * t[0] - d[0] * t[2], d[0] * t[1] - d[1] * t[0]};
    double r = radius * sqrt(u[2]);
    double ph = two_pi * u[3];
    double cp = libm_cos(ph), sp = libm_sin(ph);
    for (int k = 0; k < 3; ++k) o[k] = (center[k] - radius * d[k]) + r * (cp * t[k] + sp * b[k]);
}

/* out[0]=omega, out[1]=pa_any_hit, out[2]=pa_from_ehss, out[3]=residual_mass,
 * out[4]=unresolved count, out[5]=box tests, out[6]=sphere tests.  Weight w per ray. */
void oracle_solve_cpu(int64_t n_rays, const uint32_t* sv, const uint32_t* shift, const double* center,
                      double radius, double w, int cap, int threads, int do_pa, int64_t nn, const double* lo,
                      const double* hi, const int64_t* end, const int64_t* leaf, const int64_t* offsets,
                      const int64_t* ids, const double* world, const double* radii, int64_t pa_nodes,
                      const double* pa_lo, const double* pa_hi, const int64_t* pa_end, const int64_t* pa_first,
                      const int64_t* pa_count, const int64_t* pa_ids, double* out) {
    double omega = 0, pa = 0, pa_e = 0, resid = 0;
    int64_t unres = 0, boxes = 0, spheres = 0;
    (void)pa_nodes;
    if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(static) \
    reduction(+ : omega, pa, pa_e, resid, unres, boxes, spheres)
    for (int64_t i = 0; i < n_rays; ++i) {
        double u[4], o[3], d[3], inc[3];
        sobol_point(sv, shift, (uint64_t)i, u);
        source_ray(u, center, radius, o, d);
        inc[0] = d[0]; inc[1] = d[1]; inc[2] = d[2];
        if (do_pa) {
            int64_t node = 0, stop = pa_end[0];
            int hit = 0;
            while (node < stop && !hit) {
                if (!box_hit(o, d, pa_lo + 3 * node, pa_hi + 3 * node)) { node = pa_end[node]; continue; }
                for (int64_t k = pa_first[node]; k < pa_first[node] + pa_count[node]; ++k)
                    if (sphere_hit(o, d, world + 3 * pa_ids[k], radii[pa_ids[k]])) { hit = 1; break; }
                ++node;
            }
            if (hit) pa += w;
        }
        int64_t last = -1, count = 0;
        for (;;) {
            double best;
            int64_t atom = nearest_tree(o, d, last, nn, lo, hi, end, leaf, offsets, ids, world, radii, &best,
                                        &boxes, &spheres);
            if (atom < 0) {
                double dot = (inc[0] * d[0] + inc[1] * d[1]) + inc[2] * d[2];
                double angle = 1 - dot;
                angle = angle < 0. ? 0. : (angle > 2. ? 2. : angle);
                omega += w * angle;
                if (count > 0) pa_e += w;
                break;
            }
            if (count >= cap) { unres += 1; resid += w; pa_e += w; break; }
            reflect(o, d, world + 3 * atom, radii[atom], best);
            ++count;
            last = atom;
        }
    }
    out[0] = omega; out[1] = pa; out[2] = pa_e; out[3] = resid;
    out[4] = (double)unres; out[5] = (double)boxes; out[6] = (double)spheres;
}

int oracle_max_threads(void) { return omp_get_max_threads(); }

/* ---------------------------------------------------------------------------
 * Finite-grid PA (the CLI's adaptive PA): union of projected disks on an n x n
 * midpoint grid.
 * ------------------------------------------------------------------------- */

/* pa_dyadic_bitmask.py:55-66 (row_span): pixels [lo, hi) of row y whose midpoint
 * satisfies (lower0 + (x + .5) spacing - cx)^2 <= r^2 - dy^2, with the same
 * rounding of the first guess and the same refinement loops. */
static void row_span(double cx, double cy, double r, const double* lower, double spacing, int64_t y, int64_t n,
                     int64_t* lo_out, int64_t* hi_out) {
    double dy = lower[1] + ((double)y + .5) * spacing - cy;
    double rem = r * r - dy * dy;
    *lo_out = *hi_out = 0;
    if (rem < 0) return;
    double half = sqrt(rem);
    int64_t lo = (int64_t)ceil((cx - half - lower[0]) / spacing - .5);
    int64_t hi = (int64_t)floor((cx + half - lower[0]) / spacing - .5) + 1;
    lo = lo < n ? lo : n;
    lo = lo > 0 ? lo : 0;
    hi = hi < n ? hi : n;
    hi = hi > 0 ? hi : 0;
#define SQ(v) ((v) * (v))
    while (lo > 0 && SQ(lower[0] + ((double)lo - .5) * spacing - cx) <= rem) --lo;
    while (lo < hi && SQ(lower[0] + ((double)lo + .5) * spacing - cx) > rem) ++lo;
    while (hi < n && SQ(lower[0] + ((double)hi + .5) * spacing - cx) <= rem) ++hi;
    while (hi > lo && SQ(lower[0] + ((double)hi - .5) * spacing - cx) > rem) --hi;
#undef SQ
    *lo_out = lo;
    *hi_out = hi;
}

/* pa_dyadic_bitmask.py:68-72 (row_bounds): candidate rows [lo, hi), one row of
 * slack on each side of the rounded disk extent. */
static void row_bounds(double cy, double r, const double* lower, double spacing, int64_t n, int64_t* lo_out,
                       int64_t* hi_out) {
    int64_t lo = (int64_t)ceil((cy - r - lower[1]) / spacing - .5) - 1;
    int64_t hi = (int64_t)floor((cy + r - lower[1]) / spacing - .5) + 2;
    lo = lo < n ? lo : n;
    lo = lo > 0 ? lo : 0;
    hi = hi < n ? hi : n;
    hi = hi > 0 ? hi : 0;
    *lo_out = lo;
    *hi_out = hi;
}

void oracle_row_span(double cx, double cy, double r, const double* lower, double spacing, int64_t y, int64_t n,
                     int64_t* out) {
    row_span(cx, cy, r, lower, spacing, y, n, out, out + 1);
}

/* pa_dyadic_bitmask.py:86-97 (count_disks): number of grid pixels covered by
 * the union of the disks (the value block_count(block_disks(...)) returns,
 * pa_sparse_blocks.py:166-174).  rows_out (may be NULL) gets the row spans
 * evaluated. */
int64_t oracle_grid_count(int64_t na, const double* xy, const double* radii, const double* lower, double spacing,
                          int64_t n, int64_t* rows_out) {
    int64_t* delta = (int64_t*)calloc((size_t)(n * (n + 1)), sizeof(int64_t));
    int64_t rows = 0;
    for (int64_t a = 0; a < na; ++a) {
        int64_t first, stop;
        row_bounds(xy[2 * a + 1], radii[a], lower, spacing, n, &first, &stop);
        for (int64_t y = first; y < stop; ++y) {
            int64_t lo, hi;
            row_span(xy[2 * a], xy[2 * a + 1], radii[a], lower, spacing, y, n, &lo, &hi);
            ++rows;
            if (hi > lo) {
                delta[y * (n + 1) + lo] += 1;
                delta[y * (n + 1) + hi] -= 1;
            }
        }
    }
    int64_t count = 0;
    for (int64_t y = 0; y < n; ++y) {
        int64_t active = 0;
        for (int64_t x = 0; x < n; ++x) {
            active += delta[y * (n + 1) + x];
            count += active > 0;
        }
    }
    free(delta);
    if (rows_out) *rows_out = rows;
    return count;
}

/* Brute force over every pixel and every disk with the row_span inequality (an
 * independent twin of oracle_grid_count for small grids). */
int64_t oracle_grid_brute(int64_t na, const double* xy, const double* radii, const double* lower, double spacing,
                          int64_t n) {
    int64_t count = 0;
    for (int64_t y = 0; y < n; ++y)
        for (int64_t x = 0; x < n; ++x)
            for (int64_t a = 0; a < na; ++a) {
                double dy = lower[1] + ((double)y + .5) * spacing - xy[2 * a + 1];
                double rem = radii[a] * radii[a] - dy * dy;
                double dx = lower[0] + ((double)x + .5) * spacing - xy[2 * a];
                if (rem >= 0 && dx * dx <= rem) {
                    ++count;
                    break;
                }
            }
    return count;
}
