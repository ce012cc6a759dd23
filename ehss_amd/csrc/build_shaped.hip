// build_shaped.hip -- the tree build for structures whose atoms have distinct key triples.
//
// The median-split recursion (ehss_static.py:16-50, restated level by level in
// build.hip) makes a segment a leaf when it has at most leaf_size atoms, sits at
// depth 24, or has zero key span on every axis.  Zero span needs two atoms with
// the same key triple ((c - r) + (c + r)) / 2, so for a structure without such
// a pair a segment of size s at depth d splits iff s > leaf_size and d < 24,
// into sizes s / 2 and s - s / 2: the shape of the tree is a function of
// (n, leaf_size) alone (TreeShape), and only the split axes and the order of the
// atoms depend on the coordinates.  A walk of d steps from the root gives any
// segment's range, preorder number (the PA node index, pa_spatial.py:38-75) and
// first group (groups are numbered left to right), so every output lands at
// its final index without a pass over the whole tree, and the host knows the
// node and group counts before anything runs: the build is a short launch
// sequence on the scene's stream with no synchronisation.
//
//   k_shape_init    keys, FP64 spheres, the identity order
//   per top level   (only while segments have more than kFinishAtoms atoms)
//     k_top_axis    one workgroup per segment: key spans, split axis
//     k_top_sort    one workgroup per chunk: stable LDS sort by (key, position)
//     k_top_merge   rank across the segment's chunks by binary search, scatter
//   k_finish        one workgroup per segment of <= kFinishAtoms atoms: all its
//                   remaining levels in LDS, then its PA nodes, groups, offsets
//                   and eids, deepest level first
//   k_top_nodes     PA nodes of the top levels (unions of their children)
//   k_static_tree   the static tree over the groups (implicit topology,
//                   ehss_cpu_geometry.py:19-71): padded leaf boxes, parent unions
//
// Every sort is stable under (key, current position), as the reference's stable
// argsort of each segment; minima and maxima are exact, so the outputs equal
// build.hip's bit for bit.
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "ehss_internal.h"

namespace ehss {

TreeShape make_tree_shape(int32_t n, int32_t leaf) {
    TreeShape sh;
    memset(&sh, 0, sizeof sh);
    sh.n = n;
    sh.leaf = leaf;
    for (int d = 0; d < kShapeDepths; ++d) sh.q[d] = n >> d;
    int32_t deepest[kShapeDepths][2];
    for (int d = kShapeDepths - 1; d >= 0; --d) {
        for (int t = 0; t < 2; ++t) {
            const int32_t s = sh.q[d] + t;
            if (s <= leaf || d >= kMaxSplitDepth) {
                sh.leaves[d][t] = 1;
                deepest[d][t] = d;
            } else {
                const int32_t l = s / 2, r = s - l;
                const int32_t lt = l - sh.q[d + 1], rt = r - sh.q[d + 1];  // both 0 or 1
                sh.leaves[d][t] = sh.leaves[d + 1][lt] + sh.leaves[d + 1][rt];
                deepest[d][t] = std::max(deepest[d + 1][lt], deepest[d + 1][rt]);
            }
        }
    }
    sh.groups = sh.leaves[0][0];
    sh.max_depth = deepest[0][0];
    return sh;
}

bool has_duplicate_keys(int64_t n, const double* centers, const double* radii) {
    if (n < 2) return false;
    size_t cap = 1;
    while (cap < (size_t)(2 * n)) cap <<= 1;
    std::vector<int32_t> slot(cap, -1);
    std::vector<double> key((size_t)(3 * n));
    for (int64_t i = 0; i < n; ++i) {
        uint64_t h = 0x9E3779B97F4A7C15ull;
        for (int k = 0; k < 3; ++k) {
            const double lo = centers[3 * i + k] - radii[i], hi = centers[3 * i + k] + radii[i];
            const double v = (lo + hi) / 2;
            key[3 * i + k] = v;
            const double canon = v + 0.0;  // -0 and +0 compare equal: hash them alike
            uint64_t b;
            memcpy(&b, &canon, sizeof b);
            h = (h ^ b) * 0xBF58476D1CE4E5B9ull;
            h ^= h >> 31;
        }
        for (size_t p = (size_t)h & (cap - 1);; p = (p + 1) & (cap - 1)) {
            const int32_t j = slot[p];
            if (j < 0) {
                slot[p] = (int32_t)i;
                break;
            }
            if (key[3 * j] == key[3 * i] && key[3 * j + 1] == key[3 * i + 1] && key[3 * j + 2] == key[3 * i + 2])
                return true;
        }
    }
    return false;
}

std::vector<int32_t> frontier_of(int32_t ng) {
    struct Entry {
        int32_t node, a, b;  // node over groups [a, b)
    };
    std::vector<Entry> f{{0, 0, ng}}, next;
    for (;;) {
        next.clear();
        bool split = false;
        for (const Entry& e : f) {
            if (e.b - e.a == 1) {
                next.push_back(e);
            } else {
                const int32_t mid = e.a + (e.b - e.a) / 2;
                next.push_back({e.node + 1, e.a, mid});
                next.push_back({e.node + 2 * (mid - e.a), mid, e.b});
                split = true;
            }
        }
        if (!split || next.size() > (size_t)kFrontier) break;
        f.swap(next);
    }
    std::vector<int32_t> out;
    out.reserve(f.size());
    for (const Entry& e : f) out.push_back(e.node);
    return out;
}

namespace {

constexpr int kFinishAtoms = 4096;  // atoms of a segment finished in one workgroup's LDS
constexpr int kFinishThreads = 1024;
constexpr int kPer = kFinishAtoms / kFinishThreads;
constexpr int kChunk = 4096;        // positions per top-level sort chunk
constexpr int kChunkThreads = 1024;
constexpr int kWavesPerBlock = 1024 / kWave;
// Segments above the finishing depth have more than kFinishAtoms atoms, so they
// all split when kFinishAtoms >= every leaf size the shaped build takes.
static_assert(kFinishAtoms > kShapedMaxLeaf, "top-level segments must all split");

struct FrontierList {
    int32_t n;
    int32_t v[kFrontier];
};

__global__ void k_put_frontier(FrontierList f, int32_t* dst) {
    for (int i = threadIdx.x; i < f.n; i += blockDim.x) dst[i] = f.v[i];
}

// A segment of the shaped tree: positions [a, b), depth, PA preorder number,
// first group.
struct SegPos {
    int32_t a, b, depth, pre, group;
};

__device__ __forceinline__ bool seg_splits(const TreeShape& S, int32_t size, int32_t depth) {
    return size > S.leaf && depth < kMaxSplitDepth;
}

__device__ __forceinline__ int32_t seg_leaves(const TreeShape& S, int32_t size, int32_t depth) {
    return S.leaves[depth][size - S.q[depth]];
}

// Moves g to its left or right child (g splits): the left child follows g in
// preorder, the right child follows the left subtree's 2 leaves - 1 nodes.
__device__ __forceinline__ void seg_child(const TreeShape& S, SegPos& g, bool right) {
    const int32_t half = (g.b - g.a) / 2;
    if (right) {
        const int32_t lv = seg_leaves(S, half, g.depth + 1);
        g.pre += 2 * lv;
        g.group += lv;
        g.a += half;
    } else {
        g.pre += 1;
        g.b = g.a + half;
    }
    g.depth += 1;
}

// Descends `steps` levels along the bits of k (most significant first); false
// when a segment on the way is a leaf, i.e. there is no such segment.
__device__ __forceinline__ bool seg_path(const TreeShape& S, SegPos& g, int steps, int32_t k) {
    for (int t = steps - 1; t >= 0; --t) {
        if (!seg_splits(S, g.b - g.a, g.depth)) return false;
        seg_child(S, g, ((k >> t) & 1) != 0);
    }
    return true;
}

// Descends toward position p for at most `steps` levels, stopping at a leaf;
// returns the levels descended and the path bits in *k.
__device__ __forceinline__ int seg_find(const TreeShape& S, SegPos& g, int steps, int32_t p, int32_t* k) {
    int32_t bits = 0;
    int t = 0;
    for (; t < steps; ++t) {
        if (!seg_splits(S, g.b - g.a, g.depth)) break;
        const bool right = p >= g.a + (g.b - g.a) / 2;
        seg_child(S, g, right);
        bits = 2 * bits + (right ? 1 : 0);
    }
    *k = bits;
    return t;
}

__device__ __forceinline__ void load_shape(TreeShape& dst, const TreeShape& src) {
    const int32_t* s = reinterpret_cast<const int32_t*>(&src);
    int32_t* d = reinterpret_cast<int32_t*>(&dst);
    for (int i = threadIdx.x; i < (int)(sizeof(TreeShape) / 4); i += blockDim.x) d[i] = s[i];
}

__device__ __forceinline__ double pmin(double low, double v) { return (v < low) ? v : low; }   // Python min(low, v)
__device__ __forceinline__ double pmax(double high, double v) { return (high < v) ? v : high; } // Python max(high, v)

// np.argmax of the key spans: the first maximum.
__device__ __forceinline__ int split_axis(const double lo[3], const double hi[3]) {
    const double s0 = hi[0] - lo[0], s1 = hi[1] - lo[1], s2 = hi[2] - lo[2];
    int axis = 0;
    double best = s0;
    if (s1 > best) {
        axis = 1;
        best = s1;
    }
    if (s2 > best) axis = 2;
    return axis;
}

__device__ __forceinline__ void wave_minmax(double lo[3], double hi[3], int width) {
    for (int w = 1; w < width; w <<= 1) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            lo[c] = pmin(lo[c], __shfl_xor(lo[c], w));
            hi[c] = pmax(hi[c], __shfl_xor(hi[c], w));
        }
    }
}

// Stable merge sort of the positions of every active element's run set
// [ea, eb) by key_of(e, position), in LDS: perm holds positions in sorted order
// (identity on entry); round r merges neighbouring runs of 2^r.  An element of
// a left run is preceded by the right run's strictly smaller keys, an element
// of a right run by the left run's keys not greater than its own -- the order
// (key, position).  `big` bounds every run set's size.
template <int NPER, int NTHREADS, class KeyOf>
__device__ __forceinline__ void stable_merge(uint16_t* perm, const int32_t* ea, const int32_t* eb, const bool* act,
                                             int32_t big, KeyOf key_of) {
    for (int r = 0; (1 << r) < big; ++r) {
        int32_t np[NPER];
        uint16_t x[NPER];
#pragma unroll
        for (int e = 0; e < NPER; ++e) {
            const int32_t i = (int32_t)threadIdx.x + e * NTHREADS;
            np[e] = -1;
            x[e] = 0;
            if (act[e]) {
                x[e] = perm[i];
                const int32_t a = ea[e], b = eb[e];
                const int32_t run = (i - a) >> r;
                const int32_t ss = a + ((run ^ 1) << r);
                if (ss < b) {
                    const int32_t se = min(b, ss + (1 << r));
                    const double k = key_of(e, x[e]);
                    const bool strict = (run & 1) == 0;
                    int32_t lo = ss, hi = se;
                    while (lo < hi) {
                        const int32_t mid = (lo + hi) >> 1;
                        const double km = key_of(e, perm[mid]);
                        if (strict ? (km < k) : (km <= k))
                            lo = mid + 1;
                        else
                            hi = mid;
                    }
                    np[e] = a + ((run >> 1) << (r + 1)) + (i - (a + (run << r))) + (lo - ss);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < NPER; ++e)
            if (np[e] >= 0) perm[np[e]] = x[e];
        __syncthreads();
    }
}

__global__ void k_shape_init(int32_t n, int32_t ng, const double* c, const double* r, double* keys, double4* sph,
                             int32_t* ids, int32_t* offsets) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        x[k] = c[3 * i + k];
        const double lo = x[k] - r[i], hi = x[k] + r[i];
        keys[3 * i + k] = (lo + hi) / 2;
    }
    sph[i] = make_double4(x[0], x[1], x[2], r[i]);
    ids[i] = i;
    if (i == 0) offsets[ng] = n;
}

struct TopArgs {
    int32_t level;
    const double* keys;  // n x 3
    const int32_t* in;   // position -> atom before this level's splits
    int32_t* out;        // position -> atom after them
    uint8_t* axis;       // segment (path bits) -> split axis
    double* skey;        // chunk-sorted keys
    int32_t* spos;       // chunk-sorted positions
};

// One workgroup per segment of the level (all split): key spans, split axis.
__global__ void __launch_bounds__(1024) k_top_axis(TopArgs A, TreeShape shape) {
    __shared__ TreeShape S;
    __shared__ double part[kWavesPerBlock][6];
    load_shape(S, shape);
    __syncthreads();
    SegPos s{0, S.n, 0, 0, 0};
    seg_path(S, s, A.level, (int32_t)blockIdx.x);
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int32_t i = s.a + (int32_t)threadIdx.x; i < s.b; i += blockDim.x) {
        const int32_t atom = A.in[i];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double v = A.keys[3 * atom + c];
            lo[c] = pmin(lo[c], v);
            hi[c] = pmax(hi[c], v);
        }
    }
    wave_minmax(lo, hi, kWave);
    const int wave = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            part[wave][c] = lo[c];
            part[wave][3 + c] = hi[c];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x / kWave); ++w) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                lo[c] = pmin(lo[c], part[w][c]);
                hi[c] = pmax(hi[c], part[w][3 + c]);
            }
        }
        A.axis[blockIdx.x] = (uint8_t)split_axis(lo, hi);
    }
}

// One workgroup per chunk of kChunk positions of a segment (blockIdx.y = the
// segment, blockIdx.x = the chunk): stable sort by (key on the segment's axis,
// position); writes the sorted keys and their positions.
__global__ void __launch_bounds__(kChunkThreads) k_top_sort(TopArgs A, TreeShape shape) {
    __shared__ TreeShape S;
    __shared__ double key[kChunk];
    __shared__ uint16_t perm[kChunk];
    load_shape(S, shape);
    __syncthreads();
    SegPos s{0, S.n, 0, 0, 0};
    seg_path(S, s, A.level, (int32_t)blockIdx.y);
    const int32_t c0 = s.a + (int32_t)blockIdx.x * kChunk;
    if (c0 >= s.b) return;  // uniform over the block
    const int32_t len = min(kChunk, s.b - c0);
    const int axis = A.axis[blockIdx.y];
    for (int32_t i = threadIdx.x; i < len; i += kChunkThreads) {
        key[i] = A.keys[3 * A.in[c0 + i] + axis];
        perm[i] = (uint16_t)i;
    }
    __syncthreads();
    constexpr int kPerChunk = kChunk / kChunkThreads;
    int32_t ea[kPerChunk], eb[kPerChunk];
    bool act[kPerChunk];
#pragma unroll
    for (int e = 0; e < kPerChunk; ++e) {
        ea[e] = 0;
        eb[e] = len;
        act[e] = (int32_t)threadIdx.x + e * kChunkThreads < len;
    }
    stable_merge<kPerChunk, kChunkThreads>(perm, ea, eb, act, len,
                                           [&](int, uint16_t y) -> double { return key[y]; });
    for (int32_t i = threadIdx.x; i < len; i += kChunkThreads) {
        const int32_t y = perm[i];
        A.skey[c0 + i] = key[y];
        A.spos[c0 + i] = c0 + y;
    }
}

// Rank of every position within its segment: its index in its sorted chunk
// plus, for each other chunk of the segment, the entries that precede it (a
// binary search: earlier chunks hold earlier positions, so their equal keys
// precede it; later chunks' do not).  Scatters the atom order.
__global__ void __launch_bounds__(256) k_top_merge(TopArgs A, TreeShape shape) {
    __shared__ TreeShape S;
    load_shape(S, shape);
    __syncthreads();
    const int32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= S.n) return;
    SegPos s{0, S.n, 0, 0, 0};
    int32_t k;
    seg_find(S, s, A.level, p, &k);
    const int32_t c = (p - s.a) / kChunk;
    const double key = A.skey[p];
    int32_t rank = p - (s.a + c * kChunk);
    const int32_t chunks = (s.b - s.a + kChunk - 1) / kChunk;
    for (int32_t o = 0; o < chunks; ++o) {
        if (o == c) continue;
        const int32_t lo0 = s.a + o * kChunk, hi0 = min(s.b, lo0 + kChunk);
        const bool strict = o > c;
        int32_t lo = lo0, hi = hi0;
        while (lo < hi) {
            const int32_t mid = (lo + hi) >> 1;
            const double km = A.skey[mid];
            if (strict ? (km < key) : (km <= key))
                lo = mid + 1;
            else
                hi = mid;
        }
        rank += lo - lo0;
    }
    A.out[s.a + rank] = A.in[A.spos[p]];
}

struct FinishArgs {
    int32_t top;              // depth of the finishing segments
    const double* c;          // n x 3
    const double* r;          // n
    const double* keys;       // n x 3
    const int32_t* order_in;  // position -> atom after the top levels
    int32_t* order;           // final position -> atom (the PA id slots)
    int32_t* groups;          // atom -> group
    int32_t* offsets;         // group -> first slot
    int32_t* eids;            // slot -> atom, argsort(groups, stable)
    NodePA* pa;
};

// One workgroup per segment at depth `top` (at most kFinishAtoms atoms): every
// remaining level of its subtree in LDS -- split axes, then a stable sort of
// each splitting segment by (key on its axis, position) -- then its PA nodes,
// groups, offsets and eids.
__global__ void __launch_bounds__(kFinishThreads) k_finish(FinishArgs A, TreeShape shape) {
    __shared__ TreeShape S;
    __shared__ double K[3][kFinishAtoms];
    __shared__ int32_t sid[kFinishAtoms];
    __shared__ uint16_t perm[kFinishAtoms];
    __shared__ uint8_t axis_of[kFinishAtoms];
    __shared__ double part[kWavesPerBlock][6];
    const int tid = threadIdx.x;
    load_shape(S, shape);
    __syncthreads();
    SegPos R{0, S.n, 0, 0, 0};
    seg_path(S, R, A.top, (int32_t)blockIdx.x);
    const int32_t m = R.b - R.a;
    for (int32_t i = tid; i < m; i += kFinishThreads) {
        const int32_t atom = A.order_in[R.a + i];
        sid[i] = atom;
#pragma unroll
        for (int c = 0; c < 3; ++c) K[c][i] = A.keys[3 * atom + c];
    }
    __syncthreads();
    const SegPos L0{0, m, R.depth, R.pre, R.group};  // the subtree root in local positions
    int levels = 0;  // local levels on which a segment may split
    while (R.depth + levels < kMaxSplitDepth && S.q[R.depth + levels] + 1 > S.leaf) ++levels;

    for (int j = 0; j < levels; ++j) {
        const int32_t big = S.q[R.depth + j] + 1;  // largest segment size at this depth
        const int32_t count = 1 << j;               // path bits of the segments at this depth
        // 1. split axes of the splitting segments
        if (count <= kWavesPerBlock) {
            // kWavesPerBlock / count waves per segment, partial extrema through LDS
            const int per = kWavesPerBlock / count;
            const int wave = tid / kWave, lane = tid & (kWave - 1);
            const int32_t k = wave / per, piece = wave % per;
            SegPos s = L0;
            const bool ok = seg_path(S, s, j, k) && seg_splits(S, s.b - s.a, s.depth);
            double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            if (ok) {
                for (int32_t i = s.a + piece * kWave + lane; i < s.b; i += per * kWave) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        lo[c] = pmin(lo[c], K[c][i]);
                        hi[c] = pmax(hi[c], K[c][i]);
                    }
                }
            }
            wave_minmax(lo, hi, kWave);
            if (lane == 0) {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    part[wave][c] = lo[c];
                    part[wave][3 + c] = hi[c];
                }
            }
            __syncthreads();
            if (tid < count) {
                SegPos t = L0;
                if (seg_path(S, t, j, tid) && seg_splits(S, t.b - t.a, t.depth)) {
                    double l2[3], h2[3];
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        l2[c] = part[tid * per][c];
                        h2[c] = part[tid * per][3 + c];
                    }
                    for (int w = 1; w < per; ++w) {
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            l2[c] = pmin(l2[c], part[tid * per + w][c]);
                            h2[c] = pmax(h2[c], part[tid * per + w][3 + c]);
                        }
                    }
                    axis_of[tid] = (uint8_t)split_axis(l2, h2);
                }
            }
        } else {
            // g lanes per segment, about eight atoms per lane
            int g = 1;
            while (g < kWave && 8 * g < big) g <<= 1;
            const int per_pass = kFinishThreads / g;
            const int lane = tid & (g - 1);
            for (int32_t base = 0; base < count; base += per_pass) {
                const int32_t k = base + tid / g;
                SegPos s = L0;
                const bool ok = k < count && seg_path(S, s, j, k) && seg_splits(S, s.b - s.a, s.depth);
                double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
                if (ok) {
                    for (int32_t i = s.a + lane; i < s.b; i += g) {
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            lo[c] = pmin(lo[c], K[c][i]);
                            hi[c] = pmax(hi[c], K[c][i]);
                        }
                    }
                }
                wave_minmax(lo, hi, g);
                if (ok && lane == 0) axis_of[k] = (uint8_t)split_axis(lo, hi);
            }
        }
        __syncthreads();

        // 2. stable sort of each splitting segment by (key on its axis, position)
        int32_t ea[kPer], eb[kPer];
        int ax[kPer];
        bool act[kPer];
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            const int32_t i = tid + e * kFinishThreads;
            ea[e] = eb[e] = 0;
            ax[e] = 0;
            act[e] = false;
            if (i < m) {
                SegPos s = L0;
                int32_t k;
                if (seg_find(S, s, j, i, &k) == j && seg_splits(S, s.b - s.a, s.depth)) {
                    ea[e] = s.a;
                    eb[e] = s.b;
                    ax[e] = axis_of[k];
                    act[e] = true;
                }
                perm[i] = (uint16_t)i;
            }
        }
        __syncthreads();
        auto key_of = [&](int e, uint16_t y) -> double { return K[ax[e]][y]; };
        if (big <= 32) {
            // direct ranks: at most 32 comparisons per atom
            int32_t np[kPer];
#pragma unroll
            for (int e = 0; e < kPer; ++e) {
                const int32_t i = tid + e * kFinishThreads;
                np[e] = -1;
                if (act[e]) {
                    const double key = key_of(e, (uint16_t)i);
                    int32_t cnt = 0;
                    for (int32_t y = ea[e]; y < eb[e]; ++y) {
                        const double ky = key_of(e, (uint16_t)y);
                        cnt += (ky < key || (ky == key && y < i)) ? 1 : 0;
                    }
                    np[e] = ea[e] + cnt;
                }
            }
#pragma unroll
            for (int e = 0; e < kPer; ++e)
                if (np[e] >= 0) perm[np[e]] = (uint16_t)(tid + e * kFinishThreads);
            __syncthreads();
        } else {
            stable_merge<kPer, kFinishThreads>(perm, ea, eb, act, big, key_of);
        }

        // 3. gather: position i now holds the atom at old position perm[i]
        int32_t nid[kPer];
        double nk[kPer][3];
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            const int32_t i = tid + e * kFinishThreads;
            if (i < m) {
                const int32_t y = perm[i];
                nid[e] = sid[y];
#pragma unroll
                for (int c = 0; c < 3; ++c) nk[e][c] = K[c][y];
            }
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            const int32_t i = tid + e * kFinishThreads;
            if (i < m) {
                sid[i] = nid[e];
#pragma unroll
                for (int c = 0; c < 3; ++c) K[c][i] = nk[e][c];
            }
        }
        __syncthreads();
    }

    for (int32_t i = tid; i < m; i += kFinishThreads) A.order[R.a + i] = sid[i];

    // 4. PA nodes (unpadded boxes, pa_spatial.py:38-75), groups, offsets and
    // eids, deepest level first so that children precede their parents
    for (int j = levels; j >= 0; --j) {
        const int32_t count = 1 << j;
        for (int32_t k = tid; k < count; k += kFinishThreads) {
            SegPos s = L0;
            if (!seg_path(S, s, j, k)) continue;
            const int32_t size = s.b - s.a;
            NodePA node;
            if (!seg_splits(S, size, s.depth)) {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    double low = INFINITY, high = -INFINITY;
                    for (int32_t y = s.a; y < s.b; ++y) {
                        const int32_t atom = sid[y];
                        const double v = A.c[3 * atom + c], rr = A.r[atom];
                        low = fmin(low, v - rr);
                        high = fmax(high, v + rr);
                    }
                    node.lo[c] = low;
                    node.hi[c] = high;
                }
                node.end = s.pre + 1;
                node.count = size;
                const int32_t g = s.group, first = R.a + s.a;
                A.offsets[g] = first;
                for (int32_t y = s.a; y < s.b; ++y) {
                    const int32_t atom = sid[y];
                    A.groups[atom] = g;
                    int32_t rank = 0;  // argsort(groups, stable): atoms of a group by index
                    for (int32_t z = s.a; z < s.b; ++z) rank += sid[z] < atom ? 1 : 0;
                    A.eids[first + rank] = atom;
                }
            } else {
                const NodePA& l = A.pa[s.pre + 1];
                const NodePA& rt = A.pa[s.pre + 2 * seg_leaves(S, size / 2, s.depth + 1)];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    node.lo[c] = fmin(l.lo[c], rt.lo[c]);
                    node.hi[c] = fmax(l.hi[c], rt.hi[c]);
                }
                node.end = s.pre + 2 * seg_leaves(S, size, s.depth) - 1;
                node.count = 0;
            }
            node.first = R.a + s.a;
            node.pad = 0;
            A.pa[s.pre] = node;
        }
        __threadfence();
        __syncthreads();
    }
}

// PA nodes of the top levels (single workgroup): unions of their children.
__global__ void __launch_bounds__(1024) k_top_nodes(FinishArgs A, TreeShape shape) {
    __shared__ TreeShape S;
    load_shape(S, shape);
    __syncthreads();
    for (int l = A.top - 1; l >= 0; --l) {
        for (int32_t k = threadIdx.x; k < (1 << l); k += blockDim.x) {
            SegPos s{0, S.n, 0, 0, 0};
            seg_path(S, s, l, k);
            const int32_t size = s.b - s.a;
            const NodePA& lt = A.pa[s.pre + 1];
            const NodePA& rt = A.pa[s.pre + 2 * seg_leaves(S, size / 2, s.depth + 1)];
            NodePA node;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                node.lo[c] = fmin(lt.lo[c], rt.lo[c]);
                node.hi[c] = fmax(lt.hi[c], rt.hi[c]);
            }
            node.end = s.pre + 2 * seg_leaves(S, size, s.depth) - 1;
            node.first = s.a;
            node.count = 0;
            node.pad = 0;
            A.pa[s.pre] = node;
        }
        __threadfence();
        __syncthreads();
    }
}

struct StaticArgs {
    int32_t ng;
    double pad;  // 1e-10 max(1, r_max) (ehss_static.py:86)
    const double* c;
    const double* r;
    const int32_t* offsets;
    const int32_t* eids;
    Node64* nodes;
};

// The static tree over the groups (single workgroup): its topology is implicit
// in ng (ehss_cpu_geometry.py:19-38: a node over groups [a, b) has its left
// child next and its right child 2 (mid - a) later); padded leaf boxes
// (ehss_cpu_geometry.py:41-55), then parent unions deepest first
// (ehss_cpu_geometry.py:58-71).
__global__ void __launch_bounds__(1024) k_static_tree(StaticArgs A) {
    for (int32_t g = threadIdx.x; g < A.ng; g += blockDim.x) {
        int32_t node = 0, a = 0, b = A.ng;
        while (b - a > 1) {
            const int32_t mid = a + (b - a) / 2;
            if (g < mid) {
                node += 1;
                b = mid;
            } else {
                node += 2 * (mid - a);
                a = mid;
            }
        }
        Node64 n;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double low = INFINITY, high = -INFINITY;
            for (int32_t k = A.offsets[g]; k < A.offsets[g + 1]; ++k) {
                const int32_t atom = A.eids[k];
                low = pmin(low, A.c[3 * atom + c] - A.r[atom]);
                high = pmax(high, A.c[3 * atom + c] + A.r[atom]);
            }
            n.lo[c] = low - A.pad;
            n.hi[c] = high + A.pad;
        }
        n.end = node + 1;
        n.leaf = g;
        n.pad[0] = n.pad[1] = 0;
        A.nodes[node] = n;
    }
    __threadfence();
    __syncthreads();
    int depth = 0;  // leaves sit at depth <= ceil(log2 ng)
    while ((1 << depth) < A.ng) ++depth;
    for (int d = depth - 1; d >= 0; --d) {
        for (int32_t k = threadIdx.x; k < (1 << d); k += blockDim.x) {
            int32_t node = 0, a = 0, b = A.ng;
            bool ok = true;
            for (int t = d - 1; t >= 0; --t) {
                if (b - a <= 1) {
                    ok = false;
                    break;
                }
                const int32_t mid = a + (b - a) / 2;
                if ((k >> t) & 1) {
                    node += 2 * (mid - a);
                    a = mid;
                } else {
                    node += 1;
                    b = mid;
                }
            }
            if (!ok || b - a <= 1) continue;
            const int32_t mid = a + (b - a) / 2;
            const Node64& l = A.nodes[node + 1];
            const Node64& rt = A.nodes[node + 2 * (mid - a)];
            Node64 n;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                n.lo[c] = (rt.lo[c] < l.lo[c]) ? rt.lo[c] : l.lo[c];  // min(lo[left], lo[right])
                n.hi[c] = (l.hi[c] < rt.hi[c]) ? rt.hi[c] : l.hi[c];  // max(hi[left], hi[right])
            }
            n.end = node + 2 * (b - a) - 1;
            n.leaf = -1;
            n.pad[0] = n.pad[1] = 0;
            A.nodes[node] = n;
        }
        __threadfence();
        __syncthreads();
    }
}

}  // namespace

int launch_put_frontier(const std::vector<int32_t>& f, int32_t* dst, hipStream_t s) {
    if (f.empty() || f.size() > (size_t)kFrontier)
        return set_error(EHSS_ERR_INVALID, "frontier of %zu entries", f.size());
    FrontierList L;
    L.n = (int32_t)f.size();
    std::copy(f.begin(), f.end(), L.v);
    hipLaunchKernelGGL(k_put_frontier, dim3(1), dim3(256), 0, s, L, dst);
    EHSS_HIP(hipGetLastError());
    return EHSS_OK;
}

int build_scene_shaped(const TreeShape& shape, const double* d_centers, const double* d_radii, double rmax,
                       double4* d_sph, int32_t* d_groups, int32_t* d_offsets, int32_t* d_eids, Node64* d_nodes,
                       NodePA* d_pa, int32_t* d_pa_ids, BuildWork& work, hipStream_t s) {
    const int32_t n = shape.n;
    EHSS_CHECK(n >= 1 && shape.leaf >= 1 && shape.leaf <= kShapedMaxLeaf, "shaped build of %d atoms, leaf size %d", n,
               shape.leaf);
    int top = 0;  // depth at which every segment has at most kFinishAtoms atoms
    while ((((int64_t)n + (1LL << top) - 1) >> top) > kFinishAtoms) ++top;
    EHSS_TRY(work.keys.reserve(sizeof(double) * 3 * n));
    if (top > 0) {
        EHSS_TRY(work.ids2.reserve(sizeof(int32_t) * n));
        EHSS_TRY(work.skey.reserve(sizeof(double) * n));
        EHSS_TRY(work.rank.reserve(sizeof(int32_t) * n));
        EHSS_TRY(work.seg.reserve((size_t)1 << top));
    }
    double* keys = work.keys.as<double>();
    hipLaunchKernelGGL(k_shape_init, dim3((n + 255) / 256), dim3(256), 0, s, n, shape.groups, d_centers, d_radii, keys,
                       d_sph, d_pa_ids, d_offsets);
    int32_t* in = d_pa_ids;
    int32_t* out = work.ids2.as<int32_t>();
    for (int l = 0; l < top; ++l) {
        TopArgs T{l, keys, in, out, work.seg.as<uint8_t>(), work.skey.as<double>(), work.rank.as<int32_t>()};
        hipLaunchKernelGGL(k_top_axis, dim3(1u << l), dim3(1024), 0, s, T, shape);
        const unsigned chunks = (unsigned)((shape.q[l] + 1 + kChunk - 1) / kChunk);
        hipLaunchKernelGGL(k_top_sort, dim3(chunks, 1u << l), dim3(kChunkThreads), 0, s, T, shape);
        hipLaunchKernelGGL(k_top_merge, dim3((n + 255) / 256), dim3(256), 0, s, T, shape);
        std::swap(in, out);
    }
    FinishArgs F{top, d_centers, d_radii, keys, in, d_pa_ids, d_groups, d_offsets, d_eids, d_pa};
    hipLaunchKernelGGL(k_finish, dim3(1u << top), dim3(kFinishThreads), 0, s, F, shape);
    if (top > 0) hipLaunchKernelGGL(k_top_nodes, dim3(1), dim3(1024), 0, s, F, shape);
    StaticArgs G{shape.groups, 1e-10 * (1. < rmax ? rmax : 1.), d_centers, d_radii, d_offsets, d_eids, d_nodes};
    hipLaunchKernelGGL(k_static_tree, dim3(1), dim3(1024), 0, s, G);
    EHSS_HIP(hipGetLastError());
    return EHSS_OK;
}

}  // namespace ehss

extern "C" int ehss_tree_counts(int64_t n, int leaf_size, int64_t* out) {
    EHSS_CHECK(n >= 1 && n < (1LL << 28), "atom count %lld out of range", (long long)n);
    EHSS_CHECK(leaf_size >= 1 && out, "Positive integer leaf size required");
    const ehss::TreeShape sh = ehss::make_tree_shape((int32_t)n, leaf_size);
    out[0] = sh.groups;
    out[1] = 2 * (int64_t)sh.groups - 1;
    out[2] = 2 * (int64_t)sh.groups - 1;
    out[3] = sh.max_depth;
    return EHSS_OK;
}
