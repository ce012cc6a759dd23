// build.hip -- on-device construction of the static EHSS tree and the PA tree.
//
// Restates, exactly:
//   * _median_groups (ehss_static.py:16-50): median split along the axis of the
//     largest key span (first on ties), keys ((c - r) + (c + r)) / 2
//     (ehss_static.py:57-59), stable sort, mid = a + (b - a) / 2, leaf when
//     b - a <= leaf_size, depth >= 24 or zero span; groups numbered in
//     depth-first (left first) leaf order;
//   * offsets / ids = argsort(groups, stable) (ehss_static.py:82-83);
//   * topology over groups (ehss_cpu_geometry.py:19-38), padded leaf boxes
//     (ehss_cpu_geometry.py:41-55, pad = 1e-10 max(1, r_max), ehss_static.py:86)
//     and parent unions (ehss_cpu_geometry.py:58-71);
//   * the PA tree of pa_spatial.py:38-75 (build_tree, kind 'bvh'), which is the
//     same recursion: every segment of the median split is a PA node, numbered in
//     preorder, with unpadded boxes, first = segment start, count = leaf size.
//
// GPU formulation: the recursion is run level by level.  At each level every
// live segment computes its key span (one workgroup per segment), and every
// element of a splitting segment finds its new position as its rank under the
// order (key, current position) -- a stable sort expressed as a parallel rank
// count (2-D grid over element blocks x chunks of the segment, integer atomics,
// so the result is exact and independent of scheduling).  Topology and boxes
// are then derived in parallel; min/max unions are exact, so their order does
// not matter.
#include <stdint.h>

#include "ehss_internal.h"

namespace ehss {

namespace {

constexpr int kMaxLevels = 26;
constexpr int kRankChunk = 1024;

struct Ctl {
    int32_t seg_count;
    int32_t level_begin[kMaxLevels + 1];
    int32_t level_end[kMaxLevels + 1];
    int32_t n_groups;
    int32_t max_depth;
    int32_t pad;
    unsigned long long rmax_bits;
};

struct BuildArgs {
    int32_t n, leaf_size;
    const double* c;  // n x 3
    const double* r;  // n
    double* keys;     // n x 3
    double4* sph;     // n
    int32_t* ids;     // position -> atom (current)
    int32_t* ids2;
    int32_t* pseg;    // position -> segment (current)
    int32_t* pseg2;
    double* skey;     // position -> key on the split axis of its segment
    int32_t* rank;
    int32_t *sa, *sb, *sdepth, *sleaf, *saxis, *sleft, *sright, *ssize, *spre;  // segments
    Ctl* ctl;
    // outputs
    int32_t* groups;    // atom -> group
    int32_t* flag;      // position -> 1 at a leaf start
    int32_t* offsets;   // group -> first slot (n_groups + 1)
    int32_t* eids;      // slot -> atom, argsort(groups, stable)
    Node64* nodes;      // static tree, 2 g - 1
    int32_t* ndepth;    // node -> depth
    int32_t* nright;    // node -> right child (or -1)
    int32_t* leaf_node; // group -> node
    NodePA* pa;         // PA tree
    NodePA* pabox;      // segment -> box scratch
};

__device__ __forceinline__ double pmin(double low, double v) { return (v < low) ? v : low; }   // Python min(low, v)
__device__ __forceinline__ double pmax(double high, double v) { return (high < v) ? v : high; } // Python max(high, v)

__global__ void k_init(BuildArgs A) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < A.n) {
        double c[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            c[k] = A.c[3 * i + k];
            const double lo = c[k] - A.r[i], hi = c[k] + A.r[i];
            A.keys[3 * i + k] = (lo + hi) / 2;
        }
        A.sph[i] = make_double4(c[0], c[1], c[2], A.r[i]);
        A.ids[i] = i;
        A.pseg[i] = 0;
        atomicMax(&A.ctl->rmax_bits, (unsigned long long)__double_as_longlong(A.r[i]));
    }
    if (i == 0) {
        A.sa[0] = 0;
        A.sb[0] = A.n;
        A.sdepth[0] = 0;
        A.sleaf[0] = 0;
        A.ctl->seg_count = 1;
        A.ctl->level_begin[0] = 0;
        A.ctl->level_end[0] = 1;
        A.ctl->max_depth = 0;
    }
}

// One workgroup per segment of the level: key span, axis, leaf decision, children.
__global__ void __launch_bounds__(256) k_span(BuildArgs A, int level) {
    __shared__ double lo_s[3][256];
    __shared__ double hi_s[3][256];
    const int32_t begin = A.ctl->level_begin[level], end = A.ctl->level_end[level];
    for (int32_t s = begin + blockIdx.x; s < end; s += gridDim.x) {
        const int32_t a = A.sa[s], b = A.sb[s];
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int32_t k = a + threadIdx.x; k < b; k += blockDim.x) {
            const int32_t atom = A.ids[k];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double v = A.keys[3 * atom + c];
                lo[c] = pmin(lo[c], v);
                hi[c] = pmax(hi[c], v);
            }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            lo_s[c][threadIdx.x] = lo[c];
            hi_s[c][threadIdx.x] = hi[c];
        }
        __syncthreads();
        for (int w = blockDim.x / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    lo_s[c][threadIdx.x] = pmin(lo_s[c][threadIdx.x], lo_s[c][threadIdx.x + w]);
                    hi_s[c][threadIdx.x] = pmax(hi_s[c][threadIdx.x], hi_s[c][threadIdx.x + w]);
                }
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            double span[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) span[c] = hi_s[c][0] - lo_s[c][0];
            int axis = 0;  // np.argmax: first maximum
            if (span[1] > span[axis]) axis = 1;
            if (span[2] > span[axis]) axis = 2;
            const int32_t depth = A.sdepth[s];
            atomicMax(&A.ctl->max_depth, depth);
            if (b - a <= A.leaf_size || depth >= 24 || span[axis] == 0) {
                A.sleaf[s] = 1;
                A.sleft[s] = A.sright[s] = -1;
            } else {
                A.sleaf[s] = 0;
                A.saxis[s] = axis;
                const int32_t c0 = atomicAdd(&A.ctl->seg_count, 2);
                const int32_t mid = a + (b - a) / 2;
                A.sleft[s] = c0;
                A.sright[s] = c0 + 1;
                A.sa[c0] = a;
                A.sb[c0] = mid;
                A.sa[c0 + 1] = mid;
                A.sb[c0 + 1] = b;
                A.sdepth[c0] = A.sdepth[c0 + 1] = depth + 1;
                A.sleaf[c0] = A.sleaf[c0 + 1] = 0;
            }
        }
        __syncthreads();
    }
}

__global__ void k_level_bounds(BuildArgs A, int level) {
    A.ctl->level_begin[level + 1] = A.ctl->level_end[level];
    A.ctl->level_end[level + 1] = A.ctl->seg_count;
}

__device__ __forceinline__ bool splitting(const BuildArgs& A, int32_t s, int level) {
    return A.sdepth[s] == level && !A.sleaf[s];
}

__global__ void k_skey(BuildArgs A, int level) {
    const int32_t pos = blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= A.n) return;
    const int32_t s = A.pseg[pos];
    if (splitting(A, s, level)) {
        A.skey[pos] = A.keys[3 * A.ids[pos] + A.saxis[s]];
        A.rank[pos] = 0;
    }
}

// rank[pos] += #{ j in chunk : (skey[j], j) < (skey[pos], pos) } within the segment.
__global__ void __launch_bounds__(256) k_rank(BuildArgs A, int level) {
    const int32_t pos = blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= A.n) return;
    const int32_t s = A.pseg[pos];
    if (!splitting(A, s, level)) return;
    const int32_t a = A.sa[s], b = A.sb[s];
    const int32_t j0 = a + (int32_t)blockIdx.y * kRankChunk;
    if (j0 >= b) return;
    const int32_t j1 = min(b, j0 + kRankChunk);
    const double key = A.skey[pos];
    int32_t cnt = 0;
    for (int32_t j = j0; j < j1; ++j) {
        const double kj = A.skey[j];
        cnt += (kj < key || (kj == key && j < pos)) ? 1 : 0;
    }
    if (cnt) atomicAdd(&A.rank[pos], cnt);
}

__global__ void k_scatter(BuildArgs A, int level) {
    const int32_t pos = blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= A.n) return;
    const int32_t s = A.pseg[pos];
    if (splitting(A, s, level)) {
        const int32_t a = A.sa[s], b = A.sb[s];
        const int32_t np = a + A.rank[pos];
        const int32_t mid = a + (b - a) / 2;
        A.ids2[np] = A.ids[pos];
        A.pseg2[np] = np < mid ? A.sleft[s] : A.sright[s];
    } else {
        A.ids2[pos] = A.ids[pos];
        A.pseg2[pos] = s;
    }
}

__global__ void k_leaf_flags(BuildArgs A) {
    const int32_t pos = blockIdx.x * blockDim.x + threadIdx.x;
    if (pos < A.n) A.flag[pos] = (A.sa[A.pseg[pos]] == pos) ? 1 : 0;
}

// Single workgroup: inclusive scan of leaf-start flags -> groups and offsets.
__global__ void __launch_bounds__(1024) k_groups(BuildArgs A) {
    __shared__ int32_t part[1024];
    const int32_t n = A.n;
    const int32_t per = (n + blockDim.x - 1) / blockDim.x;
    const int32_t lo = threadIdx.x * per, hi = min(n, lo + per);
    int32_t sum = 0;
    for (int32_t p = lo; p < hi; ++p) sum += A.flag[p];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < (int)blockDim.x; off <<= 1) {
        const int32_t v = (int)threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int32_t g = (threadIdx.x ? part[threadIdx.x - 1] : 0) - 1;
    for (int32_t p = lo; p < hi; ++p) {
        if (A.flag[p]) {
            ++g;
            A.offsets[g] = p;
        }
        A.groups[A.ids[p]] = g;
    }
    if (threadIdx.x == blockDim.x - 1) {
        A.ctl->n_groups = part[threadIdx.x];
        A.offsets[part[threadIdx.x]] = n;
    }
}

// eids = argsort(groups, stable): inside each group, atoms in increasing index.
__global__ void k_group_ids(BuildArgs A) {
    const int32_t pos = blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= A.n) return;
    const int32_t s = A.pseg[pos];
    const int32_t a = A.sa[s], b = A.sb[s];
    const int32_t me = A.ids[pos];
    int32_t r = 0;
    for (int32_t k = a; k < b; ++k) r += A.ids[k] < me ? 1 : 0;
    A.eids[a + r] = me;
}

// Implicit topology over the groups (ehss_cpu_geometry.py:19-38): walk from the root.
__global__ void k_topology(BuildArgs A) {
    const int32_t ng = A.ctl->n_groups;
    const int32_t target = blockIdx.x * blockDim.x + threadIdx.x;
    if (target >= 2 * ng - 1) return;
    int32_t node = 0, a = 0, b = ng, depth = 0;
    while (node != target) {
        const int32_t mid = a + (b - a) / 2;
        const int32_t right = node + 2 * (mid - a);
        if (target < right) {
            node = node + 1;
            b = mid;
        } else {
            node = right;
            a = mid;
        }
        ++depth;
    }
    Node64& n = A.nodes[target];
    n.end = target + 2 * (b - a) - 1;
    n.pad[0] = n.pad[1] = 0;
    A.ndepth[target] = depth;
    if (b - a == 1) {
        n.leaf = a;
        A.leaf_node[a] = target;
        A.nright[target] = -1;
    } else {
        n.leaf = -1;
        const int32_t mid = a + (b - a) / 2;
        A.nright[target] = target + 2 * (mid - a);
    }
}

// Padded leaf boxes (ehss_cpu_geometry.py:41-55).
__global__ void k_leaf_boxes(BuildArgs A) {
    const int32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= A.ctl->n_groups) return;
    const double rmax = __longlong_as_double((long long)A.ctl->rmax_bits);
    const double pad = 1e-10 * (1. < rmax ? rmax : 1.);  // 1e-10 * max(1., r_max)
    Node64& n = A.nodes[A.leaf_node[g]];
    for (int c = 0; c < 3; ++c) {
        double low = INFINITY, high = -INFINITY;
        for (int32_t k = A.offsets[g]; k < A.offsets[g + 1]; ++k) {
            const int32_t atom = A.eids[k];
            low = pmin(low, A.c[3 * atom + c] - A.r[atom]);
            high = pmax(high, A.c[3 * atom + c] + A.r[atom]);
        }
        n.lo[c] = low - pad;
        n.hi[c] = high + pad;
    }
}

// Parent unions, deepest level first (single workgroup; exact min/max).
__global__ void __launch_bounds__(1024) k_parents(BuildArgs A, int max_depth) {
    const int32_t nn = 2 * A.ctl->n_groups - 1;
    for (int d = max_depth; d >= 0; --d) {
        for (int32_t node = threadIdx.x; node < nn; node += blockDim.x) {
            if (A.ndepth[node] != d || A.nright[node] < 0) continue;
            Node64& n = A.nodes[node];
            const Node64& l = A.nodes[node + 1];
            const Node64& r = A.nodes[A.nright[node]];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                n.lo[c] = (r.lo[c] < l.lo[c]) ? r.lo[c] : l.lo[c];  // min(lo[left], lo[right])
                n.hi[c] = (l.hi[c] < r.hi[c]) ? r.hi[c] : l.hi[c];  // max(hi[left], hi[right])
            }
        }
        __threadfence();
        __syncthreads();
    }
}

// PA tree from the segment tree (single workgroup): subtree sizes bottom-up,
// preorder numbers top-down, then nodes with unpadded boxes.
__global__ void __launch_bounds__(1024) k_pa_tree(BuildArgs A, int levels) {
    NodePA* tmp = A.pabox;  // boxes by segment index, moved to preorder slots below
    for (int L = levels - 1; L >= 0; --L) {
        const int32_t b0 = A.ctl->level_begin[L], b1 = A.ctl->level_end[L];
        for (int32_t s = b0 + threadIdx.x; s < b1; s += blockDim.x) {
            NodePA& t = tmp[s];
            if (A.sleaf[s]) {
                A.ssize[s] = 1;
                for (int c = 0; c < 3; ++c) {
                    double low = INFINITY, high = -INFINITY;
                    for (int32_t k = A.sa[s]; k < A.sb[s]; ++k) {
                        const int32_t atom = A.ids[k];
                        const double v = A.c[3 * atom + c];
                        low = fmin(low, v - A.r[atom]);
                        high = fmax(high, v + A.r[atom]);
                    }
                    t.lo[c] = low;
                    t.hi[c] = high;
                }
            } else {
                const int32_t l = A.sleft[s], r = A.sright[s];
                A.ssize[s] = 1 + A.ssize[l] + A.ssize[r];
                for (int c = 0; c < 3; ++c) {
                    t.lo[c] = fmin(tmp[l].lo[c], tmp[r].lo[c]);
                    t.hi[c] = fmax(tmp[l].hi[c], tmp[r].hi[c]);
                }
            }
        }
        __threadfence();
        __syncthreads();
    }
    if (threadIdx.x == 0) A.spre[0] = 0;
    __threadfence();
    __syncthreads();
    for (int L = 0; L < levels; ++L) {
        const int32_t b0 = A.ctl->level_begin[L], b1 = A.ctl->level_end[L];
        for (int32_t s = b0 + threadIdx.x; s < b1; s += blockDim.x) {
            const int32_t pre = A.spre[s];
            if (!A.sleaf[s]) {
                A.spre[A.sleft[s]] = pre + 1;
                A.spre[A.sright[s]] = pre + 1 + A.ssize[A.sleft[s]];
            }
            NodePA& n = A.pa[pre];
            const NodePA& t = tmp[s];
            for (int c = 0; c < 3; ++c) {
                n.lo[c] = t.lo[c];
                n.hi[c] = t.hi[c];
            }
            n.end = pre + A.ssize[s];
            n.first = A.sa[s];
            n.count = A.sleaf[s] ? A.sb[s] - A.sa[s] : 0;
            n.pad = 0;
        }
        __threadfence();
        __syncthreads();
    }
}

// Root segment temporarily marked as splitting [0, n) for the custom-group rank.
__global__ void k_reset_root(BuildArgs A) {
    A.ssize[0] = A.sleaf[0];  // remember the root's leaf flag in ssize[0]
    A.sleaf[0] = 0;
    A.sdepth[0] = 0;
}

__global__ void k_restore_root(BuildArgs A) { A.sleaf[0] = A.ssize[0]; }

// Caller-supplied groups: eids = argsort(groups, stable) as the rank of
// (group, atom) over all atoms, reusing k_rank on a single segment.
__global__ void k_custom_keys(BuildArgs A, const int32_t* cg) {
    const int32_t pos = blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= A.n) return;
    A.skey[pos] = (double)cg[pos];
    A.rank[pos] = 0;
    A.pseg[pos] = 0;
    A.groups[pos] = cg[pos];
}

__global__ void k_custom_place(BuildArgs A, const int32_t* cg) {
    const int32_t pos = blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= A.n) return;
    A.eids[A.rank[pos]] = pos;
}

__global__ void k_custom_offsets(BuildArgs A, const int32_t* cg, int32_t ng) {
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= A.n) return;
    const int32_t g = cg[A.eids[k]];
    if (k == 0 || cg[A.eids[k - 1]] != g) A.offsets[g] = k;
    if (k == 0) {
        A.offsets[ng] = A.n;
        A.ctl->n_groups = ng;
    }
}

}  // namespace

int build_scene_device(int32_t n, int32_t leaf_size, const double* d_centers, const double* d_radii, double4* d_sph,
                       int32_t* d_groups, int32_t* d_offsets, int32_t* d_eids, Node64* d_nodes, NodePA* d_pa,
                       int32_t* d_pa_ids, const int32_t* d_custom_groups, int32_t custom_ng, BuiltScene* out,
                       BuildWork& work, hipStream_t s) {
    // Work arrays (grow-only, kept by the caller).
    DevBuf &keys = work.keys, &ids2 = work.ids2, &pseg = work.pseg, &pseg2 = work.pseg2, &skey = work.skey,
           &rank = work.rank, &seg = work.seg, &ctl = work.ctl, &flag = work.flag, &ndepth = work.ndepth,
           &nright = work.nright, &leaf_node = work.leaf_node, &pabox = work.pabox;
    const int32_t max_segs = 2 * n + 2;
    EHSS_TRY(keys.reserve(sizeof(double) * 3 * n));
    EHSS_TRY(ids2.reserve(sizeof(int32_t) * n));
    EHSS_TRY(pseg.reserve(sizeof(int32_t) * n));
    EHSS_TRY(pseg2.reserve(sizeof(int32_t) * n));
    EHSS_TRY(skey.reserve(sizeof(double) * n));
    EHSS_TRY(rank.reserve(sizeof(int32_t) * n));
    EHSS_TRY(seg.reserve(sizeof(int32_t) * 9 * (size_t)max_segs));
    EHSS_TRY(ctl.reserve(sizeof(Ctl)));
    EHSS_TRY(flag.reserve(sizeof(int32_t) * n));
    EHSS_TRY(ndepth.reserve(sizeof(int32_t) * max_segs));
    EHSS_TRY(nright.reserve(sizeof(int32_t) * max_segs));
    EHSS_TRY(leaf_node.reserve(sizeof(int32_t) * (n + 1)));
    EHSS_TRY(pabox.reserve(sizeof(NodePA) * max_segs));
    EHSS_HIP(hipMemsetAsync(ctl.p, 0, sizeof(Ctl), s));

    BuildArgs A;
    A.n = n;
    A.leaf_size = leaf_size;
    A.c = d_centers;
    A.r = d_radii;
    A.keys = keys.as<double>();
    A.sph = d_sph;
    A.ids = d_pa_ids;  // the final segment order is the PA id order
    A.ids2 = ids2.as<int32_t>();
    A.pseg = pseg.as<int32_t>();
    A.pseg2 = pseg2.as<int32_t>();
    A.skey = skey.as<double>();
    A.rank = rank.as<int32_t>();
    int32_t* sg = seg.as<int32_t>();
    A.sa = sg;
    A.sb = sg + max_segs;
    A.sdepth = sg + 2 * max_segs;
    A.sleaf = sg + 3 * max_segs;
    A.saxis = sg + 4 * max_segs;
    A.sleft = sg + 5 * max_segs;
    A.sright = sg + 6 * max_segs;
    A.ssize = sg + 7 * max_segs;
    A.spre = sg + 8 * max_segs;
    A.ctl = ctl.as<Ctl>();
    A.groups = d_groups;
    A.flag = flag.as<int32_t>();
    A.offsets = d_offsets;
    A.eids = d_eids;
    A.nodes = d_nodes;
    A.ndepth = ndepth.as<int32_t>();
    A.nright = nright.as<int32_t>();
    A.leaf_node = leaf_node.as<int32_t>();
    A.pa = d_pa;
    A.pabox = pabox.as<NodePA>();

    const dim3 eb(256), eg((n + 255) / 256);
    hipLaunchKernelGGL(k_init, eg, eb, 0, s, A);
    EHSS_HIP(hipGetLastError());
    // Median splits halve segment sizes, so the tree is at most ceil(log2 n) + 1 deep.
    int levels = 1;
    while ((1LL << (levels - 1)) < (long long)n) ++levels;
    levels = std::min(levels + 1, 25);
    for (int L = 0; L < levels; ++L) {
        const int64_t max_size = ((int64_t)n + (1LL << L) - 1) >> L;  // ceil(n / 2^L)
        const unsigned span_blocks = (unsigned)std::min<int64_t>(std::max<int64_t>(1, (1LL << L)), 65535);
        hipLaunchKernelGGL(k_span, dim3(span_blocks), dim3(256), 0, s, A, L);
        hipLaunchKernelGGL(k_level_bounds, dim3(1), dim3(1), 0, s, A, L);
        if (L + 1 == levels) break;  // segments at the last level cannot split (size 1 or depth 24)
        hipLaunchKernelGGL(k_skey, eg, eb, 0, s, A, L);
        const unsigned chunks = (unsigned)((max_size + kRankChunk - 1) / kRankChunk);
        hipLaunchKernelGGL(k_rank, dim3(eg.x, chunks), eb, 0, s, A, L);
        hipLaunchKernelGGL(k_scatter, eg, eb, 0, s, A, L);
        EHSS_HIP(hipGetLastError());
        std::swap(A.ids, A.ids2);
        std::swap(A.pseg, A.pseg2);
    }
    // The final order must live in d_pa_ids.
    if (A.ids != d_pa_ids) {
        EHSS_HIP(hipMemcpyAsync(d_pa_ids, A.ids, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s));
        A.ids2 = A.ids;
        A.ids = d_pa_ids;
    }
    if (d_custom_groups) {
        // The PA tree above is the median-split tree; the EHSS groups are the caller's.
        // Segment 0 (the root) is reused as one segment holding every atom.
        BuildArgs C = A;
        C.pseg = A.pseg2;
        hipLaunchKernelGGL(k_reset_root, dim3(1), dim3(1), 0, s, C);
        hipLaunchKernelGGL(k_custom_keys, eg, eb, 0, s, C, d_custom_groups);
        const unsigned chunks = (unsigned)((n + kRankChunk - 1) / kRankChunk);
        hipLaunchKernelGGL(k_rank, dim3(eg.x, chunks), eb, 0, s, C, 0);
        hipLaunchKernelGGL(k_custom_place, eg, eb, 0, s, C, d_custom_groups);
        hipLaunchKernelGGL(k_custom_offsets, eg, eb, 0, s, C, d_custom_groups, custom_ng);
        hipLaunchKernelGGL(k_restore_root, dim3(1), dim3(1), 0, s, C);
    } else {
        hipLaunchKernelGGL(k_leaf_flags, eg, eb, 0, s, A);
        hipLaunchKernelGGL(k_groups, dim3(1), dim3(1024), 0, s, A);
        hipLaunchKernelGGL(k_group_ids, eg, eb, 0, s, A);
    }
    EHSS_HIP(hipGetLastError());
    Ctl hc;
    EHSS_HIP(hipMemcpyAsync(&hc, ctl.p, sizeof(Ctl), hipMemcpyDeviceToHost, s));
    EHSS_HIP(hipStreamSynchronize(s));
    const int32_t ng = hc.n_groups;
    if (ng < 1 || ng > n) return set_error(EHSS_ERR_HIP, "tree build produced %d groups for %d atoms", ng, n);
    const int32_t nn = 2 * ng - 1;
    hipLaunchKernelGGL(k_topology, dim3((nn + 255) / 256), eb, 0, s, A);
    hipLaunchKernelGGL(k_leaf_boxes, dim3((ng + 255) / 256), eb, 0, s, A);
    int tree_depth = 0;
    while ((1LL << tree_depth) < (long long)ng) ++tree_depth;
    hipLaunchKernelGGL(k_parents, dim3(1), dim3(1024), 0, s, A, tree_depth);
    hipLaunchKernelGGL(k_pa_tree, dim3(1), dim3(1024), 0, s, A, levels);
    EHSS_HIP(hipGetLastError());
    EHSS_HIP(hipStreamSynchronize(s));
    out->n_groups = ng;
    out->n_nodes = nn;
    out->pa_nodes = hc.seg_count;
    out->max_depth = hc.max_depth;
    return EHSS_OK;
}

}  // namespace ehss
