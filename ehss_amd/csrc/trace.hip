// trace.hip -- per-ray trace kernels over a device scene (one ray per lane).
//
//   kPlain (EHSS_PRECISION_FP64): the reference loop of src/native/ehss_cpu.cpp:54-79
//     with FP64 IEEE arithmetic, the same stackless preorder traversal and the
//     same box test; results and box/sphere counters are bit-identical to the
//     reference on the same rays.
//   kFast (EHSS_PRECISION_FP32_FIXUP): the same nearest-hit query answered with an
//     FP32 filter (outward-inflated FP32 boxes with a precomputed reciprocal
//     direction, FP32 sphere classification) and an exact FP64 re-test of every
//     sphere the filter cannot exclude.  The filter only ever excludes spheres
//     whose FP64 entry distance is provably infinite or larger than the running
//     FP64 best, so the selected atom, the distance and every reflection are the
//     FP64 ones: outputs are bit-identical to kPlain.
#include "ehss_query.h"

namespace ehss {

namespace {

template <int kFast>
__global__ void __launch_bounds__(256) k_trace(SceneView S, TraceIO io) {
    const int64_t ray = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t boxes = 0, spheres = 0, retests = 0;
    if (ray < io.nr) {
        V3 o = {io.p[3 * ray], io.p[3 * ray + 1], io.p[3 * ray + 2]};
        V3 d = {io.d[3 * ray], io.d[3 * ray + 1], io.d[3 * ray + 2]};
        int32_t last = (int32_t)io.lasts[ray];
        int64_t count = io.counts[ray];
        const int64_t cap = io.cap;
        if (io.record && count == 1) io.history[ray * cap] = last;
        uint8_t esc = 0, unres = 0;
        for (;;) {
            Query q;
            // Outputs do not depend on the grazing tolerance; 0 keeps the filter tightest.
            if (kFast)
                nearest_fast(S, o, d, last, 0., q, boxes, spheres, retests);
            else
                nearest_plain(S, o, d, last, 0., q, boxes, spheres);
            if (q.atom < 0) {
                esc = 1;
                break;
            }
            if (count >= cap) {
                unres = 1;
                break;
            }
            const double4 c = S.sph[q.atom];
            reflect64(o, d, c.x, c.y, c.z, c.w, q.best);
            if (io.record) io.history[ray * cap + count] = q.atom;
            ++count;
            last = q.atom;
        }
        io.p[3 * ray] = o.x;
        io.p[3 * ray + 1] = o.y;
        io.p[3 * ray + 2] = o.z;
        io.d[3 * ray] = d.x;
        io.d[3 * ray + 1] = d.y;
        io.d[3 * ray + 2] = d.z;
        io.counts[ray] = count;
        io.escaped[ray] = esc;
        io.unresolved[ray] = unres;
    }
    // Integer counters: wave reduction then one atomic per wave (order-free).
    const uint64_t b = (uint64_t)wave_sum_i64((int64_t)boxes);
    const uint64_t sp = (uint64_t)wave_sum_i64((int64_t)spheres);
    const uint64_t rt = (uint64_t)wave_sum_i64((int64_t)retests);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&io.counters[0], (unsigned long long)b);
        atomicAdd(&io.counters[1], (unsigned long long)sp);
        if (kFast) atomicAdd(&io.counters[2], (unsigned long long)rt);
    }
}

// pa_spatial.py:114-142 (_one over the PA tree from root 0, FP64).
__global__ void __launch_bounds__(256) k_pa_query(SceneView S, PAIO io) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t boxes = 0, tests = 0;
    if (i < io.nr) {
        const V3 o = {io.o[3 * i], io.o[3 * i + 1], io.o[3 * i + 2]};
        const V3 d = {io.d[3 * i], io.d[3 * i + 1], io.d[3 * i + 2]};
        const int32_t owner = io.owners ? (int32_t)io.owners[i] : -1;
        bool graze;
        io.hits[i] = pa_plain(S, o, d, owner, 0., boxes, tests, graze) ? 1 : 0;
    }
    const uint64_t b = (uint64_t)wave_sum_i64((int64_t)boxes);
    const uint64_t t = (uint64_t)wave_sum_i64((int64_t)tests);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&io.counters[0], (unsigned long long)b);
        atomicAdd(&io.counters[1], (unsigned long long)t);
    }
}

// FP32 filter arrays: boxes rounded outward and inflated by m_box, internal
// nodes linked to their right child; spheres in slot order, recentred on the
// scene centre.
__global__ void k_make_filter(SceneView S, Node32* nodes32, float4* sph32, double4* sphs, Node32* leaves32) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < S.n_nodes) {
        const Node64& n = S.nodes[i];
        const double c[3] = {S.cx, S.cy, S.cz};
        Node32 m;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            m.lo[k] = __double2float_rd(n.lo[k] - c[k]) - S.m_box;
            m.hi[k] = __double2float_ru(n.hi[k] - c[k]) + S.m_box;
        }
        m.end = n.end;
        // A caller's tree (ehss_trace) may end in an internal node with no
        // successor; it keeps -1 (only scene trees are walked by the packet
        // queries that follow the link).
        m.leaf = n.leaf >= 0 ? n.leaf : (i + 1 < S.n_nodes ? right_child(S.nodes[i + 1].end) : -1);
        nodes32[i] = m;
        if (leaves32 && n.leaf >= 0) {
            m.end = S.offsets[n.leaf];
            m.leaf = S.offsets[n.leaf + 1] - m.end;
            leaves32[n.leaf] = m;
        }
    }
    if (i < S.n_atoms) {
        const double4 s = S.sph[S.ids[i]];
        sph32[i] = make_float4((float)(s.x - S.cx), (float)(s.y - S.cy), (float)(s.z - S.cz), (float)s.w);
        sphs[i] = s;
    }
}

// Packet tree: thread i takes static node i.  The walk from the root gives the
// node's group range and, in the cut tree, its preorder index (a left step adds
// 1, a right step 1 + the cut nodes of the left sibling) and the index of its
// first leaf.  Nodes below a subtree of at most `cut` groups are dropped; the
// rest keep their FP32 filter box with links renumbered for the cut tree.
__global__ void k_packet_tree(SceneView S, int32_t cut, Node32* pk, int32_t* pk_off) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S.n_nodes) return;
    int32_t node = 0, a = 0, b = S.n_groups, at = 0, leaf = 0;
    while (node != i) {
        const int32_t k = b - a;
        if (k <= cut) return;  // inside a subtree that became one leaf
        const int32_t h = k / 2, right = node + 2 * h;
        if (i < right) {
            node += 1;
            b = a + h;
            at += 1;
        } else {
            node = right;
            at += 1 + cut_nodes(h, cut);
            leaf += cut_leaves(h, cut);
            a += h;
        }
    }
    const int32_t k = b - a;
    Node32 m = S.nodes32[i];
    if (k <= cut) {
        m.end = at + 1;
        m.leaf = leaf;
        pk_off[leaf] = S.offsets[a];
        if (b == S.n_groups) pk_off[leaf + 1] = S.offsets[b];
    } else {
        m.end = at + cut_nodes(k, cut);
        m.leaf = right_child(at + 1 + cut_nodes(k / 2, cut));
    }
    pk[at] = m;
}

}  // namespace

int launch_packet_tree(const SceneView& v, int32_t cut, Node32* pk_nodes32, int32_t* pk_offsets, hipStream_t s) {
    if (v.n_nodes == 0) return EHSS_OK;
    hipLaunchKernelGGL(k_packet_tree, dim3((v.n_nodes + 255) / 256), dim3(256), 0, s, v, cut, pk_nodes32, pk_offsets);
    EHSS_HIP(hipGetLastError());
    return EHSS_OK;
}

void set_filter_margins(SceneView& v) {
    const double u = 5.9604644775390625e-08;  // 2^-24
    v.m_box = (float)(64.0 * u * v.scale);
    v.m_p = (float)(64.0 * u * v.scale);
    v.m_disc = (float)(1024.0 * u * v.scale * v.scale);
}

int launch_make_filter(const SceneView& v, Node32* nodes32, float4* sph32, double4* sphs, Node32* leaves32,
                       hipStream_t s) {
    const int32_t n = v.n_nodes > v.n_atoms ? v.n_nodes : v.n_atoms;
    if (n == 0) return EHSS_OK;
    hipLaunchKernelGGL(k_make_filter, dim3((n + 255) / 256), dim3(256), 0, s, v, nodes32, sph32, sphs, leaves32);
    EHSS_HIP(hipGetLastError());
    return EHSS_OK;
}

int launch_trace(const SceneView& v, const TraceIO& io, int precision, hipStream_t s) {
    if (io.nr == 0) return EHSS_OK;
    const dim3 grid((unsigned)((io.nr + 255) / 256));
    if (precision == EHSS_PRECISION_FP64)
        hipLaunchKernelGGL(k_trace<0>, grid, dim3(256), 0, s, v, io);
    else
        hipLaunchKernelGGL(k_trace<1>, grid, dim3(256), 0, s, v, io);
    EHSS_HIP(hipGetLastError());
    return EHSS_OK;
}

int launch_pa_query(const SceneView& v, const PAIO& io, hipStream_t s) {
    if (io.nr == 0) return EHSS_OK;
    hipLaunchKernelGGL(k_pa_query, dim3((unsigned)((io.nr + 255) / 256)), dim3(256), 0, s, v, io);
    EHSS_HIP(hipGetLastError());
    return EHSS_OK;
}

}  // namespace ehss
