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
#include "ehss_internal.h"

namespace ehss {

namespace {

struct Query {
    int32_t atom;
    double best;
    double second;  // second-smallest finite entry distance seen (grazing test)
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

// ehss_cpu.cpp:59-72 with FP64 box tests (reference traversal, reference counters).
__device__ void nearest_plain(const SceneView& S, const V3& o, const V3& d, int32_t last, Query& q,
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

// Same query through the FP32 filter; FP64 entry for every undecided sphere.
__device__ void nearest_fast(const SceneView& S, const V3& o, const V3& d, int32_t last, Query& q,
                             uint64_t& boxes, uint64_t& spheres, uint64_t& retests) {
    q.atom = -1;
    q.best = INFINITY;
    q.second = INFINITY;
    q.graze = INFINITY;
    const float ox = (float)(o.x - S.cx), oy = (float)(o.y - S.cy), oz = (float)(o.z - S.cz);
    const float dx = (float)d.x, dy = (float)d.y, dz = (float)d.z;
    const float sx = fabsf(dx) < 1e-20f ? copysignf(1e-20f, dx) : dx;
    const float sy = fabsf(dy) < 1e-20f ? copysignf(1e-20f, dy) : dy;
    const float sz = fabsf(dz) < 1e-20f ? copysignf(1e-20f, dz) : dz;
    const float ix = 1.0f / sx, iy = 1.0f / sy, iz = 1.0f / sz;
    const float oix = ox * ix, oiy = oy * iy, oiz = oz * iz;
    const float m_p = S.m_p, m_disc = S.m_disc, m_box = S.m_box;
    float bestf = INFINITY;  // FP32 upper bound of best used for box pruning
    int32_t node = 0;
    const int32_t nn = S.n_nodes;
    while (node < nn) {
        const Node32 n = S.nodes32[node];
        ++boxes;
        const float tx0 = __builtin_fmaf(n.lo[0], ix, -oix), tx1 = __builtin_fmaf(n.hi[0], ix, -oix);
        const float ty0 = __builtin_fmaf(n.lo[1], iy, -oiy), ty1 = __builtin_fmaf(n.hi[1], iy, -oiy);
        const float tz0 = __builtin_fmaf(n.lo[2], iz, -oiz), tz1 = __builtin_fmaf(n.hi[2], iz, -oiz);
        const float near = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.f));
        const float far = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
        if (!(near <= far && near <= bestf)) {
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
                const float x0 = ox - s.x, x1 = oy - s.y, x2 = oz - s.z;
                const float p = x0 * dx + x1 * dy + x2 * dz;
                if (p >= m_p) continue;  // projection64 > 0: miss
                const float p0 = x0 - p * dx, p1 = x1 - p * dy, p2 = x2 - p * dz;
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

// Query end: a ray is grazing when a near-threshold sphere test lies at or
// before the winning distance, or when the runner-up is within tolerance.
__device__ __forceinline__ bool grazing(const Query& q, double tol) {
    return q.graze <= q.best + tol || q.second - q.best <= tol;
}

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
            if (kFast)
                nearest_fast(S, o, d, last, q, boxes, spheres, retests);
            else
                nearest_plain(S, o, d, last, q, boxes, spheres);
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
        const int64_t owner = io.owners ? io.owners[i] : -1;
        int32_t node = 0;
        const int32_t stop = S.pa[0].end;
        bool hit = false;
        while (node < stop && !hit) {
            const NodePA& n = S.pa[node];
            ++boxes;
            if (!box_hit64(o, d, n.lo, n.hi)) {
                node = n.end;
                continue;
            }
            for (int32_t k = n.first; k < n.first + n.count; ++k) {
                const int32_t j = S.pa_ids[k];
                if (j != owner) {
                    ++tests;
                    const double4 c = S.sph[j];
                    if (sphere_hit64(o, d, c.x, c.y, c.z, c.w)) {
                        hit = true;
                        break;
                    }
                }
            }
            ++node;
        }
        io.hits[i] = hit ? 1 : 0;
    }
    const uint64_t b = (uint64_t)wave_sum_i64((int64_t)boxes);
    const uint64_t t = (uint64_t)wave_sum_i64((int64_t)tests);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&io.counters[0], (unsigned long long)b);
        atomicAdd(&io.counters[1], (unsigned long long)t);
    }
}

// FP32 filter arrays: boxes rounded outward and inflated by m_box; spheres in
// slot order, recentred on the scene centre.
__global__ void k_make_filter(SceneView S, Node32* nodes32, float4* sph32) {
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
        m.leaf = n.leaf;
        nodes32[i] = m;
    }
    if (i < S.n_atoms) {
        const double4 s = S.sph[S.ids[i]];
        sph32[i] = make_float4((float)(s.x - S.cx), (float)(s.y - S.cy), (float)(s.z - S.cz), (float)s.w);
    }
}

}  // namespace

void set_filter_margins(SceneView& v) {
    const double u = 5.9604644775390625e-08;  // 2^-24
    v.m_box = (float)(64.0 * u * v.scale);
    v.m_p = (float)(64.0 * u * v.scale);
    v.m_disc = (float)(1024.0 * u * v.scale * v.scale);
}

int launch_make_filter(const SceneView& v, Node32* nodes32, float4* sph32, hipStream_t s) {
    const int32_t n = v.n_nodes > v.n_atoms ? v.n_nodes : v.n_atoms;
    if (n == 0) return EHSS_OK;
    hipLaunchKernelGGL(k_make_filter, dim3((n + 255) / 256), dim3(256), 0, s, v, nodes32, sph32);
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
