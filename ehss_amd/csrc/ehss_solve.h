// ehss_solve.h -- the per-ray body of the fused solve, shared by the single-scene
// kernel (solve.hip) and the multi-job persistent kernel (batch.hip).
//
// One replicate of solve_frame (ccs_cli.py:196-219) without leaving the GPU:
//   1. u = scrambled Sobol point (SciPy qmc.Sobol, LMS + shift, 30 bits, Gray-code
//      order: q_i = shift ^ XOR_{bit j of gray(i)} sv[j]), exact k * 2^-30;
//   2. external_source (ehss_state.py:42-52, inside_out_pa.py:45-58) in FP64 with
//      NumPy's operation order (only cos/sin come from the device library);
//   3. optional any-hit PA (pa_spatial.py:114-142);
//   4. the reflect-until-escape trace (ehss_cpu.cpp:54-79);
//   5. a per-ray record (bounces, flags, w (1 - u.v)) in ray-index order, reduced
//      afterwards in a fixed order.
//
// Ray-to-lane mapping: for n = 2^m rays, lane l of wave w traces ray
// i = (l << (m - 6)) | w.  Rays that share their low index bits share the top
// bits of every Sobol coordinate (the generator matrices are upper triangular and
// the LMS scramble is lower triangular), so a wave holds 64 rays from one small
// cell of the 4-D net: nearby directions and impact points, coherent traversal.
#pragma once
#include "ehss_query.h"

namespace ehss {

struct SolveArgs {
    int64_t n;
    int32_t power;
    int32_t cap;
    double w;      // pi R^2 / n
    double cx, cy, cz, radius;
    uint32_t* info;
    double* contrib;
    unsigned long long* counters;  // boxes, spheres, retests, pa boxes, pa tests, pa retests
};

struct Partial {
    double s, c;
    long long escaped, unresolved, pa_hits, ehss_hits, grazing, pad;
};

__device__ __forceinline__ uint32_t ray_index(uint32_t t, int power) {
    if (power < 6) return t;
    return ((t & 63u) << (power - 6)) | (t >> 6);
}

// Scrambled direction numbers read through the constant address space (scalar loads).
typedef const __attribute__((address_space(4))) uint32_t* CSobol;

__device__ __forceinline__ uint32_t sob_sv(const ehss_sobol& s, int d, int j) { return s.sv[d][j]; }
__device__ __forceinline__ uint32_t sob_shift(const ehss_sobol& s, int d) { return s.shift[d]; }
__device__ __forceinline__ uint32_t sob_sv(CSobol s, int d, int j) { return s[d * 32 + j]; }
__device__ __forceinline__ uint32_t sob_shift(CSobol s, int d) { return s[128 + d]; }

template <class Sob>
__device__ __forceinline__ void sobol_point(Sob sob, uint32_t i, double u[4]) {
    const uint32_t g = i ^ (i >> 1);
    uint32_t q0 = sob_shift(sob, 0), q1 = sob_shift(sob, 1), q2 = sob_shift(sob, 2), q3 = sob_shift(sob, 3);
#pragma unroll
    for (int j = 0; j < 30; ++j) {
        const uint32_t m = 0u - ((g >> j) & 1u);
        q0 ^= sob_sv(sob, 0, j) & m;
        q1 ^= sob_sv(sob, 1, j) & m;
        q2 ^= sob_sv(sob, 2, j) & m;
        q3 ^= sob_sv(sob, 3, j) & m;
    }
    const double scale = 9.313225746154785e-10;  // 2^-30
    u[0] = (double)q0 * scale;
    u[1] = (double)q1 * scale;
    u[2] = (double)q2 * scale;
    u[3] = (double)q3 * scale;
}

// ehss_state.py:45-49 with unit_sphere / tangent_frames (inside_out_pa.py:45-58),
// in the order NumPy evaluates the expressions.
__device__ __forceinline__ void source_ray(const double u[4], double cx, double cy, double cz, double R, V3& o,
                                           V3& d) {
    const double two_pi = 6.283185307179586;  // 2 * np.pi
    const double z = 1 - 2 * u[0];
    const double phi = two_pi * u[1];
    const double rho = sqrt(smax(0., 1 - z * z));
    double s, c;
    sincos(phi, &s, &c);
    d.x = rho * c;
    d.y = rho * s;
    d.z = z;
    double a0 = 0., a1 = 0., a2 = 1.;
    if (fabs(d.z) > .9) {
        a0 = 1.;
        a2 = 0.;
    }
    // t = cross(axes, d); t /= |t|; b = cross(d, t)  (np.cross)
    double t0 = a1 * d.z - a2 * d.y, t1 = a2 * d.x - a0 * d.z, t2 = a0 * d.y - a1 * d.x;
    const double tn = sqrt((t0 * t0 + t1 * t1) + t2 * t2);
    t0 /= tn;
    t1 /= tn;
    t2 /= tn;
    const double b0 = d.y * t2 - d.z * t1, b1 = d.z * t0 - d.x * t2, b2 = d.x * t1 - d.y * t0;
    const double r = R * sqrt(u[2]);
    const double ph = two_pi * u[3];
    double sp, cp;
    sincos(ph, &sp, &cp);
    o.x = (cx - R * d.x) + r * (cp * t0 + sp * b0);
    o.y = (cy - R * d.y) + r * (cp * t1 + sp * b1);
    o.z = (cz - R * d.z) + r * (cp * t2 + sp * b2);
}

// Kernel variants (template kMode):
//   0 plain FP64: reference traversal, separate reference PA traversal, reference counters;
//   1 FP32 filter, one ray per lane, PA fused into the first query;
//   2 FP32 filter, wave-uniform (packet) first query with fused PA, later bounces per lane;
//   3 FP32 filter, packet traversal for every bounce.
// Variants 1-3 return bit-identical per-ray results (the filter only skips
// spheres whose exact FP64 outcome is already decided).
template <int kMode, int kPA, class Sob>
__device__ __forceinline__ void solve_ray(const SceneView& S, const SolveArgs& A, Sob sob, uint32_t t,
                                          uint64_t* ctr) {
    const bool valid = t < (uint64_t)A.n;
    uint64_t boxes = 0, spheres = 0, retests = 0, pboxes = 0, ptests = 0, pretests = 0;
    const uint32_t i = valid ? ray_index(t, A.power) : 0u;
    V3 o = {0., 0., 0.}, d = {0., 0., 1.};
    if (valid) {
        double u[4];
        sobol_point(sob, i, u);
        source_ray(u, A.cx, A.cy, A.cz, A.radius, o, d);
    }
    const V3 inc = d;
    uint32_t flags = 0;
    int32_t last = -1;
    int32_t count = 0;
    bool done = !valid, escaped = false;
    Query q;
    PathError pe = path_error_start(S.scale);
    double tol = path_tol(pe, S.scale);
    auto process = [&]() {
        if (grazing(q, tol)) flags |= EHSS_RAY_GRAZING;
        if (q.atom < 0) {
            escaped = true;
            done = true;
        } else if (count >= A.cap) {
            flags |= EHSS_RAY_UNRESOLVED;
            done = true;
        } else {
            const double4 c = S.sph[q.atom];
            const double cos_inc = reflect64(o, d, c.x, c.y, c.z, c.w, q.best);
            path_error_reflect(pe, q.best, cos_inc, c.w, S.scale);
            tol = path_tol(pe, S.scale);
            ++count;
            last = q.atom;
        }
    };
    if (kMode == 0) {
        if (valid) {
            if (kPA) {
                bool pgraze = false;
                if (pa_plain(S, o, d, -1, tol, pboxes, ptests, pgraze)) flags |= EHSS_RAY_PA_HIT;
                if (pgraze) flags |= EHSS_RAY_GRAZING;
            }
            while (!done) {
                nearest_plain(S, o, d, last, tol, q, boxes, spheres);
                process();
            }
        }
    } else {
        PAState pa = {false, false, false};
        PAState* pap = kPA ? &pa : nullptr;
        if (kMode == 1) {
            if (valid) nearest_fast(S, o, d, -1, tol, q, boxes, spheres, retests, pap);
        } else {
            nearest_packet(S, o, d, -1, tol, valid, q, boxes, spheres, retests, pap);
        }
        if (valid) {
            if (kPA) {
                bool hit = pa.hit, pgraze = false;
                if (!pa.decided) {
                    if (q.atom >= 0)  // a winner without a robust any-hit: run the full any-hit query
                        hit = pa_fast(S, o, d, -1, tol, pboxes, ptests, pretests, pgraze);
                    else
                        pgraze = pa.critical;
                }
                if (hit) flags |= EHSS_RAY_PA_HIT;
                if (pgraze) flags |= EHSS_RAY_GRAZING;
            }
            process();
        }
        if (kMode == 3) {
            for (;;) {
                const bool live = !done;
                if (__ballot(live) == 0) break;
                nearest_packet(S, o, d, last, tol, live, q, boxes, spheres, retests);
                if (live) process();
            }
        } else {
            while (!done) {
                nearest_fast(S, o, d, last, tol, q, boxes, spheres, retests);
                process();
            }
        }
    }
    if (valid) {
        double contrib = 0.;
        if (escaped) {
            flags |= EHSS_RAY_ESCAPED;
            // ehss_state.py:86-87: clip(1 - sum(incoming * outgoing), 0, 2) * w
            double angle = 1 - ((inc.x * d.x + inc.y * d.y) + inc.z * d.z);
            angle = angle < 0. ? 0. : (angle > 2. ? 2. : angle);
            contrib = A.w * angle;
        }
        if (count > 0 || (flags & EHSS_RAY_UNRESOLVED)) flags |= EHSS_RAY_EHSS_HIT;
        A.info[i] = (uint32_t)count | (flags << kFlagShift);
        A.contrib[i] = contrib;
    }
    ctr[0] += boxes;
    ctr[1] += spheres;
    ctr[2] += retests;
    ctr[3] += pboxes;
    ctr[4] += ptests;
    ctr[5] += pretests;
}

// Adds per-lane counters into counters[0..5] with one atomic per wave (order-free).
__device__ __forceinline__ void flush_counters(const uint64_t* ctr, unsigned long long* counters) {
    uint64_t v[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] = (uint64_t)wave_sum_i64((int64_t)ctr[k]);
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < 6; ++k)
            if (v[k]) atomicAdd(&counters[k], (unsigned long long)v[k]);
    }
}

// Exact sum of two doubles (Knuth TwoSum).
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

__device__ __forceinline__ void add_pair(double& s, double& c, double s2, double c2) {
    double e;
    two_sum(s, s2, s, e);
    c = c + (c2 + e);
}

constexpr int kReduceRays = 2048;  // rays per reduction block (256 threads x 8)
constexpr int kLdsBins = 512;      // histogram bins kept in LDS per block
constexpr double kLedgerScale = 1099511627776.0;  // 2^40 fixed point, in units of w

// One reduction block (256 threads) over rays [block * kReduceRays, ...): fixed-order
// compensated sum and counts into partials[block]; integer histogram and fixed-point
// ledger bins by atomics (order-free).
__device__ __forceinline__ void reduce_block(int64_t block, int64_t n, int32_t cap, double w, const uint32_t* info,
                                             const double* contrib, Partial* partials, unsigned long long* hist,
                                             unsigned long long* ledger, int* max_bounces) {
    __shared__ double s_s[256], s_c[256];
    __shared__ long long s_k[5][256];
    __shared__ unsigned int h_cnt[kLdsBins];
    __shared__ unsigned long long h_led[kLdsBins];
    const int nb = min(cap + 1, kLdsBins);
    for (int k = threadIdx.x; k < nb; k += blockDim.x) {
        h_cnt[k] = 0;
        h_led[k] = 0;
    }
    __syncthreads();
    const int64_t base = block * kReduceRays;
    double s = 0., c = 0.;
    long long k_esc = 0, k_unr = 0, k_pa = 0, k_eh = 0, k_gr = 0;
    int mb = 0;
    const double inv = kLedgerScale / w;
#pragma unroll
    for (int k = 0; k < kReduceRays / 256; ++k) {
        const int64_t i = base + k * 256 + threadIdx.x;
        if (i < n) {
            const uint32_t v = info[i];
            const double x = contrib[i];
            double e;
            two_sum(s, x, s, e);
            c += e;
            const uint32_t f = v >> kFlagShift;
            const int b = (int)(v & kBounceMask);
            k_esc += (f & EHSS_RAY_ESCAPED) ? 1 : 0;
            k_unr += (f & EHSS_RAY_UNRESOLVED) ? 1 : 0;
            k_pa += (f & EHSS_RAY_PA_HIT) ? 1 : 0;
            k_eh += (f & EHSS_RAY_EHSS_HIT) ? 1 : 0;
            k_gr += (f & EHSS_RAY_GRAZING) ? 1 : 0;
            mb = max(mb, b);
            const unsigned long long q = (unsigned long long)llrint(x * inv);
            if (b < nb) {
                atomicAdd(&h_cnt[b], 1u);
                if (q) atomicAdd(&h_led[b], q);
            } else {
                atomicAdd(&hist[b], 1ull);
                if (q) atomicAdd(&ledger[b], q);
            }
        }
    }
    s_s[threadIdx.x] = s;
    s_c[threadIdx.x] = c;
    s_k[0][threadIdx.x] = k_esc;
    s_k[1][threadIdx.x] = k_unr;
    s_k[2][threadIdx.x] = k_pa;
    s_k[3][threadIdx.x] = k_eh;
    s_k[4][threadIdx.x] = k_gr;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            add_pair(s_s[threadIdx.x], s_c[threadIdx.x], s_s[threadIdx.x + off], s_c[threadIdx.x + off]);
            for (int q = 0; q < 5; ++q) s_k[q][threadIdx.x] += s_k[q][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        Partial p;
        p.s = s_s[0];
        p.c = s_c[0];
        p.escaped = s_k[0][0];
        p.unresolved = s_k[1][0];
        p.pa_hits = s_k[2][0];
        p.ehss_hits = s_k[3][0];
        p.grazing = s_k[4][0];
        p.pad = 0;
        partials[block] = p;
    }
    for (int k = threadIdx.x; k < nb; k += blockDim.x) {
        if (h_cnt[k]) atomicAdd(&hist[k], (unsigned long long)h_cnt[k]);
        if (h_led[k]) atomicAdd(&ledger[k], h_led[k]);
    }
    const int wmb = __reduce_max_sync(~0ull, mb);
    if ((threadIdx.x & 63) == 0) atomicMax(max_bounces, wmb);
}

// Combines the block partials of one solve in a fixed order (one 256-thread block).
__device__ __forceinline__ void finalize_block(const Partial* partials, int64_t nblocks, Partial* out) {
    __shared__ double s_s[256], s_c[256];
    __shared__ long long s_k[5][256];
    double s = 0., c = 0.;
    long long k[5] = {0, 0, 0, 0, 0};
    for (int64_t b = threadIdx.x; b < nblocks; b += blockDim.x) {
        const Partial p = partials[b];
        add_pair(s, c, p.s, p.c);
        k[0] += p.escaped;
        k[1] += p.unresolved;
        k[2] += p.pa_hits;
        k[3] += p.ehss_hits;
        k[4] += p.grazing;
    }
    s_s[threadIdx.x] = s;
    s_c[threadIdx.x] = c;
    for (int q = 0; q < 5; ++q) s_k[q][threadIdx.x] = k[q];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            add_pair(s_s[threadIdx.x], s_c[threadIdx.x], s_s[threadIdx.x + off], s_c[threadIdx.x + off]);
            for (int q = 0; q < 5; ++q) s_k[q][threadIdx.x] += s_k[q][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        Partial p;
        p.s = s_s[0];
        p.c = s_c[0];
        p.escaped = s_k[0][0];
        p.unresolved = s_k[1][0];
        p.pa_hits = s_k[2][0];
        p.ehss_hits = s_k[3][0];
        p.grazing = s_k[4][0];
        p.pad = 0;
        *out = p;
    }
}

}  // namespace ehss
