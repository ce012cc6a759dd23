// solve.hip -- fused EHSS/PA solve of one scrambled Sobol net on the device.
//
// One replicate of solve_frame (ccs_cli.py:196-219) without leaving the GPU:
//   1. u = scrambled Sobol point (SciPy qmc.Sobol, LMS + shift, 30 bits, Gray-code
//      order: q_i = shift ^ XOR_{bit j of gray(i)} sv[j]), exact k * 2^-30;
//   2. external_source (ehss_state.py:42-52, inside_out_pa.py:45-58) in FP64 with
//      NumPy's operation order (only cos/sin come from the device library);
//   3. optional any-hit PA (pa_spatial.py:114-142);
//   4. the reflect-until-escape trace (ehss_cpu.cpp:54-79);
//   5. per-ray records, then a fixed-order FP64 compensated reduction of
//      w (1 - u.v) (ehss_state.py:84-91), PA and tail masses, and the collision
//      histogram / contribution ledger (ccs_cli.py:212-214).
//
// Ray-to-lane mapping: for n = 2^m rays, lane l of wave w traces ray
// i = (l << (m - 6)) | w.  Rays that share their low index bits share the top
// bits of every Sobol coordinate (the generator matrices are upper triangular and
// the LMS scramble is lower triangular), so a wave holds 64 rays from one small
// cell of the 4-D net: nearby directions and impact points, coherent traversal.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "ehss_query.h"

namespace ehss {

struct SolveArgs {
    int64_t n;
    int32_t power;
    int32_t cap;
    double w;      // pi R^2 / n
    double tol;    // grazing tolerance on distances
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

__device__ __forceinline__ void sobol_point(const ehss_sobol& sob, uint32_t i, double u[4]) {
    const uint32_t g = i ^ (i >> 1);
    uint32_t q0 = sob.shift[0], q1 = sob.shift[1], q2 = sob.shift[2], q3 = sob.shift[3];
#pragma unroll
    for (int j = 0; j < 30; ++j) {
        const uint32_t m = 0u - ((g >> j) & 1u);
        q0 ^= sob.sv[0][j] & m;
        q1 ^= sob.sv[1][j] & m;
        q2 ^= sob.sv[2][j] & m;
        q3 ^= sob.sv[3][j] & m;
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

namespace {

template <int kFast, int kPA>
__global__ void __launch_bounds__(256) k_solve(SceneView S, SolveArgs A, ehss_sobol sob) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t boxes = 0, spheres = 0, retests = 0, pboxes = 0, ptests = 0, pretests = 0;
    if (t < (uint64_t)A.n) {
        const uint32_t i = ray_index(t, A.power);
        double u[4];
        sobol_point(sob, i, u);
        V3 o, d;
        source_ray(u, A.cx, A.cy, A.cz, A.radius, o, d);
        const V3 inc = d;
        uint32_t flags = 0;
        if (kPA) {
            bool pgraze = false;
            const bool hit = kFast ? pa_fast(S, o, d, -1, pboxes, ptests, pretests, pgraze)
                                   : pa_plain(S, o, d, -1, pboxes, ptests, pgraze);
            if (hit) flags |= EHSS_RAY_PA_HIT;
            if (pgraze) flags |= EHSS_RAY_GRAZING;
        }
        int32_t last = -1;
        int32_t count = 0;
        bool escaped = false;
        for (;;) {
            Query q;
            if (kFast)
                nearest_fast(S, o, d, last, q, boxes, spheres, retests);
            else
                nearest_plain(S, o, d, last, q, boxes, spheres);
            if (grazing(q, A.tol)) flags |= EHSS_RAY_GRAZING;
            if (q.atom < 0) {
                escaped = true;
                break;
            }
            if (count >= A.cap) {
                flags |= EHSS_RAY_UNRESOLVED;
                break;
            }
            const double4 c = S.sph[q.atom];
            reflect64(o, d, c.x, c.y, c.z, c.w, q.best);
            ++count;
            last = q.atom;
        }
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
    const uint64_t v0 = (uint64_t)wave_sum_i64((int64_t)boxes), v1 = (uint64_t)wave_sum_i64((int64_t)spheres),
                   v2 = (uint64_t)wave_sum_i64((int64_t)retests), v3 = (uint64_t)wave_sum_i64((int64_t)pboxes),
                   v4 = (uint64_t)wave_sum_i64((int64_t)ptests), v5 = (uint64_t)wave_sum_i64((int64_t)pretests);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&A.counters[0], (unsigned long long)v0);
        atomicAdd(&A.counters[1], (unsigned long long)v1);
        atomicAdd(&A.counters[2], (unsigned long long)v2);
        atomicAdd(&A.counters[3], (unsigned long long)v3);
        atomicAdd(&A.counters[4], (unsigned long long)v4);
        atomicAdd(&A.counters[5], (unsigned long long)v5);
    }
}

__global__ void k_zero(unsigned long long* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
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

__global__ void __launch_bounds__(256) k_reduce(int64_t n, int32_t cap, double w, const uint32_t* info,
                                                const double* contrib, Partial* partials,
                                                unsigned long long* hist, unsigned long long* ledger,
                                                int* max_bounces) {
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
    const int64_t base = (int64_t)blockIdx.x * kReduceRays;
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
        partials[blockIdx.x] = p;
    }
    for (int k = threadIdx.x; k < nb; k += blockDim.x) {
        if (h_cnt[k]) atomicAdd(&hist[k], (unsigned long long)h_cnt[k]);
        if (h_led[k]) atomicAdd(&ledger[k], h_led[k]);
    }
    const int wmb = __reduce_max_sync(~0ull, mb);
    if ((threadIdx.x & 63) == 0) atomicMax(max_bounces, wmb);
}

// Combines the block partials in a fixed order.
__global__ void __launch_bounds__(256) k_finalize(const Partial* partials, int64_t nblocks, Partial* out) {
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

__global__ void k_source_rays(int64_t n, double cx, double cy, double cz, double R, ehss_sobol sob, double* u_out,
                              double* o_out, double* d_out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double u[4];
    sobol_point(sob, (uint32_t)i, u);
    V3 o, d;
    source_ray(u, cx, cy, cz, R, o, d);
    if (u_out)
        for (int k = 0; k < 4; ++k) u_out[4 * i + k] = u[k];
    if (o_out) {
        o_out[3 * i] = o.x;
        o_out[3 * i + 1] = o.y;
        o_out[3 * i + 2] = o.z;
    }
    if (d_out) {
        d_out[3 * i] = d.x;
        d_out[3 * i + 1] = d.y;
        d_out[3 * i + 2] = d.z;
    }
}

}  // namespace

int launch_zero(unsigned long long* p, int64_t n, hipStream_t s) {
    if (n <= 0) return EHSS_OK;
    const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(k_zero, dim3(blocks), dim3(256), 0, s, p, n);
    EHSS_HIP(hipGetLastError());
    return EHSS_OK;
}

}  // namespace ehss

using namespace ehss;

namespace {

// Workspace layout in scene->counters: [0..5] kernel counters, [6] max bounces,
// [8..] final Partial.
constexpr int kCounterWords = 8 + (int)(sizeof(Partial) / 8);

int solve_impl(ehss_scene* sc, const ehss_sobol* sob, int power, int cap, int precision, unsigned flags,
               ehss_totals* totals, int64_t* histogram, double* ledger_out) {
    EHSS_CHECK(sc && sob && totals, "null scene/sobol/totals");
    EHSS_CHECK(power >= 0 && power <= 30, "power %d out of range 0..30", power);
    EHSS_CHECK(cap >= 1 && cap <= (int)kBounceMask - 1, "cap %d out of range 1..%u", cap, kBounceMask - 1);
    EHSS_CHECK(precision == EHSS_PRECISION_FP64 || precision == EHSS_PRECISION_FP32_FIXUP, "unknown precision %d",
               precision);
    hipStream_t s = sc->stream;
    const int64_t n = 1LL << power;
    const int64_t nblocks = (n + kReduceRays - 1) / kReduceRays;
    EHSS_TRY(sc->ray_info.reserve(sizeof(uint32_t) * n));
    EHSS_TRY(sc->ray_contrib.reserve(sizeof(double) * n));
    EHSS_TRY(sc->partials.reserve(sizeof(Partial) * nblocks));
    EHSS_TRY(sc->hist.reserve(sizeof(unsigned long long) * (cap + 1)));
    EHSS_TRY(sc->ledger.reserve(sizeof(unsigned long long) * (cap + 1)));
    EHSS_TRY(sc->counters.reserve(sizeof(unsigned long long) * kCounterWords));
    unsigned long long* ctr = sc->counters.as<unsigned long long>();
    EHSS_TRY(launch_zero(ctr, kCounterWords, s));
    EHSS_TRY(launch_zero(sc->hist.as<unsigned long long>(), cap + 1, s));
    EHSS_TRY(launch_zero(sc->ledger.as<unsigned long long>(), cap + 1, s));

    SolveArgs A;
    A.n = n;
    A.power = power;
    A.cap = cap;
    A.w = M_PI * sc->radius * sc->radius / (double)n;  // np.pi*radius*radius/n
    A.tol = kGrazeTol * sc->scale;
    A.cx = sc->center[0];
    A.cy = sc->center[1];
    A.cz = sc->center[2];
    A.radius = sc->radius;
    A.info = sc->ray_info.as<uint32_t>();
    A.contrib = sc->ray_contrib.as<double>();
    A.counters = ctr;
    const bool pa = (flags & EHSS_FLAG_PA) != 0;
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    EHSS_HIP(hipEventRecord(sc->ev0, s));
    if (precision == EHSS_PRECISION_FP64) {
        if (pa)
            hipLaunchKernelGGL((k_solve<0, 1>), grid, block, 0, s, sc->view, A, *sob);
        else
            hipLaunchKernelGGL((k_solve<0, 0>), grid, block, 0, s, sc->view, A, *sob);
    } else {
        if (pa)
            hipLaunchKernelGGL((k_solve<1, 1>), grid, block, 0, s, sc->view, A, *sob);
        else
            hipLaunchKernelGGL((k_solve<1, 0>), grid, block, 0, s, sc->view, A, *sob);
    }
    EHSS_HIP(hipGetLastError());
    EHSS_HIP(hipEventRecord(sc->ev1, s));
    hipLaunchKernelGGL(k_reduce, dim3((unsigned)nblocks), dim3(256), 0, s, n, cap, A.w, A.info, A.contrib,
                       sc->partials.as<Partial>(), sc->hist.as<unsigned long long>(),
                       sc->ledger.as<unsigned long long>(), reinterpret_cast<int*>(ctr + 6));
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, sc->partials.as<Partial>(), nblocks,
                       reinterpret_cast<Partial*>(ctr + 8));
    EHSS_HIP(hipGetLastError());
    unsigned long long host_ctr[kCounterWords];
    EHSS_HIP(hipMemcpyAsync(host_ctr, ctr, sizeof host_ctr, hipMemcpyDeviceToHost, s));
    std::vector<unsigned long long> h_hist(histogram ? cap + 1 : 0), h_led(ledger_out ? cap + 1 : 0);
    if (histogram)
        EHSS_HIP(hipMemcpyAsync(h_hist.data(), sc->hist.p, sizeof(unsigned long long) * (cap + 1),
                                hipMemcpyDeviceToHost, s));
    if (ledger_out)
        EHSS_HIP(hipMemcpyAsync(h_led.data(), sc->ledger.p, sizeof(unsigned long long) * (cap + 1),
                                hipMemcpyDeviceToHost, s));
    EHSS_HIP(hipStreamSynchronize(s));
    float ms = 0.f;
    EHSS_HIP(hipEventElapsedTime(&ms, sc->ev0, sc->ev1));
    Partial fin;
    memcpy(&fin, host_ctr + 8, sizeof fin);
    ehss_totals& T = *totals;
    T.weight = A.w;
    T.omega = fin.s + fin.c;
    T.omega_err = fin.c;
    T.rays = n;
    T.escaped = fin.escaped;
    T.unresolved = fin.unresolved;
    T.pa_hits = pa ? fin.pa_hits : 0;
    T.ehss_hits = fin.ehss_hits;
    T.grazing_rays = fin.grazing;
    T.pa = pa ? A.w * (double)fin.pa_hits : 0.;
    T.pa_from_ehss = A.w * (double)fin.ehss_hits;
    T.residual_mass = A.w * (double)fin.unresolved;
    T.tail_bound = 2 * T.residual_mass;
    T.box_tests = (int64_t)host_ctr[0];
    T.sphere_tests = (int64_t)host_ctr[1];
    T.fp64_retests = (int64_t)(host_ctr[2] + host_ctr[5]);
    T.pa_box_tests = (int64_t)host_ctr[3];
    T.pa_sphere_tests = (int64_t)host_ctr[4];
    T.max_bounces = (int64_t)(int)host_ctr[6];
    T.kernel_ms = ms;
    if (histogram)
        for (int k = 0; k <= cap; ++k) histogram[k] = (int64_t)h_hist[k];
    if (ledger_out)
        for (int k = 0; k <= cap; ++k) ledger_out[k] = (double)h_led[k] / kLedgerScale * A.w;
    sc->last_rays = (flags & EHSS_FLAG_KEEP_RAYS) ? n : 0;
    sc->last_weight = A.w;
    return EHSS_OK;
}

}  // namespace

extern "C" {

int ehss_solve(ehss_scene* scene, const ehss_sobol* sobol, int power, int cap, int precision, unsigned flags,
               ehss_totals* totals, int64_t* histogram, double* ledger) {
    return solve_impl(scene, sobol, power, cap, precision, flags, totals, histogram, ledger);
}

int ehss_scene_last_rays(ehss_scene* scene, int64_t n, uint32_t* flags, double* contrib, int32_t* bounces) {
    EHSS_CHECK(scene, "null scene");
    EHSS_CHECK(scene->last_rays > 0, "no per-ray records kept: solve with EHSS_FLAG_KEEP_RAYS first");
    EHSS_CHECK(n == scene->last_rays, "record count %lld != %lld rays of the last solve", (long long)n,
               (long long)scene->last_rays);
    hipStream_t s = scene->stream;
    std::vector<uint32_t> info((size_t)n);
    EHSS_HIP(hipMemcpyAsync(info.data(), scene->ray_info.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
    if (contrib) EHSS_HIP(hipMemcpyAsync(contrib, scene->ray_contrib.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    EHSS_HIP(hipStreamSynchronize(s));
    for (int64_t i = 0; i < n; ++i) {
        if (flags) flags[i] = info[i] >> kFlagShift;
        if (bounces) bounces[i] = (int32_t)(info[i] & kBounceMask);
    }
    return EHSS_OK;
}

int ehss_source_rays(ehss_scene* scene, const ehss_sobol* sobol, int power, double* u, double* origins,
                     double* directions) {
    EHSS_CHECK(scene && sobol, "null scene/sobol");
    EHSS_CHECK(power >= 0 && power <= 26, "power %d out of range 0..26", power);
    hipStream_t s = scene->stream;
    const int64_t n = 1LL << power;
    DevBuf bu, bo, bd;
    if (u) EHSS_TRY(bu.alloc(sizeof(double) * 4 * n));
    if (origins) EHSS_TRY(bo.alloc(sizeof(double) * 3 * n));
    if (directions) EHSS_TRY(bd.alloc(sizeof(double) * 3 * n));
    hipLaunchKernelGGL(k_source_rays, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, scene->center[0],
                       scene->center[1], scene->center[2], scene->radius, *sobol, bu.as<double>(), bo.as<double>(),
                       bd.as<double>());
    EHSS_HIP(hipGetLastError());
    if (u) EHSS_HIP(hipMemcpyAsync(u, bu.p, sizeof(double) * 4 * n, hipMemcpyDeviceToHost, s));
    if (origins) EHSS_HIP(hipMemcpyAsync(origins, bo.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
    if (directions) EHSS_HIP(hipMemcpyAsync(directions, bd.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
    EHSS_HIP(hipStreamSynchronize(s));
    return EHSS_OK;
}

}  // extern "C"
