// sources.hip -- alternative incidence sources generated, traced and reduced on the device.
//
// Restates, in NumPy's operation order (so that the states equal the reference's
// NumPy ones bit for bit wherever the reference does not go through BLAS):
//   * CompactBoundary.source, kinds 'box' and 'ellipsoid' (ehss_compact_source.py:44-87);
//   * sample_source (inside_out_pa.py:72-90), the inside-out any-hit test
//     (inside_out_pa.py:93-118: every sphere but the owner) and reverse_pa_source
//     (ehss_state.py:55-69: the unblocked states reflected once);
//   * the EHSSResult readout with per-ray weights (ehss_state.py:84-103).
// unit_sphere / tangent_frames are inside_out_pa.py:45-58; the cos/sin of the
// Sobol angles are the host libm's bits (ehss_trig::sincos_sobol), as in the
// impact-disk source.
//
// One solve: k_source writes each ray's initial state into the drain queue (one
// ballot and one atomic per wave), the fused solve's drain traces the queue to
// the end (drain_queued, records with weight 1, i.e. the angle), and
// k_weight_reduce scales each record by its ray's weight and sums the block in
// a fixed order; k_weight_finalize combines the blocks in a fixed order.
// Records sit at the ray's index, so the sums do not depend on queue order.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "ehss_internal.h"
#include "ehss_solve.h"

namespace ehss {
namespace {

struct SourceParams {
    int32_t kind, n_atoms;
    int64_t n;             // rays
    double center[3];
    double rot[9];         // row-major
    double axes[3];
    double face_area[3];   // box: 4 prod(axes) / axes
    double lift;           // box: 1e-10 max(1, max(axes))
    double inv_n;          // divisor of the per-ray area: n (as a double)
    double w_reverse;      // reverse PA: area / (4 n)
    const double* cdf;     // reverse PA: device copy of the searchsorted table
    ehss_trig::CompatTable trig;
};

struct SrcState {
    V3 o, d, inc;
    double w;
    int32_t owner, count;
    bool blocked;
};

// Point i of random_base2: Gray index g = i ^ (i >> 1), q = shift ^ XOR of the
// direction numbers of g's set bits.
__device__ __forceinline__ void net_point(const ehss_sobol5& sob, uint32_t i, uint32_t q[5]) {
    const uint32_t g = i ^ (i >> 1);
#pragma unroll
    for (int k = 0; k < 5; ++k) q[k] = sob.shift[k];
#pragma unroll
    for (int j = 0; j < 30; ++j) {
        const uint32_t m = 0u - ((g >> j) & 1u);
#pragma unroll
        for (int k = 0; k < 5; ++k) q[k] ^= sob.sv[k][j] & m;
    }
}

__device__ __forceinline__ double word_u(uint32_t q) { return (double)q * 9.313225746154785e-10; }  // q 2^-30

// unit_sphere row (inside_out_pa.py:45-49): z = 1 - 2 uz, phi = 2pi uphi.
__device__ __forceinline__ V3 unit_row(uint32_t qz, uint32_t qphi, const ehss_trig::CompatTable& trig) {
    const double two_pi = 6.283185307179586;  // 2 * np.pi
    const double z = 1 - 2 * word_u(qz);
    const double phi = two_pi * word_u(qphi);
    const double rho = sqrt(smax(0., 1 - z * z));
    double s, c;
    ehss_trig::sincos_sobol(phi, qphi, trig, &s, &c);
    return V3{rho * c, rho * s, z};
}

// tangent_frames row (inside_out_pa.py:52-58): t = cross(axes, n) / |.|, b = cross(n, t).
__device__ __forceinline__ void frames_row(const V3& n, V3& t, V3& b) {
    double a0 = 0., a1 = 0., a2 = 1.;
    if (fabs(n.z) > .9) {
        a0 = 1.;
        a2 = 0.;
    }
    double t0 = a1 * n.z - a2 * n.y, t1 = a2 * n.x - a0 * n.z, t2 = a0 * n.y - a1 * n.x;
    const double tn = sqrt((t0 * t0 + t1 * t1) + t2 * t2);
    t = V3{t0 / tn, t1 / tn, t2 / tn};
    b = V3{n.y * t.z - n.z * t.y, n.z * t.x - n.x * t.z, n.x * t.y - n.y * t.x};
}

// Row of v @ rot (column j of the rotation).
__device__ __forceinline__ double vrot(const V3& v, const double* rot, int j) {
    return (v.x * rot[j] + v.y * rot[3 + j]) + v.z * rot[6 + j];
}

__device__ __forceinline__ double np_sign(double v) { return v > 0 ? 1. : (v < 0 ? -1. : 0.); }

// ehss_compact_source.py:56-70 for one ray.
__device__ __forceinline__ void box_state(const SourceParams& P, const uint32_t q[5], SrcState& st) {
    const V3 d = unit_row(q[0], q[1], P.trig);
    const double ld[3] = {vrot(d, P.rot, 0), vrot(d, P.rot, 1), vrot(d, P.rot, 2)};
    const double p0 = fabs(ld[0]) * P.face_area[0], p1 = fabs(ld[1]) * P.face_area[1],
                 p2 = fabs(ld[2]) * P.face_area[2];
    const double c1 = p0 + p1;
    const double area = c1 + p2;
    const double x = word_u(q[2]) * area;
    int face = (x >= p0) + (x >= c1) + (x >= area);
    face = face > 2 ? 2 : face;  // u < 1: x < area
    double p[3];
    p[face] = -np_sign(ld[face]) * P.axes[face];
    const int o0 = face == 0 ? 1 : 0, o1 = face == 2 ? 1 : 2;
    p[o0] = (2 * word_u(q[3]) - 1) * P.axes[o0];
    p[o1] = (2 * word_u(q[4]) - 1) * P.axes[o1];
    const double* R = P.rot;
    const double w0 = (p[0] * R[0] + p[1] * R[1]) + p[2] * R[2];  // p @ rotation.T
    const double w1 = (p[0] * R[3] + p[1] * R[4]) + p[2] * R[5];
    const double w2 = (p[0] * R[6] + p[1] * R[7]) + p[2] * R[8];
    st.o = V3{(P.center[0] + w0) - P.lift * d.x, (P.center[1] + w1) - P.lift * d.y, (P.center[2] + w2) - P.lift * d.z};
    st.d = d;
    st.inc = d;
    st.w = area / P.inv_n;
    st.owner = -1;
    st.count = 0;
    st.blocked = false;
}

// ehss_compact_source.py:71-84 for one ray.
__device__ __forceinline__ void ellipsoid_state(const SourceParams& P, const uint32_t q[5], SrcState& st) {
    const V3 d = unit_row(q[0], q[1], P.trig);
    V3 t, b;
    frames_row(d, t, b);
    double lt[3], lb[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        lt[j] = vrot(t, P.rot, j) * P.axes[j];
        lb[j] = vrot(b, P.rot, j) * P.axes[j];
    }
    const double g00 = (lt[0] * lt[0] + lt[1] * lt[1]) + lt[2] * lt[2];
    const double g01 = (lt[0] * lb[0] + lt[1] * lb[1]) + lt[2] * lb[2];
    const double g11 = (lb[0] * lb[0] + lb[1] * lb[1]) + lb[2] * lb[2];
    const double l00 = sqrt(g00);
    const double l10 = g01 / l00;
    const double l11 = sqrt(smax(0., g11 - l10 * l10));
    const double two_pi = 6.283185307179586;
    const double rho = sqrt(word_u(q[2]));
    const double phi = two_pi * word_u(q[3]);
    double s, c;
    ehss_trig::sincos_sobol(phi, q[3], P.trig, &s, &c);
    const double x = rho * c, y = rho * s;
    const double a = l00 * x, bb = l10 * x + l11 * y;
    const V3 off{a * t.x + bb * b.x, a * t.y + bb * b.y, a * t.z + bb * b.z};
    double e[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) e[j] = vrot(d, P.rot, j) * P.axes[j];
    const double extent = sqrt((e[0] * e[0] + e[1] * e[1]) + e[2] * e[2]);
    const double back = extent * (1 + 1e-10);
    st.o = V3{(P.center[0] + off.x) - back * d.x, (P.center[1] + off.y) - back * d.y,
              (P.center[2] + off.z) - back * d.z};
    st.d = d;
    st.inc = d;
    st.w = (3.141592653589793 * l00) * l11 / P.inv_n;
    st.owner = -1;
    st.count = 0;
    st.blocked = false;
}

// sample_source + inside-out any-hit + reverse_pa_source's reflection for one ray.
__device__ __forceinline__ void reverse_pa_state(const SceneView& S, const SourceParams& P, const uint32_t q[5],
                                                 SrcState& st) {
    // ids = searchsorted(cdf, u0, side='right'): the first entry above u0
    const double u0 = word_u(q[0]);
    int32_t lo = 0, hi = P.n_atoms;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (P.cdf[mid] <= u0)
            lo = mid + 1;
        else
            hi = mid;
    }
    const int32_t id = lo < P.n_atoms ? lo : P.n_atoms - 1;
    const V3 nrm = unit_row(q[1], q[2], P.trig);
    V3 t, b;
    frames_row(nrm, t, b);
    const double two_pi = 6.283185307179586;
    const double u3 = word_u(q[3]);
    const double phi = two_pi * word_u(q[4]);
    const double rho = sqrt(u3);
    double s, c;
    ehss_trig::sincos_sobol(phi, q[4], P.trig, &s, &c);
    const double sn = sqrt(1 - u3), rc = rho * c, rs = rho * s;
    const V3 dir{(sn * nrm.x + rc * t.x) + rs * b.x, (sn * nrm.y + rc * t.y) + rs * b.y,
                 (sn * nrm.z + rc * t.z) + rs * b.z};
    const double4 sp = S.sph[id];
    const V3 o{sp.x + sp.w * nrm.x, sp.y + sp.w * nrm.y, sp.z + sp.w * nrm.z};
    uint64_t boxes = 0, tests = 0;
    bool graze = false;
    st.blocked = pa_plain(S, o, dir, id, 0., boxes, tests, graze);
    st.o = o;
    st.owner = id;
    st.count = 1;
    if (st.blocked) {
        st.d = dir;
        st.inc = V3{-dir.x, -dir.y, -dir.z};
        st.w = 0.;
        return;
    }
    const V3 inc{-dir.x, -dir.y, -dir.z};
    V3 nn{(o.x - sp.x) / sp.w, (o.y - sp.y) / sp.w, (o.z - sp.z) / sp.w};
    const double nl = sqrt((nn.x * nn.x + nn.y * nn.y) + nn.z * nn.z);
    nn = V3{nn.x / nl, nn.y / nl, nn.z / nl};
    const double two_dot = 2 * ((inc.x * nn.x + inc.y * nn.y) + inc.z * nn.z);
    V3 out{inc.x - two_dot * nn.x, inc.y - two_dot * nn.y, inc.z - two_dot * nn.z};
    const double ol = sqrt((out.x * out.x + out.y * out.y) + out.z * out.z);
    st.d = V3{out.x / ol, out.y / ol, out.z / ol};
    st.inc = inc;
    st.w = P.w_reverse;
}

struct StateOut {
    double *o, *d, *inc, *w;
    int64_t *owner, *count;
    uint8_t* blocked;
};

// One thread per ray: the initial state.  With a queue, unblocked states are
// queued (one ballot and one atomic per wave), blocked ones recorded as zero and
// counted; with `out`, the states are written out instead (parity tests).
template <int kKind>
__global__ void __launch_bounds__(256) k_source(SceneView S, SourceParams P, ehss_sobol5 sob, SolveArgs A,
                                                RayState* queue, uint32_t* queue_words, double* weights,
                                                unsigned long long* blocked_count, StateOut out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < P.n;
    SrcState st;
    st.blocked = false;
    if (valid) {
        uint32_t q[5];
        net_point(sob, (uint32_t)i, q);
        if constexpr (kKind == EHSS_SOURCE_BOX) box_state(P, q, st);
        else if constexpr (kKind == EHSS_SOURCE_ELLIPSOID) ellipsoid_state(P, q, st);
        else reverse_pa_state(S, P, q, st);
    }
    if (out.o) {
        if (valid) {
            const double* v[3] = {&st.o.x, &st.d.x, &st.inc.x};
            double* dst[3] = {out.o, out.d, out.inc};
            for (int a = 0; a < 3; ++a)
                if (dst[a])
                    for (int k = 0; k < 3; ++k) dst[a][3 * i + k] = v[a][k];
            if (out.w) out.w[i] = st.w;
            if (out.owner) out.owner[i] = st.owner;
            if (out.count) out.count[i] = st.count;
            if (out.blocked) out.blocked[i] = st.blocked ? 1 : 0;
        }
        return;
    }
    const bool traced = valid && !st.blocked;
    const uint64_t m = __ballot(traced);
    if (valid) {
        weights[i] = st.w;
        if (st.blocked) {
            A.info[i] = 0u;
            A.contrib[i] = 0.;
        }
    }
    const uint64_t mb = __ballot(valid && st.blocked);
    if ((threadIdx.x & 63) == 0 && mb) atomicAdd(blocked_count, (unsigned long long)__popcll(mb));
    if (m == 0) return;
    uint32_t base = 0;
    if ((threadIdx.x & 63) == 0) base = atomicAdd(queue_words, (uint32_t)__popcll(m));
    base = __shfl(base, 0, 64);
    if (traced) {
        RayCtx R;
        R.o = st.o;
        R.d = st.d;
        R.inc = st.inc;
        R.pe = path_error_start(S.scale);
        R.i = (uint32_t)i;
        R.flags = 0;
        R.last = st.owner;
        R.count = st.count;
        const uint64_t below = m & ((1ull << (threadIdx.x & 63)) - 1ull);
        ray_save(R, 0, queue + base + (uint32_t)__popcll(below));
    }
}

// Weighted sums of one solve: omega, residual mass, pa_from_ehss, total weight.
struct WPartial {
    double s[4], c[4];
    long long escaped, unresolved, ehss_hits, grazing;
};

constexpr int kWReduceRays = 2048;

// One block of kWReduceRays rays: contrib[i] = angle * weight[i] (written back),
// then fixed-order compensated sums and counts into partials[block].
__global__ void __launch_bounds__(256) k_weight_reduce(int64_t n, int32_t cap, const uint32_t* info, double* contrib,
                                                       const double* weights, WPartial* partials,
                                                       unsigned long long* hist, int* max_bounces) {
    __shared__ double s_s[4][256], s_c[4][256];
    __shared__ long long s_k[4][256];
    double s[4] = {0., 0., 0., 0.}, c[4] = {0., 0., 0., 0.};
    long long k_esc = 0, k_unr = 0, k_eh = 0, k_gr = 0;
    int mb = 0;
    const int64_t base = (int64_t)blockIdx.x * kWReduceRays;
    for (int k = 0; k < kWReduceRays / 256; ++k) {
        const int64_t i = base + k * 256 + threadIdx.x;
        if (i >= n) continue;
        const uint32_t v = info[i];
        const double w = weights[i];
        const double x = contrib[i] * w;  // weights * angle * escaped (ehss_state.py:86-87)
        contrib[i] = x;
        const uint32_t f = v >> kFlagShift;
        const int b = (int)(v & kBounceMask);
        const double add[4] = {x, (f & EHSS_RAY_UNRESOLVED) ? w : 0., (f & EHSS_RAY_EHSS_HIT) ? w : 0., w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            double e;
            two_sum(s[q], add[q], s[q], e);
            c[q] += e;
        }
        k_esc += (f & EHSS_RAY_ESCAPED) ? 1 : 0;
        k_unr += (f & EHSS_RAY_UNRESOLVED) ? 1 : 0;
        k_eh += (f & EHSS_RAY_EHSS_HIT) ? 1 : 0;
        k_gr += (f & EHSS_RAY_GRAZING) ? 1 : 0;
        mb = max(mb, b);
        if (w != 0. && b <= cap) atomicAdd(&hist[b], 1ull);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        s_s[q][threadIdx.x] = s[q];
        s_c[q][threadIdx.x] = c[q];
    }
    s_k[0][threadIdx.x] = k_esc;
    s_k[1][threadIdx.x] = k_unr;
    s_k[2][threadIdx.x] = k_eh;
    s_k[3][threadIdx.x] = k_gr;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            for (int q = 0; q < 4; ++q) {
                add_pair(s_s[q][threadIdx.x], s_c[q][threadIdx.x], s_s[q][threadIdx.x + off],
                         s_c[q][threadIdx.x + off]);
                s_k[q][threadIdx.x] += s_k[q][threadIdx.x + off];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        WPartial p;
        for (int q = 0; q < 4; ++q) {
            p.s[q] = s_s[q][0];
            p.c[q] = s_c[q][0];
        }
        p.escaped = s_k[0][0];
        p.unresolved = s_k[1][0];
        p.ehss_hits = s_k[2][0];
        p.grazing = s_k[3][0];
        partials[blockIdx.x] = p;
    }
    const int wmb = __reduce_max_sync(~0ull, mb);
    if ((threadIdx.x & 63) == 0) atomicMax(max_bounces, wmb);
}

// Combines the block partials in a fixed order (one 256-thread block).
__global__ void __launch_bounds__(256) k_weight_finalize(const WPartial* partials, int64_t nblocks, WPartial* out) {
    __shared__ double s_s[4][256], s_c[4][256];
    __shared__ long long s_k[4][256];
    double s[4] = {0., 0., 0., 0.}, c[4] = {0., 0., 0., 0.};
    long long k[4] = {0, 0, 0, 0};
    for (int64_t b = threadIdx.x; b < nblocks; b += blockDim.x) {
        const WPartial p = partials[b];
        for (int q = 0; q < 4; ++q) add_pair(s[q], c[q], p.s[q], p.c[q]);
        k[0] += p.escaped;
        k[1] += p.unresolved;
        k[2] += p.ehss_hits;
        k[3] += p.grazing;
    }
    for (int q = 0; q < 4; ++q) {
        s_s[q][threadIdx.x] = s[q];
        s_c[q][threadIdx.x] = c[q];
        s_k[q][threadIdx.x] = k[q];
    }
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            for (int q = 0; q < 4; ++q) {
                add_pair(s_s[q][threadIdx.x], s_c[q][threadIdx.x], s_s[q][threadIdx.x + off],
                         s_c[q][threadIdx.x + off]);
                s_k[q][threadIdx.x] += s_k[q][threadIdx.x + off];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        WPartial p;
        for (int q = 0; q < 4; ++q) {
            p.s[q] = s_s[q][0];
            p.c[q] = s_c[q][0];
        }
        p.escaped = s_k[0][0];
        p.unresolved = s_k[1][0];
        p.ehss_hits = s_k[2][0];
        p.grazing = s_k[3][0];
        *out = p;
    }
}

// Validates the source and fills the device parameters (the searchsorted table
// is uploaded into the scene's buffer on stream s).
int source_params(ehss_scene* sc, const ehss_source* src, int power, SourceParams* P, hipStream_t s) {
    EHSS_CHECK(src, "null source");
    EHSS_CHECK(power >= 0 && power <= 30, "power %d out of range 0..30", power);
    memset(P, 0, sizeof *P);
    P->kind = src->kind;
    P->n_atoms = sc->n;
    P->n = 1LL << power;
    P->inv_n = (double)P->n;
    P->trig = trig_compat();
    memcpy(P->center, src->center, sizeof P->center);
    memcpy(P->rot, src->rotation, sizeof P->rot);
    memcpy(P->axes, src->axes, sizeof P->axes);
    switch (src->kind) {
        case EHSS_SOURCE_BOX: {
            for (int k = 0; k < 3; ++k)
                EHSS_CHECK(isfinite(src->axes[k]) && src->axes[k] > 0, "box half-widths must be positive");
            const double prod = (src->axes[0] * src->axes[1]) * src->axes[2];  // np.prod
            for (int k = 0; k < 3; ++k) P->face_area[k] = 4 * prod / src->axes[k];  // 4*np.prod(axes)/axes
            const double amax = std::max({src->axes[0], src->axes[1], src->axes[2]});
            P->lift = 1e-10 * std::max(1., amax);
            return EHSS_OK;
        }
        case EHSS_SOURCE_ELLIPSOID:
            for (int k = 0; k < 3; ++k)
                EHSS_CHECK(isfinite(src->axes[k]) && src->axes[k] > 0, "ellipsoid semi-axes must be positive");
            return EHSS_OK;
        case EHSS_SOURCE_REVERSE_PA: {
            EHSS_CHECK(src->cdf, "reverse-PA source needs the searchsorted table");
            EHSS_CHECK(isfinite(src->area) && src->area > 0, "reverse-PA source needs the spheres' area");
            EHSS_TRY(sc->src_cdf.reserve(sizeof(double) * sc->n));
            EHSS_HIP(hipMemcpyAsync(sc->src_cdf.p, src->cdf, sizeof(double) * sc->n, hipMemcpyHostToDevice, s));
            P->cdf = sc->src_cdf.as<double>();
            P->w_reverse = src->area / (double)(4 * P->n);  // source.area / (4 * len(source.owners))
            return EHSS_OK;
        }
        default:
            return set_error(EHSS_ERR_INVALID, "unknown source kind %d", src->kind);
    }
}

template <class F>
int by_kind(int kind, F f) {
    switch (kind) {
        case EHSS_SOURCE_BOX: return f(std::integral_constant<int, EHSS_SOURCE_BOX>());
        case EHSS_SOURCE_ELLIPSOID: return f(std::integral_constant<int, EHSS_SOURCE_ELLIPSOID>());
        default: return f(std::integral_constant<int, EHSS_SOURCE_REVERSE_PA>());
    }
}

}  // namespace
}  // namespace ehss

using namespace ehss;

extern "C" {

int ehss_source_states(ehss_scene* scene, const ehss_source* source, const ehss_sobol5* sobol, int power,
                       double* origins, double* directions, double* incoming, double* weights, int64_t* owners,
                       int64_t* bounces, uint8_t* blocked) {
    EHSS_TRY(ensure_init());
    EHSS_CHECK(scene && sobol, "null scene/sobol");
    std::lock_guard<std::mutex> guard(scene->lock);
    if (scene->pending) return set_error(EHSS_ERR_BUSY, "a solve is pending on this scene");
    hipStream_t s = scene->stream;
    SourceParams P;
    EHSS_TRY(source_params(scene, source, power, &P, s));
    const int64_t n = P.n;
    DevBuf dv, dw, di, db;
    EHSS_TRY(dv.alloc(sizeof(double) * 9 * n));
    EHSS_TRY(dw.alloc(sizeof(double) * n));
    EHSS_TRY(di.alloc(sizeof(int64_t) * 2 * n));
    EHSS_TRY(db.alloc(n));
    StateOut out;
    out.o = dv.as<double>();
    out.d = out.o + 3 * n;
    out.inc = out.o + 6 * n;
    out.w = dw.as<double>();
    out.owner = di.as<int64_t>();
    out.count = out.owner + n;
    out.blocked = db.as<uint8_t>();
    const dim3 grid((unsigned)((n + 255) / 256));
    EHSS_TRY(by_kind(P.kind, [&](auto k) -> int {
        hipLaunchKernelGGL(k_source<decltype(k)::value>, grid, dim3(256), 0, s, scene->view, P, *sobol, SolveArgs(),
                           nullptr, nullptr, nullptr, nullptr, out);
        return EHSS_OK;
    }));
    EHSS_HIP(hipGetLastError());
    EHSS_HIP(hipStreamSynchronize(s));
    if (origins) EHSS_HIP(hipMemcpy(origins, out.o, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
    if (directions) EHSS_HIP(hipMemcpy(directions, out.d, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
    if (incoming) EHSS_HIP(hipMemcpy(incoming, out.inc, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
    if (weights) EHSS_HIP(hipMemcpy(weights, out.w, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (owners) EHSS_HIP(hipMemcpy(owners, out.owner, sizeof(int64_t) * n, hipMemcpyDeviceToHost));
    if (bounces) EHSS_HIP(hipMemcpy(bounces, out.count, sizeof(int64_t) * n, hipMemcpyDeviceToHost));
    if (blocked) EHSS_HIP(hipMemcpy(blocked, out.blocked, n, hipMemcpyDeviceToHost));
    return EHSS_OK;
}

int ehss_solve_source(ehss_scene* scene, const ehss_source* source, const ehss_sobol5* sobol, int power, int cap,
                      ehss_totals* totals, int64_t* histogram, uint32_t* info, double* contrib) {
    EHSS_TRY(ensure_init());
    EHSS_CHECK(scene && sobol && totals, "null scene/sobol/totals");
    EHSS_CHECK(cap >= 1 && cap <= (int)kBounceMask - 1, "cap %d out of range 1..%u", cap, kBounceMask - 1);
    std::lock_guard<std::mutex> guard(scene->lock);
    if (scene->pending) return set_error(EHSS_ERR_BUSY, "a solve is already pending on this scene");
    hipStream_t s = scene->stream;
    SourceParams P;
    EHSS_TRY(source_params(scene, source, power, &P, s));
    const int64_t n = P.n;
    const int64_t nblocks = (n + kWReduceRays - 1) / kWReduceRays;
    const int64_t bins = cap + 1;
    EHSS_TRY(scene->ray_info.reserve(sizeof(uint32_t) * n));
    EHSS_TRY(scene->ray_contrib.reserve(sizeof(double) * n));
    EHSS_TRY(scene->src_weights.reserve(sizeof(double) * n));
    EHSS_TRY(scene->queue.reserve(sizeof(RayState) * n));
    EHSS_TRY(scene->partials.reserve(sizeof(WPartial) * (nblocks + 1)));
    EHSS_TRY(scene->hist.reserve(sizeof(unsigned long long) * bins));
    // counters: [0..5] kernel counters, [6] max bounces, [7] blocked samples, [8..9] queue words
    const int64_t words = kCounterWords;
    EHSS_TRY(scene->counters.reserve(sizeof(unsigned long long) * words));
    unsigned long long* ctr = scene->counters.as<unsigned long long>();
    EHSS_TRY(launch_zero(ctr, words, s));
    EHSS_TRY(launch_zero(scene->hist.as<unsigned long long>(), bins, s));
    uint32_t* qwords = reinterpret_cast<uint32_t*>(ctr + 8);

    SolveArgs A;
    memset(&A, 0, sizeof A);
    A.n = n;
    A.power = power;
    A.cap = cap;
    A.w = 1.;  // records hold the angle; k_weight_reduce applies each ray's weight
    A.cx = scene->center[0];
    A.cy = scene->center[1];
    A.cz = scene->center[2];
    A.radius = scene->radius;
    A.info = scene->ray_info.as<uint32_t>();
    A.contrib = scene->ray_contrib.as<double>();
    A.counters = ctr;
    A.coop = coop_from_flags(0);
    A.group = group_from_flags(0, n);
    A.packets = packets_from_flags(0);
    A.trig = trig_compat();
    A.cont = nullptr;
    A.cont_count = nullptr;
    A.times = nullptr;
    scene->cont_valid = false;  // the records are overwritten
    scene->last_rays = 0;
    scene->last_times = 0;
    double* wts = scene->src_weights.as<double>();
    StateOut none;
    memset(&none, 0, sizeof none);
    EHSS_HIP(hipEventRecord(scene->ev0, s));
    const dim3 grid((unsigned)((n + 255) / 256));
    EHSS_TRY(by_kind(P.kind, [&](auto k) -> int {
        hipLaunchKernelGGL(k_source<decltype(k)::value>, grid, dim3(256), 0, s, scene->view, P, *sobol, A,
                           scene->queue.as<RayState>(), qwords, wts, ctr + 7, none);
        return EHSS_OK;
    }));
    EHSS_HIP(hipGetLastError());
    EHSS_TRY(drain_queued(scene, A, scene->queue.as<RayState>(), qwords, n, 0u, s));
    EHSS_HIP(hipEventRecord(scene->ev1, s));
    WPartial* parts = scene->partials.as<WPartial>();
    hipLaunchKernelGGL(k_weight_reduce, dim3((unsigned)nblocks), dim3(256), 0, s, n, cap, A.info, A.contrib, wts,
                       parts, scene->hist.as<unsigned long long>(), reinterpret_cast<int*>(ctr + 6));
    hipLaunchKernelGGL(k_weight_finalize, dim3(1), dim3(256), 0, s, parts, nblocks, parts + nblocks);
    EHSS_HIP(hipGetLastError());
    EHSS_HIP(hipStreamSynchronize(s));
    unsigned long long hc[8];
    WPartial fin;
    EHSS_HIP(hipMemcpy(hc, ctr, sizeof hc, hipMemcpyDeviceToHost));
    EHSS_HIP(hipMemcpy(&fin, parts + nblocks, sizeof fin, hipMemcpyDeviceToHost));
    float ms = 0.f;
    EHSS_HIP(hipEventElapsedTime(&ms, scene->ev0, scene->ev1));
    ehss_totals& T = *totals;
    memset(&T, 0, sizeof T);
    T.omega = fin.s[0] + fin.c[0];
    T.omega_err = fin.c[0];
    T.residual_mass = fin.s[1] + fin.c[1];
    T.tail_bound = 2 * T.residual_mass;
    T.pa_from_ehss = fin.s[2] + fin.c[2];
    T.weight = fin.s[3] + fin.c[3];
    T.rays = n;
    T.escaped = fin.escaped;
    T.unresolved = fin.unresolved;
    T.ehss_hits = fin.ehss_hits;
    T.grazing_rays = fin.grazing;
    T.box_tests = (int64_t)hc[0];
    T.sphere_tests = (int64_t)hc[1];
    T.fp64_retests = (int64_t)hc[2];
    T.max_bounces = (int64_t)(int)hc[6];
    if (P.kind == EHSS_SOURCE_REVERSE_PA) {
        T.pa_hits = (int64_t)hc[7];
        // source.area * mean(~blocked) / 4 (inside_out_pa.py:66-67)
        const double mean = (double)(n - (int64_t)hc[7]) / (double)n;
        T.pa = source->area * mean / 4;
    }
    T.kernel_ms = ms;
    if (histogram) {
        std::vector<unsigned long long> h((size_t)bins);
        EHSS_HIP(hipMemcpy(h.data(), scene->hist.p, sizeof(unsigned long long) * bins, hipMemcpyDeviceToHost));
        for (int64_t k = 0; k < bins; ++k) histogram[k] = (int64_t)h[k];
    }
    if (info) EHSS_HIP(hipMemcpy(info, A.info, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    if (contrib) EHSS_HIP(hipMemcpy(contrib, A.contrib, sizeof(double) * n, hipMemcpyDeviceToHost));
    return EHSS_OK;
}

}  // extern "C"
