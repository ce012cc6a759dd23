// ehss_solve.h -- the per-ray body of the fused solve, shared by the single-scene
// kernels (solve.hip) and the multi-structure plan kernels (multi.hip).
//
// One replicate of solve_frame (ccs_cli.py:196-219) without leaving the GPU:
//   1. u = scrambled Sobol point (SciPy qmc.Sobol, LMS + shift, 30 bits, Gray-code
//      order: q_i = shift ^ XOR_{bit j of gray(i)} sv[j]), exact k * 2^-30;
//   2. external_source (ehss_state.py:42-52, inside_out_pa.py:45-58) in FP64 with
//      NumPy's operation order; cos/sin are correctly rounded (ehss_trig.h) and
//      corrected to the host libm's bits by its compatibility table;
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
#include <type_traits>

#include "ehss_query.h"
#include "ehss_trig.h"

// Minimum waves per SIMD the packet kernels and the queue drains are compiled
// for (the register budget follows: 512 / waves VGPRs per lane).  Both kernels
// are latency-bound, so resident waves are what hides their memory round trips.
// Measured on the 650-atom globule (2^20 rays) and on 18 globules of 200-5000
// atoms (2^20 rays each): packet kernels at 7 waves (72 VGPRs, 108 B spilled
// per lane) run 9.68 ms per 16 x 2^20 rays against 9.97 at 6 and 9.79 at 8; the
// single-scene drain, whose LDS-staged scene already limits it to 3 blocks per
// CU, gains nothing from more; the multi-scene drain, which reads the scenes
// from HBM, gains 5-8 % at 4.
// Packet queries of the fast kernel visit children front to back (0: the
// stackless preorder walk).
#ifndef EHSS_ORDERED_PACKETS
#define EHSS_ORDERED_PACKETS 1
#endif
// Packet queries of the fast kernel in two passes (nearest_packet_2p: FP32 walk
// collecting candidates, then lock-step FP64 re-tests; 0: re-tests during the
// walk, nearest_packet_ordered).
#ifndef EHSS_TWO_PASS_PACKETS
#define EHSS_TWO_PASS_PACKETS 1
#endif
#ifndef EHSS_SOLVE_WAVES
#define EHSS_SOLVE_WAVES 7
#endif
#ifndef EHSS_DRAIN_WAVES
#define EHSS_DRAIN_WAVES 3
#endif
#ifndef EHSS_PLAN_DRAIN_WAVES
#define EHSS_PLAN_DRAIN_WAVES 4
#endif
// The plain FP64 kernel (variant 0) is compiled for 4 waves per SIMD: it runs
// 17.0 ms per 2^20-ray solve there against 20.5 ms at 6 (its spills cost more
// than the extra waves hide).
#ifndef EHSS_PLAIN_WAVES
#define EHSS_PLAIN_WAVES 4
#endif
#define EHSS_SOLVE_KERNEL                 \
    __launch_bounds__(256) __attribute__( \
        (amdgpu_waves_per_eu(kMode == 0 ? EHSS_PLAIN_WAVES : EHSS_SOLVE_WAVES)))
#define EHSS_DRAIN_KERNEL __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(EHSS_DRAIN_WAVES)))
#define EHSS_PLAN_DRAIN_KERNEL __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(EHSS_PLAN_DRAIN_WAVES)))

namespace ehss {

struct RayState;

struct SolveArgs {
    int64_t n;
    int32_t power;
    int32_t cap;
    double w;      // pi R^2 / n
    double cx, cy, cz, radius;
    uint32_t* info;
    double* contrib;
    unsigned long long* counters;  // boxes, spheres, retests, pa boxes, pa tests, pa retests
    int32_t coop;                  // a wave with at most this many live rays finishes them cooperatively
    int32_t group;                 // lanes per ray in the queue drain (power of two 1..64)
    long long* times;              // optional per-ray completion time stamps (device wall clock)
    int32_t packets;               // variant 4: packet queries before queueing (1..15)
    int32_t pad_;
    ehss_trig::CompatTable trig;   // host-libm compatibility table of the Sobol angles (may be null)
    // Continuation (EHSS_FLAG_CONTINUABLE): the state of every ray that ends
    // unresolved is also appended here, for ehss_solve_extend (null: off).
    RayState* cont;
    uint32_t* cont_count;
    int32_t rep;                   // replicate of these arguments (queue entries carry it)
    int32_t pad2_;
};

// Words per replicate in the solve workspace: [0..5] kernel counters, [6] max
// bounces, [8..15] final Partial.
constexpr int kCounterWords = 16;

struct Partial {
    double s, c;
    long long escaped, unresolved, pa_hits, ehss_hits, grazing, pad;
};
static_assert(sizeof(Partial) == 8 * 8, "Partial occupies words 8..15 of a replicate's counters");

// Wave-coherent ray map.  For n = 2^m rays the 64 rays of wave w are the points of
// one elementary cell of the scrambled net: the top a_d output bits of every
// coordinate d are fixed (sum a_d = m - 6).  Those bits are GF(2)-linear in the
// Gray-code index g, so with T the invertible m x m matrix whose first m - 6 rows
// are the corresponding rows of the scrambled generator matrices,
// g = T^-1 (bits of w || bits of lane); col[k] is column k of T^-1.  The ray index
// (record position) is the inverse Gray code of g.
//
// The device holds the map composed with the scramble: q[k][d] = XOR over the
// set bits j of column k of T^-1 of the scrambled direction number sv[d][j]
// (for m < 6, of the Gray code's column k).  The Sobol words of the point that
// thread t = (w << 6) | lane traces are then
//   shift_d ^ XOR_{bit k of w} q[k][d] ^ XOR_{bit k of lane} q[m - 6 + k][d],
// the wave part on the scalar unit and the lane part from six columns, instead
// of 30 bit tests per coordinate on every lane.
struct RayMap {
    uint32_t q[32][4];
};

__device__ __forceinline__ uint32_t inverse_gray(uint32_t g) {
    g ^= g >> 1;
    g ^= g >> 2;
    g ^= g >> 4;
    g ^= g >> 8;
    g ^= g >> 16;
    return g;
}

// The map of a net of 2^m points composed with its scramble (host).
void ray_map_words(const ehss_sobol& sob, int m, RayMap& out);

typedef const __attribute__((address_space(4))) uint32_t* CMap;
__device__ __forceinline__ uint32_t map_word(const RayMap& m, int k, int d) { return m.q[k][d]; }
__device__ __forceinline__ uint32_t map_word(CMap m, int k, int d) { return m[4 * k + d]; }

// Scrambled direction numbers read through the constant address space (scalar loads).
typedef const __attribute__((address_space(4))) uint32_t* CSobol;

__device__ __forceinline__ uint32_t sob_sv(const ehss_sobol& s, int d, int j) { return s.sv[d][j]; }
__device__ __forceinline__ uint32_t sob_shift(const ehss_sobol& s, int d) { return s.shift[d]; }
__device__ __forceinline__ uint32_t sob_sv(CSobol s, int d, int j) { return s[d * 32 + j]; }
__device__ __forceinline__ uint32_t sob_shift(CSobol s, int d) { return s[128 + d]; }
// Per-lane scrambles (vector loads).
__device__ __forceinline__ uint32_t sob_sv(const ehss_sobol* s, int d, int j) { return s->sv[d][j]; }
__device__ __forceinline__ uint32_t sob_shift(const ehss_sobol* s, int d) { return s->shift[d]; }

// Cross-replicate schedule of a multi-replicate launch.  With the spatial rows of
// a partition shared by every replicate (their rank does not depend on the
// scramble), the point of replicate r whose partition cell is c has Gray index
// g = T2_r^-1 (c ^ cell_shift_r); its record slot is the slot the replicate's own
// wave-coherent map gives it, t = T1_r g.  A wave takes one cell c for 64
// replicates, so its rays share a cell 64 times smaller than a wave of one
// replicate's net would.
struct CrossMap {
    uint32_t inv2[32];   // columns of T2^-1
    uint32_t rows1[32];  // rows of T1 (slot of Gray index g: bit k = parity(rows1[k] & g))
    uint32_t cell_shift;
    uint32_t pad[3];
};

// The same schedule as linear maps of the cell index c, in tables laid out
// replicate-fastest so that a wave (64 replicates) reads each entry with one
// coalesced load: the 30-bit Sobol words of the point are
// q_d = off[d][r] ^ XOR_{bit k of c} col[k][d][r], and the slot bits (before the
// wave/lane rotation) are u = uoff[r] ^ XOR_{bit k of c} ucol[k][r].
struct CrossTables {
    const uint32_t* col;   // [m][4][rp]
    const uint32_t* off;   // [4][rp]
    const uint32_t* ucol;  // [m][rp]
    const uint32_t* uoff;  // [rp]
    int32_t rp;            // replicates padded to whole waves
    int32_t pad;
};

// A precomputed point (30-bit words per dimension), used in place of a scramble.
struct PointQ {
    uint32_t q[4];
};

// Sobol words of the point thread t traces (see RayMap).
template <class Map, class Sob>
__device__ __forceinline__ PointQ map_point(Map map, Sob sob, uint32_t t, int power) {
    PointQ p;
#pragma unroll
    for (int d = 0; d < 4; ++d) p.q[d] = sob_shift(sob, d);
    const int lb = power < 6 ? 0 : power - 6;
    const uint32_t w = __builtin_amdgcn_readfirstlane(t >> 6);
    for (int k = 0; k < lb; ++k)
        if ((w >> k) & 1u)
#pragma unroll
            for (int d = 0; d < 4; ++d) p.q[d] ^= map_word(map, k, d);
    const uint32_t lane = t & 63u;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const uint32_t m = 0u - ((lane >> k) & 1u);
#pragma unroll
        for (int d = 0; d < 4; ++d) p.q[d] ^= map_word(map, lb + k, d) & m;
    }
    return p;
}

__device__ __forceinline__ void sobol_point_gray(const PointQ& p, uint32_t, double u[4]) {
    const double scale = 9.313225746154785e-10;  // 2^-30
    u[0] = (double)p.q[0] * scale;
    u[1] = (double)p.q[1] * scale;
    u[2] = (double)p.q[2] * scale;
    u[3] = (double)p.q[3] * scale;
}

// Point with Gray-code index g: q = shift ^ XOR_{bit j of g} sv[j].
template <class Sob>
__device__ __forceinline__ void sobol_point_gray(Sob sob, uint32_t g, double u[4]) {
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
// in the order NumPy evaluates the expressions.  The angles are 2pi u for the
// Sobol coordinates u = q 2^-30; their cos/sin are the host libm's bits
// (ehss_trig::sincos_sobol), so the rays equal NumPy's bit for bit.
__device__ __forceinline__ void source_ray(const double u[4], double cx, double cy, double cz, double R,
                                           ehss_trig::CompatTable trig, V3& o, V3& d) {
    const double two_pi = 6.283185307179586;  // 2 * np.pi
    const double z = 1 - 2 * u[0];
    const double phi = two_pi * u[1];
    const double rho = sqrt(smax(0., 1 - z * z));
    double s, c;
    ehss_trig::sincos_sobol(phi, (uint32_t)(u[1] * 1073741824.0), trig, &s, &c);
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
    ehss_trig::sincos_sobol(ph, (uint32_t)(u[3] * 1073741824.0), trig, &sp, &cp);
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
// State of one ray between queries.
struct RayCtx {
    V3 o, d, inc;
    PathError pe;
    double tol;
    uint32_t i, flags;  // i: record slot (thread index t; ray index = inverse Gray of the map's g(t))
    int32_t last, count;
    bool done, escaped;
};

// Queue entry of a ray that is still bouncing (two-phase variant); `job` selects
// the batch job (0 for a single solve).
struct alignas(8) RayState {
    double o[3], d[3], inc[3];
    double ep, ed;
    uint32_t i, flags;
    int32_t last, count;
    int32_t job, pad;
};

// Grazing tolerance of the ray's next query.  The flag is sticky and no
// physical output depends on tol, so a ray already flagged queries with tol 0:
// its filter then prunes as tightly as on a short path.
__device__ __forceinline__ double ray_tol(const RayCtx& R, double scale) {
    return (R.flags & EHSS_RAY_GRAZING) ? 0. : path_tol(R.pe, scale);
}

// One query result applied to the ray: escape, cap (unresolved) or reflection
// (ehss_cpu.cpp:73-77), plus the grazing flag and the path error.
__device__ __forceinline__ void ray_process(const SceneView& S, int32_t cap, RayCtx& R, const Query& q) {
    if (grazing(q, R.tol)) R.flags |= EHSS_RAY_GRAZING;
    if (q.atom < 0) {
        R.escaped = true;
        R.done = true;
    } else if (R.count >= cap) {
        R.flags |= EHSS_RAY_UNRESOLVED;
        R.done = true;
    } else {
        const double4 c = S.sph[q.atom];
        const double cos_inc = reflect64(R.o, R.d, c.x, c.y, c.z, c.w, q.best);
        path_error_reflect(R.pe, q.best, cos_inc, c.w, S.scale);
        R.tol = ray_tol(R, S.scale);
        ++R.count;
        R.last = q.atom;
    }
}

// Queue entry of a ray (its state between queries).
__device__ __forceinline__ void ray_save(const RayCtx& R, int32_t job, RayState* q, uint32_t flags) {
    RayState st;
    st.o[0] = R.o.x;
    st.o[1] = R.o.y;
    st.o[2] = R.o.z;
    st.d[0] = R.d.x;
    st.d[1] = R.d.y;
    st.d[2] = R.d.z;
    st.inc[0] = R.inc.x;
    st.inc[1] = R.inc.y;
    st.inc[2] = R.inc.z;
    st.ep = R.pe.ep;
    st.ed = R.pe.ed;
    st.i = R.i;
    st.flags = flags;
    st.last = R.last;
    st.count = R.count;
    st.job = job;
    st.pad = 0;
    *q = st;
}

__device__ __forceinline__ void ray_save(const RayCtx& R, int32_t job, RayState* q) { ray_save(R, job, q, R.flags); }

// The ray's record (ehss_state.py:84-87): bounces and flags, w (1 - u.v) if escaped,
// stored at the ray's slot.  Slots follow thread order, a fixed bijection with
// ray indices (ehss_ray_map), so the fixed-order reduction over slots does not
// depend on scheduling; host readback permutes to ray-index order.
__device__ __forceinline__ void ray_record(const SolveArgs& A, RayCtx& R) {
    // A continuable solve keeps the state of an unresolved ray as it stood
    // before its last query (count == cap): continuing from there with a
    // larger cap repeats that query and then follows the path a full retrace
    // at the larger cap would follow, bit for bit.
    if (A.cont && (R.flags & EHSS_RAY_UNRESOLVED))
        ray_save(R, A.rep, A.cont + atomicAdd(A.cont_count, 1u), R.flags & ~(uint32_t)EHSS_RAY_UNRESOLVED);
    double contrib = 0.;
    if (R.escaped) {
        R.flags |= EHSS_RAY_ESCAPED;
        // clip(1 - sum(incoming * outgoing), 0, 2) * w
        double angle = 1 - ((R.inc.x * R.d.x + R.inc.y * R.d.y) + R.inc.z * R.d.z);
        angle = angle < 0. ? 0. : (angle > 2. ? 2. : angle);
        contrib = A.w * angle;
    }
    if (R.count > 0 || (R.flags & EHSS_RAY_UNRESOLVED)) R.flags |= EHSS_RAY_EHSS_HIT;
    A.info[R.i] = (uint32_t)R.count | (R.flags << kFlagShift);
    A.contrib[R.i] = contrib;
    if (A.times) A.times[R.i] = (long long)wall_clock64();
}

__device__ __forceinline__ int32_t ray_job(const RayState* q) { return q->job; }

__device__ __forceinline__ int32_t ray_load(const RayState* q, const SceneView& S, RayCtx& R) {
    const RayState st = *q;
    R.o = V3{st.o[0], st.o[1], st.o[2]};
    R.d = V3{st.d[0], st.d[1], st.d[2]};
    R.inc = V3{st.inc[0], st.inc[1], st.inc[2]};
    R.pe = PathError{st.ep, st.ed};
    R.tol = ray_tol(R, S.scale);
    R.i = st.i;
    R.flags = st.flags;
    R.last = st.last;
    R.count = st.count;
    R.done = false;
    R.escaped = false;
    return st.job;
}

// Tail threshold from ehss_solve / ehss_plan_create flags (bits 12..18).
constexpr int kDefaultCoop = 8;
inline int coop_from_flags(unsigned flags) {
    const int f = (int)((flags >> EHSS_COOP_SHIFT) & 0x7Fu);
    return f == 0 ? kDefaultCoop : f - 1;
}

// Drain shape from flags: lanes per ray of the first drain (bits 19..21,
// log2 + 1) and the queries a ray gets there before it moves to the second
// drain (bits 22..25, q + 1; 0 queries = one drain) with kStage2Group lanes per
// ray.  Defaults (field 0): from 2^18 rays, where the queue fills the chip,
// kAutoGroup (the drain picks its lanes per ray from its queue length,
// drain_group_for) for 6 queries, then the second drain; below, one drain with
// a wave per ray.  Lanes per ray of a grouped drain are otherwise a power of
// two in 2..64; 1 is the per-lane drain.
constexpr int kAutoGroup = 0;
#ifndef EHSS_DRAIN_GROUP
#define EHSS_DRAIN_GROUP 0
#endif
inline int group_from_flags(unsigned flags, int64_t rays) {
    const int f = (int)((flags >> EHSS_GROUP_SHIFT) & 0x7u);
    if (f == 0) return rays >= (1 << 18) ? EHSS_DRAIN_GROUP : 64;
    return 1 << (f - 1);
}

// Lanes per ray of the second drain stage (the long paths) and of the drain of
// continued rays (kAutoGroup: from the queue length).
#ifndef EHSS_STAGE2_GROUP
#define EHSS_STAGE2_GROUP 0
#endif
constexpr int kStage2Group = EHSS_STAGE2_GROUP;
static_assert(kStage2Group == kAutoGroup ||
                  (kStage2Group >= 2 && kStage2Group <= kWave && (kStage2Group & (kStage2Group - 1)) == 0),
              "second-stage lanes per ray: kAutoGroup or a power of two in 2..64");

// Lanes per ray of a drain left to choose (kAutoGroup): the fewest, at least 2,
// with which its `len` queued rays cover all `lanes` lanes of its grid.  A long
// queue is then drained for throughput (a two-pass grouped query costs less per
// ray with fewer lanes), a short one for latency (more lanes per ray shorten
// each query of the long paths that set its length).  Wave-uniform.
__device__ __forceinline__ int drain_group_for(uint32_t len, uint32_t lanes) {
    int G = 2;
    while (G < kWave && (uint64_t)len * (uint64_t)G < lanes) G <<= 1;
    return G;
}

// Packet queries of variant 4 (bits 26..29; 0 = default).  Measured with the
// front-to-back walk at 2^20 rays per replicate, 16 replicates of the 650-atom
// globule (6 first-stage drain queries): 2 / 3 / 4 / 5 queries take 9.96 /
// 9.66 / 9.87 / 10.14 ms per launch; the 2000-atom ellipsoid (4 x 2^22 rays,
// cap 128) keeps gaining up to 5 (15.39 / 12.65 / 11.58 / 11.39 ms), which the
// autotuner picks up for such sets.
constexpr int kDefaultPackets = 3;
inline int packets_from_flags(unsigned flags) {
    const int f = (int)((flags >> EHSS_PACKETS_SHIFT) & 0xFu);
    return f == 0 ? kDefaultPackets : f;
}

// Measured at 16 x 2^20 and 8192 x 2^11 rays of the 650-atom globule with 4
// lanes per ray: 1 / 2 / 4 / 6 / 8 / 12 first-stage queries give 12.40 / 12.10 /
// 11.89 / 11.83 / 12.02 / 11.91 ms per launch (and 11.43 ms at 6 for the small
// nets, against 11.75 at 2); with the front-to-back walk, 4 / 6 / 8 give 9.77 /
// 9.66 / 9.68 ms.
inline int stage1_queries(unsigned flags, int64_t rays) {
    const int f = (int)((flags >> EHSS_STAGE1_SHIFT) & 0xFu);
    if (f == 0) return rays >= (1 << 18) ? 6 : 0;
    return f - 1;
}

// Value of lane `src` (wave-uniform) on every lane.
__device__ __forceinline__ int32_t lane_value(int32_t x, int src) { return __builtin_amdgcn_readlane(x, src); }
__device__ __forceinline__ double lane_value(double x, int src) {
    const long long b = __double_as_longlong(x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, src);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), src);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Tail mode: the wave's live rays (lanes with !R.done) are finished one at a time,
// every query spread over the whole wave (nearest_coop).  A long path then costs
// one short cooperative query per bounce instead of one full per-lane traversal
// while 63 lanes idle.  All lanes of the wave call this; `job` is the lane's job,
// `account(job, boxes, spheres, retests)` receives each lane's share of a query's
// work.  Results are those of the per-lane queries.
template <class SceneOf, class ArgsOf, class Account>
__device__ __forceinline__ void finish_coop(RayCtx& R, int32_t job, SceneOf scene_of, ArgsOf args_of,
                                            Account account) {
    const int lane = threadIdx.x & (kWave - 1);
    for (;;) {
        const uint64_t m = __ballot(!R.done);
        if (m == 0) break;
        const int src = __builtin_amdgcn_readfirstlane(__builtin_ctzll(m));
        const int32_t j = lane_value(job, src);
        const SceneView& S = scene_of(j);
        const V3 o{lane_value(R.o.x, src), lane_value(R.o.y, src), lane_value(R.o.z, src)};
        const V3 d{lane_value(R.d.x, src), lane_value(R.d.y, src), lane_value(R.d.z, src)};
        uint32_t boxes = 0, spheres = 0, retests = 0;  // one query
        Query q;
        nearest_coop(S, o, d, lane_value(R.last, src), lane_value(R.tol, src), q, boxes, spheres, retests);
        account(j, boxes, spheres, retests);
        if (lane == src) ray_process(S, args_of(j).cap, R, q);
    }
}

// Number of wave-uniform queries a ray gets in the two-phase variant (4) before
// it is queued for the per-lane drain kernel.

// Kernel variants (template kMode):
//   0 plain FP64: reference traversal, separate reference PA traversal, reference counters;
//   1 FP32 filter, one ray per lane, PA fused into the first query;
//   2 FP32 filter, wave-uniform (packet) first query with fused PA, later bounces per lane;
//   3 FP32 filter, packet traversal for every bounce;
//   4 FP32 filter, packet traversal for the first A.packets queries, then the rays
//     still bouncing are compacted into a queue (wave ballot, one atomic per wave)
//     and finished by the persistent per-lane drain kernel.
// Variants 1-4 return bit-identical per-ray results (the filter only skips spheres
// whose exact FP64 outcome is already decided).
// The ray with Gray index g of the net, recorded at slot t (valid: the lane
// has a ray; every lane of the wave must call this).
template <int kMode, int kPA, int kLds = 0, class Sob>
__device__ __forceinline__ void solve_ray_at(const SceneView& S, const SolveArgs& A, Sob sob, bool valid, uint32_t g,
                                             uint32_t t, uint64_t* ctr, RayState* queue, uint32_t* queue_count,
                                             int32_t job) {
    // Work counts of this ray in this kernel.  Variant 4 runs a bounded number of
    // queries here (the packet queries and the PA test), so 32 bits hold them and
    // save registers; the full-path variants keep 64 bits.
    using Cnt = std::conditional_t<kMode == 4, uint32_t, uint64_t>;
    Cnt boxes = 0, spheres = 0, retests = 0, pboxes = 0, ptests = 0, pretests = 0;
    RayCtx R;
    R.i = t;
    R.o = V3{0., 0., 0.};
    R.d = V3{0., 0., 1.};
    if (valid) {
        double u[4];
        sobol_point_gray(sob, g, u);
        source_ray(u, A.cx, A.cy, A.cz, A.radius, A.trig, R.o, R.d);
    }
    R.inc = R.d;
    R.flags = 0;
    R.last = -1;
    R.count = 0;
    R.done = !valid;
    R.escaped = false;
    R.pe = path_error_start(S.scale);
    R.tol = ray_tol(R, S.scale);
    Query q;
    bool record = valid;
    if (kMode == 0) {
        if (valid) {
            if (kPA) {
                bool pgraze = false;
                if (pa_plain(S, R.o, R.d, -1, R.tol, pboxes, ptests, pgraze)) R.flags |= EHSS_RAY_PA_HIT;
                if (pgraze) R.flags |= EHSS_RAY_GRAZING;
            }
            while (!R.done) {
                nearest_plain(S, R.o, R.d, R.last, R.tol, q, boxes, spheres);
                ray_process(S, A.cap, R, q);
            }
        }
    } else {
        PAState pa = {false, false, false};
        PAState* pap = kPA ? &pa : nullptr;
        // Two-pass packet queries: a ray whose candidate list overflowed has no
        // answer from the query; it keeps its state, takes no further packet
        // query, and is queued for the drain, which repeats that query.
        bool deferred = false;
        if (kMode == 1) {
            if (valid) nearest_fast(S, R.o, R.d, -1, R.tol, q, boxes, spheres, retests, pap);
        } else {
            if constexpr (kMode == 4 && EHSS_ORDERED_PACKETS && EHSS_TWO_PASS_PACKETS)
                nearest_packet_2p<kLds>(S, R.o, R.d, -1, R.tol, valid, q, boxes, spheres, retests, deferred, pap);
            else if constexpr (kMode == 4 && EHSS_ORDERED_PACKETS)
                nearest_packet_ordered<kLds>(S, R.o, R.d, -1, R.tol, valid, q, boxes, spheres, retests, pap);
            else
                nearest_packet<kLds>(S, R.o, R.d, -1, R.tol, valid, q, boxes, spheres, retests, pap);
        }
        if (valid) {
            if (kPA) {
                bool hit = pa.hit, pgraze = false;
                if (!pa.decided) {
                    // A winner without a robust any-hit (or no answer from the
                    // query): run the full any-hit query.
                    if (deferred || q.atom >= 0)
                        hit = pa_fast(S, R.o, R.d, -1, R.tol, pboxes, ptests, pretests, pgraze);
                    else
                        pgraze = pa.critical;
                }
                if (hit) R.flags |= EHSS_RAY_PA_HIT;
                if (pgraze) R.flags |= EHSS_RAY_GRAZING;
            }
            if (!deferred) ray_process(S, A.cap, R, q);
        }
        auto scene_of = [&](int32_t) -> const SceneView& { return S; };
        auto args_of = [&](int32_t) -> const SolveArgs& { return A; };
        auto account = [&](int32_t, uint64_t b, uint64_t s, uint64_t r) {
            boxes += b;
            spheres += s;
            retests += r;
        };
        if (kMode == 3 || kMode == 4) {
            for (int k = 1; kMode == 3 || k < A.packets; ++k) {
                const bool live = !R.done && !deferred;
                const uint64_t m = __ballot(live);
                if (m == 0) break;
                if (kMode == 3 && __popcll(m) <= (unsigned)A.coop) {
                    finish_coop(R, 0, scene_of, args_of, account);
                    break;
                }
                bool overflowed = false;
                if constexpr (kMode == 4 && EHSS_ORDERED_PACKETS && EHSS_TWO_PASS_PACKETS)
                    nearest_packet_2p<kLds>(S, R.o, R.d, R.last, R.tol, live, q, boxes, spheres, retests,
                                            overflowed);
                else if constexpr (kMode == 4 && EHSS_ORDERED_PACKETS)
                    nearest_packet_ordered<kLds>(S, R.o, R.d, R.last, R.tol, live, q, boxes, spheres, retests);
                else
                    nearest_packet<kLds>(S, R.o, R.d, R.last, R.tol, live, q, boxes, spheres, retests);
                if (live) {
                    if (overflowed) deferred = true;
                    else ray_process(S, A.cap, R, q);
                }
            }
            if (kMode == 4) {
                const bool alive = !R.done;
                const uint64_t m = __ballot(alive);
                if (m) {
                    uint32_t base = 0;
                    if ((threadIdx.x & 63) == 0) base = atomicAdd(queue_count, (uint32_t)__popcll(m));
                    base = __shfl(base, 0, 64);
                    if (alive) {
                        const uint64_t below = m & ((1ull << (threadIdx.x & 63)) - 1ull);
                        ray_save(R, job, queue + base + (uint32_t)__popcll(below));
                        record = false;  // finished by the drain kernel
                    }
                }
            }
        } else {
            for (;;) {
                const uint64_t m = __ballot(!R.done);
                if (m == 0) break;
                if (__popcll(m) <= (unsigned)A.coop) {
                    finish_coop(R, 0, scene_of, args_of, account);
                    break;
                }
                if (!R.done) {
                    nearest_fast(S, R.o, R.d, R.last, R.tol, q, boxes, spheres, retests);
                    ray_process(S, A.cap, R, q);
                }
            }
        }
    }
    if (record) ray_record(A, R);
    ctr[0] += boxes;
    ctr[1] += spheres;
    ctr[2] += retests;
    ctr[3] += pboxes;
    ctr[4] += ptests;
    ctr[5] += pretests;
}

// Thread t of a solve traces the ray the wave-coherent map assigns to it and
// records it at slot t, so a wave's record writes are contiguous.
template <int kMode, int kPA, int kLds = 0, class Sob, class Map>
__device__ __forceinline__ void solve_ray(const SceneView& S, const SolveArgs& A, Sob sob, Map map, uint32_t t,
                                          uint64_t* ctr, RayState* queue = nullptr, uint32_t* queue_count = nullptr,
                                          int32_t job = 0) {
    const bool valid = t < (uint64_t)A.n;
    solve_ray_at<kMode, kPA, kLds>(S, A, map_point(map, sob, t, A.power), valid, 0u, t, ctr, queue, queue_count,
                                   job);
}

__host__ __device__ inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// LDS bytes of the arrays every staged view holds.
__host__ __device__ inline size_t staged_base_bytes(const SceneView& v) {
    return sizeof(Node32) * (v.n_nodes + v.n_groups) + sizeof(float4) * v.n_atoms +
           sizeof(int32_t) * (v.n_groups + 1) + sizeof(int32_t) * v.n_atoms;
}

// Copies the scene arrays the FP32 filter reads (tree, leaf boxes, leaf
// offsets, slot ids, FP32 spheres, and for kPacket the packet tree and its leaf
// offsets) into the block's LDS and returns the view pointing there.  All
// threads of the block call it.
template <bool kPacket = false>
__device__ __forceinline__ SceneView stage_scene(const SceneView& S, unsigned char* lds) {
    Node32* nodes = reinterpret_cast<Node32*>(lds);
    Node32* leaves = nodes + S.n_nodes;
    float4* sph = reinterpret_cast<float4*>(leaves + S.n_groups);
    int32_t* offsets = reinterpret_cast<int32_t*>(sph + S.n_atoms);
    int32_t* ids = offsets + (S.n_groups + 1);
    // Packet tree, when it is not the static tree itself.
    const bool own_pk = kPacket && S.pk_nodes32 != S.nodes32;
    Node32* pk = reinterpret_cast<Node32*>(lds + align_up(staged_base_bytes(S), alignof(Node32)));
    int32_t* pk_off = reinterpret_cast<int32_t*>(pk + S.pk_nodes);
    if (own_pk) {
        const int32_t leaves_pk = (S.pk_nodes + 1) / 2;
        for (int32_t k = threadIdx.x; k < S.pk_nodes; k += blockDim.x) pk[k] = S.pk_nodes32[k];
        for (int32_t k = threadIdx.x; k <= leaves_pk; k += blockDim.x) pk_off[k] = S.pk_offsets[k];
    }
    for (int32_t k = threadIdx.x; k < S.n_nodes; k += blockDim.x) nodes[k] = S.nodes32[k];
    for (int32_t k = threadIdx.x; k < S.n_groups; k += blockDim.x) leaves[k] = S.leaves32[k];
    for (int32_t k = threadIdx.x; k < S.n_atoms; k += blockDim.x) {
        sph[k] = S.sph32[k];
        ids[k] = S.ids[k];
    }
    for (int32_t k = threadIdx.x; k <= S.n_groups; k += blockDim.x) offsets[k] = S.offsets[k];
    __syncthreads();
    SceneView L = S;
    L.nodes32 = nodes;
    L.leaves32 = leaves;
    L.sph32 = sph;
    L.offsets = offsets;
    L.ids = ids;
    if (own_pk) {
        L.pk_nodes32 = pk;
        L.pk_offsets = pk_off;
    } else {
        L.pk_nodes32 = nodes;
        L.pk_offsets = offsets;
    }
    return L;
}

// LDS bytes stage_scene<kPacket> uses for a scene.
__host__ __device__ inline size_t staged_scene_bytes(const SceneView& v, bool packet = false) {
    if (!packet || v.pk_nodes32 == v.nodes32) return staged_base_bytes(v);
    return align_up(staged_base_bytes(v), alignof(Node32)) + sizeof(Node32) * v.pk_nodes +
           sizeof(int32_t) * ((v.pk_nodes + 1) / 2 + 1);
}

// Persistent per-lane drain of queued rays.  Idle lanes refill from the queue
// (one atomic per wave); each live lane runs one query per iteration, so lanes
// never wait for another lane's long path.  Once the queue is exhausted and at
// most `coop` rays are left in the wave, they are finished cooperatively.  Every
// wave exits once the queue is exhausted and its lanes are idle.
// `scene_of(job)` / `args_of(job)` give the scene and solve arguments of a queued
// ray's job; `account(job, boxes, spheres, retests)` receives the work of each
// finished ray (per lane) and of each cooperative query (each lane's share).
template <class SceneOf, class ArgsOf, class Account>
__device__ __forceinline__ void drain_queue(const RayState* queue, uint32_t queue_len, uint32_t* queue_head,
                                            int32_t coop, SceneOf scene_of, ArgsOf args_of, Account account) {
    RayCtx R;
    R.done = true;
    int32_t job = 0;
    bool have = false;
    const int lane = threadIdx.x & (kWave - 1);
    uint64_t boxes = 0, spheres = 0, retests = 0;  // of the lane's current ray
    bool exhausted = false;  // wave-uniform: the queue head has passed the end
    for (;;) {
        const uint64_t idle = __ballot(!have);
        if (idle && !exhausted) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(queue_head, (uint32_t)__popcll(idle));
            base = __builtin_amdgcn_readfirstlane(__shfl(base, 0, kWave));
            // One grab past the end at most per wave, so the head cannot wrap.
            if (base + (uint32_t)__popcll(idle) >= queue_len) exhausted = true;
            if (!have) {
                const uint32_t slot = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                if (slot < queue_len) {
                    job = ray_load(queue + slot, scene_of(ray_job(queue + slot)), R);
                    have = true;
                }
            }
        }
        const uint64_t live = __ballot(have);
        if (live == 0) break;
        if (exhausted && __popcll(live) <= (unsigned)coop) {
            // Nothing left to refill from: finish the wave's rays cooperatively.
            if (have) account(job, boxes, spheres, retests);
            finish_coop(R, job, scene_of, args_of, account);
            if (have) ray_record(args_of(job), R);
            break;
        }
        if (have) {
            const SceneView& S = scene_of(job);
            Query q;
            nearest_fast(S, R.o, R.d, R.last, R.tol, q, boxes, spheres, retests);
            const SolveArgs& A = args_of(job);
            ray_process(S, A.cap, R, q);
            if (R.done) {
                ray_record(A, R);
                account(job, boxes, spheres, retests);
                boxes = spheres = retests = 0;
                have = false;
            }
        }
    }
}

// Sum over the G lanes of an aligned group, on every lane of the group.
__device__ __forceinline__ uint64_t group_sum(uint64_t v, int G) {
    for (int off = G >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Grouped drain queries: the two-pass form (nearest_group_2p) unless built
// with EHSS_GROUP_TWO_PASS=0.
#ifndef EHSS_GROUP_TWO_PASS
#define EHSS_GROUP_TWO_PASS 1
#endif
constexpr bool kGroupTwoPass = EHSS_GROUP_TWO_PASS != 0;

// Queue drain with G lanes per ray (G >= 2).  Every lane of a group holds the
// group's ray state and applies the same query result to it, so the copies stay
// identical; each query is a nearest_group over the group's lanes.  An idle
// group takes the next queued ray (one atomic per wave per refill round), so a
// long path holds only its own group.  `account` receives each finished ray's
// work (summed over the group, on the group's first lane).
//
// With max_queries > 0 a ray that is still bouncing after that many queries in
// this drain is appended to `spill` (count in *spill_count) instead, for a
// second drain with more lanes per ray: the first drain keeps its lanes on the
// many short paths, the second gives the few long ones more lanes each.
template <class SceneOf, class ArgsOf, class Account>
__device__ __forceinline__ void drain_groups(const RayState* queue, uint32_t queue_len, uint32_t* queue_head,
                                             int G, SceneOf scene_of, ArgsOf args_of, Account account,
                                             int max_queries = 0, RayState* spill = nullptr,
                                             uint32_t* spill_count = nullptr) {
    const int lane = threadIdx.x & (kWave - 1);
    const int sub = lane & (G - 1);
    const int leader = lane - sub;
    RayCtx R;
    R.done = true;
    int32_t job = 0;
    bool have = false;       // the same on every lane of a group
    bool exhausted = false;  // wave-uniform: the queue head has passed the end
    int queries = 0;         // queries of the group's current ray in this drain
    uint64_t boxes = 0, spheres = 0, retests = 0;  // this lane's share of the current ray
    for (;;) {
        const uint64_t idle = __ballot(!have && sub == 0);  // one bit per idle group, at its first lane
        if (idle && !exhausted) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(queue_head, (uint32_t)__popcll(idle));
            base = __builtin_amdgcn_readfirstlane(__shfl(base, 0, kWave));
            // One grab past the end at most per wave, so the head cannot wrap.
            if (base + (uint32_t)__popcll(idle) >= queue_len) exhausted = true;
            const uint32_t mine = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
            const uint32_t slot = (uint32_t)__shfl((int)mine, leader, kWave);
            if (!have && slot < queue_len) {
                job = ray_load(queue + slot, scene_of(ray_job(queue + slot)), R);
                have = true;
                queries = 0;
            }
        }
        if (__ballot(have) == 0) break;
        if (have) {
            const SceneView& S = scene_of(job);
            Query q;
            if constexpr (kGroupTwoPass) nearest_group_2p(S, R.o, R.d, R.last, R.tol, G, sub, q, boxes, spheres, retests);
            else nearest_group(S, R.o, R.d, R.last, R.tol, G, sub, q, boxes, spheres, retests);
            const SolveArgs& A = args_of(job);
            ray_process(S, A.cap, R, q);
            if (!R.done && ++queries == max_queries) {
                // Long path: hand it to the second drain.
                boxes = group_sum(boxes, G);
                spheres = group_sum(spheres, G);
                retests = group_sum(retests, G);
                if (sub == 0) {
                    account(job, boxes, spheres, retests);
                    ray_save(R, job, spill + atomicAdd(spill_count, 1u));
                }
                boxes = spheres = retests = 0;
                have = false;
            } else if (R.done) {
                boxes = group_sum(boxes, G);
                spheres = group_sum(spheres, G);
                retests = group_sum(retests, G);
                if (sub == 0) {
                    ray_record(A, R);
                    account(job, boxes, spheres, retests);
                }
                boxes = spheres = retests = 0;
                have = false;
            }
        }
    }
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

// Traces the ray states a kernel queued (one replicate; words[0] = their count,
// written on the device, words[1..3] = head, second-stage count and head, all
// zero on entry) to their ends on stream s with the fused solve's drain
// schedule for `capacity` rays and `flags` (A.group lanes per ray for the
// first-stage queries, then kStage2Group lanes per ray), recording each ray at its
// slot through A.  Defined in solve.hip.
int drain_queued(ehss_scene* sc, const SolveArgs& A, const RayState* queue, uint32_t* words, int64_t capacity,
                 unsigned flags, hipStream_t s);

}  // namespace ehss
