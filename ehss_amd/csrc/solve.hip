// solve.hip -- single-scene fused solve: kernel launch, fixed-order reduction and
// the ehss_solve / ehss_scene_last_rays / ehss_source_rays entry points.  The
// per-ray body and the reduction bodies are shared with the batch path
// (ehss_solve.h).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "ehss_solve.h"

namespace ehss {

namespace {

// Counter stripes (see k_drain): kDrainStripes per replicate, kStripeWords words each.
constexpr int kDrainStripes = 64;
constexpr int kStripeWords = 8;

// Replicate r of a multi-replicate solve: its slices of the record, counter,
// histogram and ledger arrays (replicates are laid out one after another).
__device__ __forceinline__ SolveArgs replicate_args(const SolveArgs& A, int32_t r) {
    SolveArgs B = A;
    B.rep = r;
    B.info = A.info + (int64_t)r * A.n;
    B.contrib = A.contrib + (int64_t)r * A.n;
    B.counters = A.counters + (int64_t)r * kCounterWords;
    if (A.times) B.times = A.times + (int64_t)r * A.n;
    return B;
}

// One thread per (ray slot, replicate): blockIdx.y is the replicate, whose
// scrambled net and ray map are read (scalar loads) from device tables.
// kLds: the block first stages the scene's filter arrays in LDS (stage_scene).
template <int kMode, int kPA, int kLds>
__global__ void EHSS_SOLVE_KERNEL k_solve(SceneView S, SolveArgs A, const ehss_sobol* sobs, const RayMap* maps,
                                               RayState* queue, uint32_t* queue_count) {
    extern __shared__ __align__(16) unsigned char lds[];
    const SceneView V = kLds ? stage_scene<true>(S, lds) : S;
    const int32_t r = (int32_t)blockIdx.y;
    const SolveArgs B = replicate_args(A, r);
    uint64_t ctr[6] = {0, 0, 0, 0, 0, 0};
    solve_ray<kMode, kPA, kLds>(V, B, (CSobol)(sobs + r), (CMap)(maps + r), blockIdx.x * blockDim.x + threadIdx.x,
                                ctr, queue, queue_count, r);
    flush_counters(ctr, B.counters);
}

// Sorted schedule of launches of many replicates of small nets.  The shared
// partition gives every replicate exactly one point per cell, so its cells
// cannot be finer than one replicate's net (2^-11 of the unit cube at 2^11
// rays), however many replicates a launch holds.  Here the launch's rays are
// instead ordered by the 4-D cell of their Sobol points: the top kSortBits bits
// of each 30-bit coordinate word, interleaved from the top (Morton order), by a
// counting sort over the kSortBins cells.  The scatter writes each ray's Sobol
// words, replicate and slot bits at its sorted position, so wave w of the
// packet kernel reads rays 64 w .. 64 w + 63 with two coalesced loads; they
// come from about (64 / rays per cell) neighbouring cells of 2^-20 of the cube.
// Which rays share a wave does not change any ray's result (each is traced
// exactly), and records stay at their replicate's slots, so every replicate's
// totals are those of a separate solve.
constexpr int kSortBits = 5;
constexpr int64_t kSortBins = 1LL << (4 * kSortBits);
constexpr int kSortChunk = 1024;  // bins per block of the local scan (256 threads x 4)
constexpr int kSortBlocks = (int)(kSortBins / kSortChunk);
static_assert(kSortBlocks <= 1024, "top-level scan runs in one block of 1024 threads");

__device__ __forceinline__ uint32_t cell_key(const PointQ& p) {
    uint32_t key = 0;
    for (int j = 0; j < kSortBits; ++j)
        for (int d = 0; d < 4; ++d) key = (key << 1) | ((p.q[d] >> (29 - j)) & 1u);
    return key;
}

// Sobol words p and slot bits u of point c of replicate r under the cross
// tables.  The sort passes number the launch's rays replicate-fastest (ray g is
// point g / reps of replicate g mod reps): a wave's table loads are coalesced,
// and its points share a cell of the shared partition, so its bin atomics land
// close together.
__device__ __forceinline__ void sort_point(const CrossTables& X, int32_t reps, int32_t bits, uint32_t g, int32_t& r,
                                           PointQ& p, uint32_t& u) {
    r = (int32_t)(g % (uint32_t)reps);
    const uint32_t c = g / (uint32_t)reps;
    u = X.uoff[r];
    for (int d = 0; d < 4; ++d) p.q[d] = X.off[d * X.rp + r];
    for (int k = 0; k < bits; ++k) {
        if (!((c >> k) & 1u)) continue;
        u ^= X.ucol[(int64_t)k * X.rp + r];
        for (int d = 0; d < 4; ++d) p.q[d] ^= X.col[((int64_t)k * 4 + d) * X.rp + r];
    }
}

__global__ void __launch_bounds__(256) k_sort_count(CrossTables X, int32_t reps, int32_t bits, int64_t total,
                                                    uint32_t* bins) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    int32_t r;
    PointQ p;
    uint32_t u;
    sort_point(X, reps, bits, (uint32_t)g, r, p, u);
    atomicAdd(&bins[cell_key(p)], 1u);
}

// Exclusive scan of each chunk of kSortChunk bins in place; chunk totals to tops.
__global__ void __launch_bounds__(256) k_sort_scan_local(uint32_t* bins, uint32_t* tops) {
    __shared__ uint32_t s[256];
    const int t = threadIdx.x;
    uint4* q = reinterpret_cast<uint4*>(bins + (int64_t)blockIdx.x * kSortChunk) + t;
    const uint4 v = *q;
    const uint32_t sum = v.x + v.y + v.z + v.w;
    s[t] = sum;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const uint32_t a = t >= off ? s[t - off] : 0u;
        __syncthreads();
        s[t] += a;
        __syncthreads();
    }
    uint4 o;
    o.x = s[t] - sum;
    o.y = o.x + v.x;
    o.z = o.y + v.y;
    o.w = o.z + v.z;
    *q = o;
    if (t == 255) tops[blockIdx.x] = s[255];
}

// Exclusive scan of the kSortBlocks chunk totals (one block).
__global__ void __launch_bounds__(1024) k_sort_scan_top(uint32_t* tops) {
    __shared__ uint32_t s[1024];
    const int t = threadIdx.x;
    const uint32_t v = t < kSortBlocks ? tops[t] : 0u;
    s[t] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const uint32_t a = t >= off ? s[t - off] : 0u;
        __syncthreads();
        s[t] += a;
        __syncthreads();
    }
    if (t < kSortBlocks) tops[t] = s[t] - v;
}

// Ray g's words, replicate and slot bits at its sorted position (the order
// within a cell is whatever the atomics give).
__global__ void __launch_bounds__(256) k_sort_scatter(CrossTables X, int32_t reps, int32_t bits, int64_t total,
                                                      uint32_t* bins, const uint32_t* tops, uint4* sq, uint2* sru) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    int32_t r;
    PointQ p;
    uint32_t u;
    sort_point(X, reps, bits, (uint32_t)g, r, p, u);
    const uint32_t key = cell_key(p);
    const uint32_t at = tops[key / kSortChunk] + atomicAdd(&bins[key], 1u);
    sq[at] = make_uint4(p.q[0], p.q[1], p.q[2], p.q[3]);
    sru[at] = make_uint2((uint32_t)r, u);
}

// Cross-replicate schedule (see CrossMap / CrossTables): the replicates share
// one spatial partition of m - b bits, so each has exactly 2^b points in every
// cell.  Wave w takes cell c = w / groups for 64 / 2^b replicates (group
// w % groups): lane l traces point (c, l mod 2^b) of replicate
// (w % groups) 64 / 2^b + l / 2^b.  b = 0 for many replicates of small nets
// (one point of each of 64 replicates per cell); b = 6 - log2(reps) for a
// power-of-two count of at most 64 replicates, so that the wave's rays come
// from a cell 64 / 2^b times smaller than one replicate's own wave-coherent
// map gives.  Each ray is recorded at the slot its replicate's own map gives
// it, so every replicate's records, and therefore its totals, are those of a
// separate solve.  Work counters go to the ray's replicate, once per group of
// 2^b lanes.
template <int kMode, int kPA, int kLds>
__global__ void EHSS_SOLVE_KERNEL k_solve_x(SceneView S0, SolveArgs A, CrossTables X, int32_t reps,
                                                 int32_t groups, int32_t sub_bits, RayState* queue,
                                                 uint32_t* queue_count, unsigned long long* stripes,
                                                 const uint4* sq, const uint2* sru) {
    extern __shared__ __align__(16) unsigned char lds[];
    const SceneView S = kLds ? stage_scene<true>(S0, lds) : S0;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int m = A.power;
    int32_t r;
    uint32_t jsub = 0;
    bool valid;
    PointQ p;
    uint32_t u;
    if (sq) {
        // Sorted schedule (sub_bits 0): the words of the launch's i-th ray in
        // cell order (k_sort_scatter).
        const int64_t i = wave * 64 + lane;
        valid = i < ((int64_t)reps << m);
        const uint4 q = valid ? sq[i] : make_uint4(0u, 0u, 0u, 0u);
        const uint2 ru = valid ? sru[i] : make_uint2(0u, 0u);
        p.q[0] = q.x;
        p.q[1] = q.y;
        p.q[2] = q.z;
        p.q[3] = q.w;
        u = ru.y;
        r = valid ? (int32_t)ru.x : reps;
    } else {
        const uint32_t c = __builtin_amdgcn_readfirstlane((uint32_t)(wave / groups));
        r = (int32_t)(wave % groups) * (64 >> sub_bits) + (lane >> sub_bits);
        jsub = (uint32_t)lane & ((1u << sub_bits) - 1u);
        const int cell_bits = m - sub_bits;
        const int32_t r0 = r < reps ? r : 0;
        valid = r < reps && (int64_t)c < (A.n >> sub_bits);
        u = X.uoff[r0];
        for (int d = 0; d < 4; ++d) p.q[d] = X.off[d * X.rp + r0];
        for (int k = 0; k < cell_bits; ++k) {
            if (!((c >> k) & 1u)) continue;  // wave-uniform
            u ^= X.ucol[(int64_t)k * X.rp + r0];
            for (int d = 0; d < 4; ++d) p.q[d] ^= X.col[((int64_t)k * 4 + d) * X.rp + r0];
        }
        for (int k = 0; k < sub_bits; ++k) {
            if (!((jsub >> k) & 1u)) continue;
            const int64_t kk = cell_bits + k;
            u ^= X.ucol[kk * X.rp + r0];
            for (int d = 0; d < 4; ++d) p.q[d] ^= X.col[(kk * 4 + d) * X.rp + r0];
        }
    }
    const int32_t rr = r < reps ? r : 0;
    const SolveArgs B = replicate_args(A, rr);
    uint64_t ctr[6] = {0, 0, 0, 0, 0, 0};
    const uint32_t t = m >= 6 ? (((u & ((1u << (m - 6)) - 1u)) << 6) | (u >> (m - 6))) : u;
    solve_ray_at<kMode, kPA, kLds>(S, B, p, valid, 0u, t, ctr, queue, queue_count, rr);
    // Counters: summed over the lanes of a replicate, then added to one of its
    // stripes (picked by the wave): every wave of the launch holds the same
    // replicates, and one address per counter would take an atomic per wave.
    const int G = 1 << sub_bits;
    unsigned long long* cs = stripes + ((int64_t)rr * kDrainStripes + wave % kDrainStripes) * kStripeWords;
    for (int k = 0; k < 6; ++k) {
        const uint64_t v = G > 1 ? group_sum(ctr[k], G) : ctr[k];
        if (r < reps && jsub == 0 && v) atomicAdd(&cs[k], (unsigned long long)v);
    }
}

// Persistent drain of the rays the two-phase variant queued (every replicate's
// rays share one queue; a ray's job is its replicate): one lane per ray
// (per-lane tree traversal, cooperative tail) or A.group lanes per ray.  With
// kLds the FP32 filter tree, the leaf boxes, leaf offsets, slot ids and FP32
// spheres are first copied into LDS, so every node, leaf and sphere fetch of the
// drain is an LDS access instead of a dependent L2 round trip.
// Several replicates: the drain's work counters of replicate r go to
// stripes[(r * kDrainStripes + s) * kStripeWords + k], stripe s picked by the
// wave, so that the atomics of a launch spread over many addresses; k_finalize
// adds a replicate's stripes into its counters.  (The cross-replicate solve
// kernel adds its six counters the same way.)

template <int kLds>
__global__ void EHSS_DRAIN_KERNEL k_drain(SceneView S, SolveArgs A, int32_t replicates, const RayState* queue,
                                               const uint32_t* queue_count, uint32_t* queue_head, int32_t group,
                                               int32_t max_queries, RayState* spill, uint32_t* spill_count,
                                               unsigned long long* stripes) {
    extern __shared__ __align__(16) unsigned char lds[];
    uint64_t ctr[6] = {0, 0, 0, 0, 0, 0};
    const uint32_t len = __builtin_amdgcn_readfirstlane(*queue_count);
    if (len == 0) return;
    const SceneView L = kLds ? stage_scene(S, lds) : S;
    if (group == kAutoGroup) group = drain_group_for(len, gridDim.x * blockDim.x);
    auto scene_of = [&](int32_t) -> const SceneView& { return L; };
    auto args_of = [&](int32_t j) -> SolveArgs { return replicate_args(A, j); };
    // Per-lane sums.  One replicate: flushed once per wave at the end.  Several:
    // the lane's sums belong to replicate `cur` and are added to its counters
    // when the lane moves to a ray of another replicate (queued rays of one
    // replicate mostly come in runs), never one atomic per ray.
    int32_t cur = -1;
    const int64_t stripe = (int64_t)((blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave) % kDrainStripes);
    auto flush_lane = [&]() {
        if (cur < 0) return;
        unsigned long long* c = stripes + ((int64_t)cur * kDrainStripes + stripe) * kStripeWords;
        for (int k = 0; k < 3; ++k)
            if (ctr[k]) atomicAdd(&c[k], (unsigned long long)ctr[k]);
        ctr[0] = ctr[1] = ctr[2] = 0;
    };
    auto account = [&](int32_t j, uint64_t b, uint64_t sp, uint64_t r) {
        if (replicates > 1 && j != cur) {
            flush_lane();
            cur = j;
        }
        ctr[0] += b;
        ctr[1] += sp;
        ctr[2] += r;
    };
    if (group > 1)
        drain_groups(queue, len, queue_head, group, scene_of, args_of, account, max_queries, spill, spill_count);
    else
        drain_queue(queue, len, queue_head, A.coop, scene_of, args_of, account);
    if (replicates == 1) {
        flush_counters(ctr, A.counters);
        return;
    }
    // Remaining per-lane sums: one wave reduction and one atomic per distinct
    // replicate held by the wave's lanes (usually one or two).
    const int lane = threadIdx.x & (kWave - 1);
    for (;;) {
        const uint64_t pending = __ballot(cur >= 0);
        if (!pending) break;
        const int32_t j = __builtin_amdgcn_readlane(cur, __builtin_ctzll(pending));
        const bool mine = cur == j;
        unsigned long long* c = stripes + ((int64_t)j * kDrainStripes + stripe) * kStripeWords;
        for (int k = 0; k < 3; ++k) {
            const int64_t v = wave_sum_i64(mine ? (int64_t)ctr[k] : 0);
            if (lane == 0 && v) atomicAdd(&c[k], (unsigned long long)v);
        }
        if (mine) cur = -1;
    }
}

// LDS bytes of a scene staged by the kLds kernels.
size_t drain_lds_bytes(const SceneView& v) { return staged_scene_bytes(v); }

// Allocation-free solve: the scramble and ray map arrive as kernel arguments
// (captured by value under graph capture) and are written to the work array.
__global__ void k_put_tables(ehss_sobol sob, RayMap map, ehss_sobol* d_sob, RayMap* d_map) {
    const int i = threadIdx.x;
    if (i < 4 * 32) d_sob->sv[i / 32][i % 32] = sob.sv[i / 32][i % 32];
    if (i < 4) d_sob->shift[i] = sob.shift[i];
    if (i < 4 * 32) d_map->q[i / 4][i % 4] = map.q[i / 4][i % 4];
}

// Allocation-free solve: totals, histogram and ledger from the device counters,
// as solve_collect forms them on the host.
__global__ void k_write_totals(const unsigned long long* c, const unsigned long long* hist,
                               const unsigned long long* ledger, int64_t bins, int64_t n, double w, int pa,
                               ehss_totals* totals, int64_t* hist_out, double* ledger_out) {
    for (int64_t k = threadIdx.x; k < bins; k += blockDim.x) {
        if (hist_out) hist_out[k] = (int64_t)hist[k];
        if (ledger_out) ledger_out[k] = (double)ledger[k] / kLedgerScale * w;
    }
    if (threadIdx.x != 0) return;
    const Partial* fin = reinterpret_cast<const Partial*>(c + 8);
    ehss_totals T;
    memset(&T, 0, sizeof T);
    T.weight = w;
    T.omega = fin->s + fin->c;
    T.omega_err = fin->c;
    T.rays = n;
    T.escaped = fin->escaped;
    T.unresolved = fin->unresolved;
    T.pa_hits = pa ? fin->pa_hits : 0;
    T.ehss_hits = fin->ehss_hits;
    T.grazing_rays = fin->grazing;
    T.pa = pa ? w * (double)fin->pa_hits : 0.;
    T.pa_from_ehss = w * (double)fin->ehss_hits;
    T.residual_mass = w * (double)fin->unresolved;
    T.tail_bound = 2 * T.residual_mass;
    T.box_tests = (int64_t)c[0];
    T.sphere_tests = (int64_t)c[1];
    T.fp64_retests = (int64_t)(c[2] + c[5]);
    T.pa_box_tests = (int64_t)c[3];
    T.pa_sphere_tests = (int64_t)c[4];
    T.max_bounces = (int64_t)(int)c[6];
    T.kernel_ms = 0.;
    *totals = T;
}

// Continuation words {count, head} of queue `in` and `out`: the input keeps its
// count and restarts its head, the output starts empty.
__global__ void k_cont_reset(uint32_t* w, int in) {
    if (threadIdx.x == 0) {
        w[2 * in + 1] = 0u;
        w[2 * (1 - in)] = 0u;
        w[2 * (1 - in) + 1] = 0u;
    }
}

__global__ void k_zero(unsigned long long* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// Reduction block blockIdx.x of replicate blockIdx.y.
__global__ void __launch_bounds__(256) k_reduce(SolveArgs A, int64_t nblocks, Partial* partials,
                                                unsigned long long* hist, unsigned long long* ledger) {
    const int32_t r = (int32_t)blockIdx.y;
    const SolveArgs B = replicate_args(A, r);
    const int64_t bins = A.cap + 1;
    reduce_block(blockIdx.x, A.n, A.cap, A.w, B.info, B.contrib, partials + r * nblocks, hist + r * bins,
                 ledger + r * bins, reinterpret_cast<int*>(B.counters + 6));
}

// Final fixed-order combination of replicate blockIdx.x, and the sum of its
// drain counter stripes (if any) into its counters.
__global__ void __launch_bounds__(256) k_finalize(const Partial* partials, int64_t nblocks,
                                                  unsigned long long* counters, const unsigned long long* stripes) {
    const int32_t r = (int32_t)blockIdx.x;
    finalize_block(partials + r * nblocks, nblocks,
                   reinterpret_cast<Partial*>(counters + (int64_t)r * kCounterWords + 8));
    if (stripes && threadIdx.x < 6) {
        unsigned long long sum = 0;
        for (int s = 0; s < kDrainStripes; ++s)
            sum += stripes[((int64_t)r * kDrainStripes + s) * kStripeWords + threadIdx.x];
        counters[(int64_t)r * kCounterWords + threadIdx.x] += sum;
    }
}

__global__ void k_source_rays(int64_t n, double cx, double cy, double cz, double R, ehss_trig::CompatTable trig,
                              ehss_sobol sob, double* u_out, double* o_out, double* d_out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double u[4];
    const uint32_t ii = (uint32_t)i;
    sobol_point_gray(sob, ii ^ (ii >> 1), u);
    V3 o, d;
    source_ray(u, cx, cy, cz, R, trig, o, d);
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

// Builds the wave-coherent ray map (see RayMap).  Row p (MSB-first) of the
// scrambled generator matrix of dimension d, restricted to the first m index
// bits, is bit (29 - p) of sv[d][j] for j < m.  The cell widths a_d are chosen
// as balanced as possible (angles first) among partitions whose stacked rows are
// linearly independent; without one, the identity map (lanes = low index bits)
// is used.
bool map_for_partition(const ehss_sobol& sob, int m, const std::array<int, 4>& a, uint32_t cols[32],
                       uint32_t* rows) {
    for (int k = 0; k < 32; ++k) cols[k] = 0;
    auto row = [&](int d, int p) {
        uint32_t r = 0;
        for (int j = 0; j < m; ++j) r |= ((sob.sv[d][j] >> (29 - p)) & 1u) << j;
        return r;
    };
    // T rows: spatial rows first, then unit rows completing the basis.
    uint32_t T[32];
    int nr = 0;
    for (int d = 0; d < 4; ++d)
        for (int p = 0; p < a[d]; ++p) T[nr++] = row(d, p);
    // Gaussian elimination on a copy to test rank and find pivot columns.
    uint32_t E[32];
    for (int k = 0; k < nr; ++k) E[k] = T[k];
    uint32_t pivots = 0;
    int rank = 0;
    for (int c = 0; c < m && rank < nr; ++c) {
        int sel = -1;
        for (int k = rank; k < nr; ++k)
            if ((E[k] >> c) & 1u) {
                sel = k;
                break;
            }
        if (sel < 0) continue;
        std::swap(E[rank], E[sel]);
        for (int k = 0; k < nr; ++k)
            if (k != rank && ((E[k] >> c) & 1u)) E[k] ^= E[rank];
        pivots |= 1u << c;
        ++rank;
    }
    if (rank < nr) return false;
    for (int c = 0; c < m; ++c)
        if (!((pivots >> c) & 1u)) T[nr++] = 1u << c;  // completes to full rank
    // Invert T (m x m over GF(2)): [T | I] -> [I | T^-1]; rows as bit masks.
    uint32_t A[32], B[32];
    for (int k = 0; k < m; ++k) {
        A[k] = T[k];
        B[k] = 1u << k;
    }
    for (int c = 0; c < m; ++c) {
        int sel = -1;
        for (int k = c; k < m; ++k)
            if ((A[k] >> c) & 1u) {
                sel = k;
                break;
            }
        if (sel < 0) return false;
        std::swap(A[c], A[sel]);
        std::swap(B[c], B[sel]);
        for (int k = 0; k < m; ++k)
            if (k != c && ((A[k] >> c) & 1u)) {
                A[k] ^= A[c];
                B[k] ^= B[c];
            }
    }
    // Row c of T^-1 is B[c]; column k of T^-1 collects bit k of every row.
    for (int k = 0; k < m; ++k) {
        uint32_t col = 0;
        for (int c = 0; c < m; ++c) col |= ((B[c] >> k) & 1u) << c;
        cols[k] = col;
    }
    if (rows)
        for (int k = 0; k < 32; ++k) rows[k] = k < m ? T[k] : 0u;
    return true;
}

// GF(2) map of a net of 2^m points.  T's first rows are the spatial rows -- the
// top a_d output bits of dimension d, for the first partition (most balanced,
// angular dimensions favoured) with sum a_d = spatial whose rows are linearly
// independent -- completed to a basis by unit rows.  Writes the columns of T^-1
// (cols), the rows of T (rows, may be NULL) and the partition; false if no
// partition of that size has independent rows.  Whether the rows are
// independent, and which unit rows complete them, depend only on their row
// space, which the LMS scramble does not change (its top-left blocks are unit
// lower triangular): every scramble of the same net size gets the same
// partition.
// T for one partition; false if its spatial rows are dependent.
bool map_for_partition(const ehss_sobol& sob, int m, const std::array<int, 4>& a, uint32_t cols[32],
                       uint32_t* rows);

// fn(i) for i < n on up to 8 host threads when n is large enough to pay for
// them (the per-replicate ray maps and tables of launches with many replicates).
template <class F>
void parallel_for(int n, F fn) {
    const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
    const int threads = std::min({hw, 8, n / 128});
    if (threads <= 1) {
        for (int i = 0; i < n; ++i) fn(i);
        return;
    }
    std::atomic<int> next{0};
    auto work = [&] {
        for (int i; (i = next.fetch_add(32)) < n;)
            for (int k = i; k < std::min(n, i + 32); ++k) fn(k);
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
}

// Partition chosen for (m, spatial), found once on the first scramble and reused:
// it does not depend on the scramble.  State 0 unknown, 1 found, 2 none; the
// value is written before the state is published, so readers need no lock.
std::atomic<int> g_part_state[31][31];
std::array<int, 4> g_part_value[31][31];
std::mutex g_part_mutex;  // serialises the searches

void publish_partition(int m, int spatial, const std::array<int, 4>& a, int state) {
    std::lock_guard<std::mutex> lock(g_part_mutex);
    if (g_part_state[m][spatial].load(std::memory_order_acquire) != 0) return;
    g_part_value[m][spatial] = a;
    g_part_state[m][spatial].store(state, std::memory_order_release);
}

bool build_map(const ehss_sobol& sob, int m, int spatial, uint32_t cols[32], uint32_t* rows,
               std::array<int, 4>* part) {
    const int state = g_part_state[m][spatial].load(std::memory_order_acquire);
    if (state == 2) return false;
    if (state == 1) {
        const std::array<int, 4> cached = g_part_value[m][spatial];
        if (map_for_partition(sob, m, cached, cols, rows)) {
            if (part) *part = cached;
            return true;
        }
    }
    std::vector<std::array<int, 4>> parts;
    for (int a0 = 0; a0 <= spatial; ++a0)
        for (int a1 = 0; a0 + a1 <= spatial; ++a1)
            for (int a2 = 0; a0 + a1 + a2 <= spatial; ++a2) {
                const int a3 = spatial - a0 - a1 - a2;
                if (a0 > 30 || a1 > 30 || a2 > 30 || a3 > 30) continue;
                parts.push_back({a0, a1, a2, a3});
            }
    auto cost = [](const std::array<int, 4>& a) {
        const int mx = std::max({a[0], a[1], a[2], a[3]}), mn = std::min({a[0], a[1], a[2], a[3]});
        return (mx - mn) * 16 - (a[1] + a[3]);
    };
    std::stable_sort(parts.begin(), parts.end(),
                     [&](const std::array<int, 4>& x, const std::array<int, 4>& y) { return cost(x) < cost(y); });
    for (const auto& a : parts) {
        if (!map_for_partition(sob, m, a, cols, rows)) continue;
        if (part) *part = a;
        publish_partition(m, spatial, a, 1);
        return true;
    }
    publish_partition(m, spatial, {-1, -1, -1, -1}, 2);
    return false;
}

// Wave-coherent map of one net: a wave is one cell of m - 6 spatial bits and its
// lanes are the 64 points inside (identity-like fallback without such a cell).
void make_ray_map(const ehss_sobol& sob, int m, uint32_t cols[32], uint32_t* rows) {
    for (int k = 0; k < 32; ++k) cols[k] = 0;
    if (rows)
        for (int k = 0; k < 32; ++k) rows[k] = 0;
    if (m < 6) {
        for (int k = 0; k < m; ++k) {
            cols[k] = 1u << k;
            if (rows) rows[k] = 1u << k;
        }
        return;
    }
    if (build_map(sob, m, m - 6, cols, rows, nullptr)) return;
    const int cells = m - 6;
    for (int k = 0; k < m; ++k) {
        cols[k] = k < cells ? 1u << (k + 6) : 1u << (k - cells);
        if (rows) rows[k] = k < cells ? 1u << (k + 6) : 1u << (k - cells);  // a permutation: T = (T^-1)^T
    }
}

void make_ray_map(const ehss_sobol& sob, int m, uint32_t cols[32]) { make_ray_map(sob, m, cols, nullptr); }

void ray_map_words(const ehss_sobol& sob, int m, RayMap& out) {
    uint32_t cols[32];
    make_ray_map(sob, m, cols, nullptr);
    if (m < 6)  // the Gray code g = t ^ (t >> 1): column k holds bits k and k - 1
        for (int k = 0; k < m; ++k) cols[k] = (1u << k) | (k > 0 ? 1u << (k - 1) : 0u);
    for (int k = 0; k < 32; ++k)
        for (int d = 0; d < 4; ++d) {
            uint32_t v = 0;
            if (k < m)
                for (int j = 0; j < 30; ++j)
                    if ((cols[k] >> j) & 1u) v ^= sob.sv[d][j];
            out.q[k][d] = v;
        }
}

// Replicate-fastest tables of the cross schedule (CrossTables) for replicates
// padded to rp: words [m][4][rp] col, [4][rp] off, [m][rp] ucol, [rp] uoff.
size_t cross_table_words(int m, int rp) { return (size_t)rp * (4 * m + 4 + m + 1); }

void make_cross_tables(const ehss_sobol* sobs, const CrossMap* maps, int reps, int rp, int m, uint32_t* w) {
    memset(w, 0, sizeof(uint32_t) * cross_table_words(m, rp));
    uint32_t* col = w;
    uint32_t* off = col + (size_t)m * 4 * rp;
    uint32_t* ucol = off + (size_t)4 * rp;
    uint32_t* uoff = ucol + (size_t)m * rp;
    parallel_for(reps, [&](int r) {
        const CrossMap& X = maps[r];
        const ehss_sobol& sb = sobs[r];
        for (int k = 0; k < m; ++k) {
            const uint32_t g = X.inv2[k];  // Gray index of cell bit k
            for (int d = 0; d < 4; ++d) {
                uint32_t q = 0;
                for (uint32_t bits = g; bits; bits &= bits - 1u) q ^= sb.sv[d][__builtin_ctz(bits)];
                col[((size_t)k * 4 + d) * rp + r] = q;
            }
            uint32_t u = 0;
            for (int i = 0; i < m; ++i) u |= (uint32_t)(__builtin_popcount(X.rows1[i] & g) & 1) << i;
            ucol[(size_t)k * rp + r] = u;
        }
        for (int d = 0; d < 4; ++d) {
            uint32_t q = sb.shift[d];
            for (int k = 0; k < m; ++k)
                if ((X.cell_shift >> k) & 1u) q ^= col[((size_t)k * 4 + d) * rp + r];
            off[(size_t)d * rp + r] = q;
        }
        uint32_t u = 0;
        for (int k = 0; k < m; ++k)
            if ((X.cell_shift >> k) & 1u) u ^= ucol[(size_t)k * rp + r];
        uoff[r] = u;
    });
}

// Cross-replicate maps for `reps` scrambles of 2^m points: the finest shared
// spatial partition (m, m - 1, ... spatial bits, or exactly `only` bits when
// only >= 0) whose rows are independent.  False when none of at least m - 6
// bits exists or the replicates disagree.
bool make_cross_maps(const ehss_sobol* sobs, int reps, int m, CrossMap* out, int only = -1) {
    const int lo = only >= 0 ? only : std::max(1, m - 6);
    for (int spatial = only >= 0 ? only : m; spatial >= lo; --spatial) {
        // Replicate 0 fixes the partition; the others must share it.
        auto one = [&](int r, std::array<int, 4>* part) -> bool {
            CrossMap& X = out[r];
            memset(&X, 0, sizeof X);
            uint32_t rows2[32];
            if (!build_map(sobs[r], m, spatial, X.inv2, rows2, part)) return false;
            uint32_t cols1[32];
            make_ray_map(sobs[r], m, cols1, X.rows1);
            if (m < 6)  // the single-solve slot is the inverse Gray code of g
                for (int i = 0; i < 32; ++i) X.rows1[i] = i < m ? ((1u << m) - 1u) & ~((1u << i) - 1u) : 0u;
            // Digital-shift bits of the spatial rows, in T2's row order.
            int k = 0;
            for (int d = 0; d < 4; ++d)
                for (int p = 0; p < (*part)[d]; ++p, ++k)
                    X.cell_shift |= ((sobs[r].shift[d] >> (29 - p)) & 1u) << k;
            return true;
        };
        std::array<int, 4> part0{};
        if (!one(0, &part0)) continue;
        std::atomic<bool> all{true};
        parallel_for(reps - 1, [&](int i) {
            if (!all.load(std::memory_order_relaxed)) return;
            std::array<int, 4> part{};
            if (!one(i + 1, &part) || part != part0) all.store(false, std::memory_order_relaxed);
        });
        if (all.load()) return true;
    }
    return false;
}
// Persistent drain grid: enough waves to keep every SIMD busy, never more than
// the rays that could be queued.
// The CU count is read once per process (call_once: entry points may run on
// several threads at once; the library is bound to one device).
int64_t drain_blocks(int64_t n) {
    static std::once_flag once;
    static int cus = 256;
    std::call_once(once, [] {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
            prop.multiProcessorCount > 0)
            cus = prop.multiProcessorCount;
    });
    return std::max<int64_t>(1, std::min<int64_t>((int64_t)cus * 8, (n + 255) / 256));
}

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
constexpr int kDefaultVariant = 4;
constexpr int kMaxReplicates = 16384;
constexpr int kCrossMinReps = 16;  // replicates per launch from which small nets are scheduled across lanes
constexpr int kSortMinReps = 256;  // replicates per launch from which their rays are sorted by cell
constexpr int64_t kSortMaxRays = 1LL << 26;  // sorted launches hold 24 B per ray of sorted words
// Power-of-two replicate counts up to 64 scheduled across lanes with 2^b points
// of each per wave (k_solve_x); 0: each replicate on its own wave-coherent map.
#ifndef EHSS_MIXED_CELLS
#define EHSS_MIXED_CELLS 1
#endif
constexpr bool kMixedCells = EHSS_MIXED_CELLS;
#ifndef EHSS_DRAIN_LDS_KB
#define EHSS_DRAIN_LDS_KB 144
#endif
constexpr size_t kDrainLdsLimit = (size_t)EHSS_DRAIN_LDS_KB * 1024;  // of the 160 KB per CU
static_assert(kDrainLdsLimit + sizeof(GroupCandLists) <= 160 * 1024,
              "staged scene and the grouped queries' candidate rows exceed a workgroup's LDS");
// Packet kernels: a staged scene of at most this size keeps several blocks per
// CU and fits beside the kernels' static LDS in a workgroup's 64 KB.
constexpr size_t kSolveLdsLimit = 40 * 1024;
static_assert(kSolveLdsLimit + kPacketStaticLds <= 64 * 1024, "packet kernel LDS exceeds 64 KB per workgroup");

// Raises k_drain<1>'s dynamic LDS limit to kDrainLdsLimit (once per process;
// the library is bound to one device).
int set_drain_lds_limit() {
    static std::once_flag once;
    static hipError_t status = hipSuccess;
    std::call_once(once, [] {
        status = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_drain<1>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)kDrainLdsLimit);
    });
    EHSS_HIP(status);
    return EHSS_OK;
}

// Ray index of record slot t (host copy of the device mapping: Gray index from
// the ray map, then inverse Gray code).
uint32_t slot_ray(const uint32_t cols[32], int power, uint32_t t) {
    uint32_t g = 0;
    if (power < 6) {
        g = t ^ (t >> 1);
    } else {
        const uint32_t w = t >> 6, lane = t & 63u;
        for (int k = 0; k < power - 6; ++k)
            if ((w >> k) & 1u) g ^= cols[k];
        for (int k = 0; k < 6; ++k)
            if ((lane >> k) & 1u) g ^= cols[power - 6 + k];
    }
    g ^= g >> 1;
    g ^= g >> 2;
    g ^= g >> 4;
    g ^= g >> 8;
    g ^= g >> 16;
    return g;
}

// Enqueues `reps` solves of 2^power rays (one scramble each) as one launch on the
// scene's stream, with the copies of their totals, histograms and ledgers into
// the scene's pinned host buffer; solve_collect waits for them.  One launch per
// scene may be pending.
int solve_enqueue(ehss_scene* sc, int reps, const ehss_sobol* sobs, int power, int cap, int precision,
                  unsigned flags) {
    EHSS_TRY(ensure_init());  // the calling thread may not have selected the device
    EHSS_CHECK(sc && sobs, "null scene/sobol");
    std::lock_guard<std::mutex> guard(sc->lock);
    if (sc->pending) return set_error(EHSS_ERR_BUSY, "a solve is already pending on this scene");
    EHSS_CHECK(reps >= 1 && reps <= kMaxReplicates, "replicate count %d out of range 1..%d", reps, kMaxReplicates);
    EHSS_CHECK(power >= 0 && power <= 30, "power %d out of range 0..30", power);
    EHSS_CHECK((int64_t)reps << power <= (1LL << 30), "%d replicates of 2^%d rays exceed 2^30 rays per launch", reps,
               power);
    EHSS_CHECK(cap >= 1 && cap <= (int)kBounceMask - 1, "cap %d out of range 1..%u", cap, kBounceMask - 1);
    EHSS_CHECK(precision == EHSS_PRECISION_FP64 || precision == EHSS_PRECISION_FP32_FIXUP, "unknown precision %d",
               precision);
    hipStream_t s = sc->stream;
    const int64_t n = 1LL << power;
    const int64_t total = n * reps;
    const int64_t nblocks = (n + kReduceRays - 1) / kReduceRays;
    const int64_t bins = cap + 1;
    const int64_t words = (int64_t)kCounterWords * reps + 2;  // + queue and spill length / head
    EHSS_TRY(sc->ray_info.reserve(sizeof(uint32_t) * total));
    EHSS_TRY(sc->ray_contrib.reserve(sizeof(double) * total));
    EHSS_TRY(sc->partials.reserve(sizeof(Partial) * nblocks * reps));
    EHSS_TRY(sc->hist.reserve(sizeof(unsigned long long) * bins * reps));
    EHSS_TRY(sc->ledger.reserve(sizeof(unsigned long long) * bins * reps));
    EHSS_TRY(sc->counters.reserve(sizeof(unsigned long long) * words));
    const int32_t groups = (reps + 63) / 64;
    const int32_t rp = groups * 64;
    // Host/device tables: scrambles, cross maps and cross words (always
    // uploaded), then the per-replicate ray maps (uploaded when the replicates
    // are not scheduled across lanes, which is the only schedule that reads them).
    const size_t base_bytes = align_up((sizeof(ehss_sobol) + sizeof(CrossMap)) * reps +
                                           sizeof(uint32_t) * cross_table_words(power, rp),
                                       alignof(RayMap));
    const size_t table_bytes = base_bytes + sizeof(RayMap) * reps;
    EHSS_TRY(sc->tables.reserve(table_bytes));
    EHSS_TRY(sc->host_table.reserve(table_bytes));
    unsigned long long* ctr = sc->counters.as<unsigned long long>();
    EHSS_TRY(launch_zero(ctr, words, s));
    EHSS_TRY(launch_zero(sc->hist.as<unsigned long long>(), bins * reps, s));
    EHSS_TRY(launch_zero(sc->ledger.as<unsigned long long>(), bins * reps, s));

    // Scrambles and ray maps, staged in pinned memory (not reused while pending).
    ehss_sobol* h_sob = sc->host_table.as<ehss_sobol>();
    CrossMap* h_cross = reinterpret_cast<CrossMap*>(h_sob + reps);
    RayMap* h_map = reinterpret_cast<RayMap*>(reinterpret_cast<unsigned char*>(h_sob) + base_bytes);
    parallel_for(reps, [&](int r) { h_sob[r] = sobs[r]; });
    // Many replicates of small nets: one point of each of 64 replicates per cell
    // when a shared partition exists (measured on the 650-atom globule: box
    // tests per ray 598 -> 327 at 2^11 rays).  Two to 64 replicates (a power of
    // two): 2^(6 - log2 reps) points of each per wave, from cells 64 / 2^b times
    // smaller than one replicate's wave-coherent map gives (16 x 2^20 rays:
    // 164 -> 126 box tests per ray, 8.2 -> 6.9 ms per launch).  The work
    // counters of these launches go through the stripes: one address per
    // counter, taking an atomic from every wave, had made the same launch take
    // 12.1 ms.
    const bool cross_shape = (power <= 12 && reps >= kCrossMinReps) || (power <= 16 && reps >= 64);
    // A power-of-two count of 2..64 replicates: 2^b points of each per wave from
    // one cell of power - b spatial bits (b = 6 - log2 reps), when the
    // replicates share such a partition.
    int sub_bits = 0;
    bool cross = false;
    if (!(flags & EHSS_FLAG_PER_REPLICATE) && power >= 1) {
        const int lr = __builtin_ctz((unsigned)reps);
        if (kMixedCells && reps >= 2 && reps <= 64 && (reps & (reps - 1)) == 0 && power >= 6 - lr + 1) {
            sub_bits = 6 - lr;
            cross = make_cross_maps(sobs, reps, power, h_cross, power - sub_bits);
        }
        if (!cross) {
            sub_bits = 0;
            cross = cross_shape && make_cross_maps(sobs, reps, power, h_cross);
        }
    }
    const int32_t xgroups = cross ? (reps + (64 >> sub_bits) - 1) / (64 >> sub_bits) : 0;
    uint32_t* h_words = reinterpret_cast<uint32_t*>(h_cross + reps);
    if (cross) make_cross_tables(sobs, h_cross, reps, rp, power, h_words);
    else parallel_for(reps, [&](int r) { ray_map_words(sobs[r], power, h_map[r]); });
    EHSS_HIP(hipMemcpyAsync(sc->tables.p, h_sob, cross ? base_bytes : table_bytes, hipMemcpyHostToDevice, s));
    const ehss_sobol* d_sob = sc->tables.as<ehss_sobol>();
    const RayMap* d_map =
        reinterpret_cast<const RayMap*>(reinterpret_cast<const unsigned char*>(sc->tables.p) + base_bytes);
    const uint32_t* d_words = reinterpret_cast<const uint32_t*>(reinterpret_cast<const CrossMap*>(d_sob + reps) + reps);
    CrossTables XT;
    XT.rp = rp;
    XT.pad = 0;
    XT.col = d_words;
    XT.off = XT.col + (size_t)power * 4 * rp;
    XT.ucol = XT.off + (size_t)4 * rp;
    XT.uoff = XT.ucol + (size_t)power * rp;

    SolveArgs A;
    A.n = n;
    A.power = power;
    A.cap = cap;
    A.w = M_PI * sc->radius * sc->radius / (double)n;  // np.pi*radius*radius/n
    A.cx = sc->center[0];
    A.cy = sc->center[1];
    A.cz = sc->center[2];
    A.radius = sc->radius;
    A.info = sc->ray_info.as<uint32_t>();
    A.contrib = sc->ray_contrib.as<double>();
    A.counters = ctr;
    A.coop = coop_from_flags(flags);
    A.group = group_from_flags(flags, total);
    A.packets = packets_from_flags(flags);
    A.pad_ = 0;
    A.trig = trig_compat();
    A.cont = nullptr;
    A.cont_count = nullptr;
    A.rep = 0;
    A.pad2_ = 0;
    sc->cont_valid = false;  // this launch overwrites the records
    if (flags & EHSS_FLAG_CONTINUABLE) {
        EHSS_TRY(sc->cont_q[0].reserve(sizeof(RayState) * total));
        EHSS_TRY(sc->cont_words.reserve(4 * sizeof(uint32_t)));
        EHSS_TRY(launch_zero(sc->cont_words.as<unsigned long long>(), 2, s));
        A.cont = sc->cont_q[0].as<RayState>();
        A.cont_count = sc->cont_words.as<uint32_t>();
        sc->cont_which = 0;
    }
    A.times = nullptr;
    if (flags & EHSS_FLAG_TIMESTAMPS) {
        EHSS_TRY(sc->ray_times.reserve(sizeof(long long) * total));
        A.times = sc->ray_times.as<long long>();
    }
    EHSS_CHECK(A.coop <= kWave, "tail threshold %d out of range 0..64", A.coop);
    const bool pa = (flags & EHSS_FLAG_PA) != 0;
    const dim3 grid((unsigned)((n + 255) / 256), (unsigned)reps), block(256);
    int variant = (int)((flags >> EHSS_VARIANT_SHIFT) & 0xFu);
    if (precision == EHSS_PRECISION_FP64) variant = 0;
    else if (variant == 0) variant = kDefaultVariant;
    EHSS_CHECK(variant >= 0 && variant <= 4, "unknown kernel variant %d", variant);
    RayState* qbuf = nullptr;
    uint32_t* qcount = reinterpret_cast<uint32_t*>(ctr + (int64_t)kCounterWords * reps);
    RayState* spill = nullptr;
    uint32_t* scount = qcount + 2;
    const int stage1 = stage1_queries(flags, total);
    if (variant == 4) {
        EHSS_TRY(sc->queue.reserve(sizeof(RayState) * total));
        qbuf = sc->queue.as<RayState>();
        if (stage1 > 0 && A.group != 1) {
            EHSS_TRY(sc->queue2.reserve(sizeof(RayState) * total));
            spill = sc->queue2.as<RayState>();
        }
    }
    // Counter stripes of several replicates (cross kernel and drains).
    unsigned long long* stripes = nullptr;
    if (reps > 1) {
        const int64_t words = (int64_t)reps * kDrainStripes * kStripeWords;
        EHSS_TRY(sc->drain_stripes.reserve(sizeof(unsigned long long) * words));
        stripes = sc->drain_stripes.as<unsigned long long>();
        EHSS_TRY(launch_zero(stripes, words, s));
    }
    EHSS_HIP(hipEventRecord(sc->ev0, s));
    const int64_t cell_waves = n >> sub_bits;
    const dim3 xgrid((unsigned)((cell_waves * xgroups * 64 + 255) / 256));
    // Many replicates of small nets: order the launch's rays by cell (see
    // k_sort_count) when a wave would otherwise cover a cell coarser than
    // 2^-20 of the cube by a factor of four or more.
    const uint4* sq = nullptr;
    const uint2* sru = nullptr;
    if (cross && sub_bits == 0 && reps >= kSortMinReps && total <= kSortMaxRays && !(flags & EHSS_FLAG_UNSORTED)) {
        const size_t bin_bytes = sizeof(uint32_t) * (kSortBins + kSortBlocks);
        EHSS_TRY(sc->sort_bins.reserve(bin_bytes));
        EHSS_TRY(sc->sort_order.reserve((sizeof(uint4) + sizeof(uint2)) * total));
        uint32_t* bins = sc->sort_bins.as<uint32_t>();
        uint32_t* tops = bins + kSortBins;
        uint4* q = sc->sort_order.as<uint4>();
        uint2* ru = reinterpret_cast<uint2*>(q + total);
        const unsigned tb = (unsigned)((total + 255) / 256);
        EHSS_TRY(launch_zero(sc->sort_bins.as<unsigned long long>(), (int64_t)(bin_bytes / 8), s));
        hipLaunchKernelGGL(k_sort_count, dim3(tb), dim3(256), 0, s, XT, reps, power, total, bins);
        hipLaunchKernelGGL(k_sort_scan_local, dim3(kSortBlocks), dim3(256), 0, s, bins, tops);
        hipLaunchKernelGGL(k_sort_scan_top, dim3(1), dim3(1024), 0, s, tops);
        hipLaunchKernelGGL(k_sort_scatter, dim3(tb), dim3(256), 0, s, XT, reps, power, total, bins, tops, q, ru);
        EHSS_HIP(hipGetLastError());
        sq = q;
        sru = ru;
    }
    // Packet variants stage the scene in LDS on request (EHSS_FLAG_LDS_PACKETS).
    const size_t stage_bytes = staged_scene_bytes(sc->view, true);
    const bool lds_packets = (variant >= 2) && stage_bytes <= kSolveLdsLimit && (flags & EHSS_FLAG_LDS_PACKETS);
    const size_t solve_lds = lds_packets ? stage_bytes : 0;
#define EHSS_LAUNCH_L(M, P, K)                                                                               \
    do {                                                                                                     \
        if (cross)                                                                                           \
            hipLaunchKernelGGL((k_solve_x<M, P, K>), xgrid, block, solve_lds, s, sc->view, A, XT, reps, xgroups,  \
                               sub_bits, qbuf, qcount, stripes, sq, sru);                                    \
        else                                                                                                 \
            hipLaunchKernelGGL((k_solve<M, P, K>), grid, block, solve_lds, s, sc->view, A, d_sob, d_map, qbuf,    \
                               qcount);                                                                      \
    } while (0)
#define EHSS_LAUNCH(M)                                                                                       \
    do {                                                                                                     \
        if (lds_packets) {                                                                                   \
            if (pa) EHSS_LAUNCH_L(M, 1, 1);                                                                  \
            else EHSS_LAUNCH_L(M, 0, 1);                                                                     \
        } else {                                                                                             \
            if (pa) EHSS_LAUNCH_L(M, 1, 0);                                                                  \
            else EHSS_LAUNCH_L(M, 0, 0);                                                                     \
        }                                                                                                    \
    } while (0)
    switch (variant) {
        case 0: EHSS_LAUNCH(0); break;
        case 1: EHSS_LAUNCH(1); break;
        case 2: EHSS_LAUNCH(2); break;
        case 3: EHSS_LAUNCH(3); break;
        default: EHSS_LAUNCH(4); break;
    }
#undef EHSS_LAUNCH
#undef EHSS_LAUNCH_L
    if (variant == 4) {
        const size_t lds = drain_lds_bytes(sc->view);
        const dim3 dgrid((unsigned)drain_blocks(total));
        if (lds <= kDrainLdsLimit) EHSS_TRY(set_drain_lds_limit());
        // First drain: A.group lanes per ray; with a spill buffer, rays still
        // bouncing after stage1 queries move to a second drain with
        // kStage2Group lanes per ray.
        auto drain = [&](const RayState* in, const uint32_t* count, uint32_t* head, int group, int max_q,
                         RayState* out, uint32_t* out_count) {
            if (lds <= kDrainLdsLimit)
                hipLaunchKernelGGL(k_drain<1>, dgrid, dim3(256), lds, s, sc->view, A, reps, in, count, head, group,
                                   max_q, out, out_count, stripes);
            else
                hipLaunchKernelGGL(k_drain<0>, dgrid, dim3(256), 0, s, sc->view, A, reps, in, count, head, group,
                                   max_q, out, out_count, stripes);
        };
        drain(qbuf, qcount, qcount + 1, A.group, spill ? stage1 : 0, spill, scount);
        if (spill) drain(spill, scount, scount + 1, kStage2Group, 0, nullptr, nullptr);
    }
    EHSS_HIP(hipGetLastError());
    EHSS_HIP(hipEventRecord(sc->ev1, s));
    hipLaunchKernelGGL(k_reduce, dim3((unsigned)nblocks, (unsigned)reps), dim3(256), 0, s, A, nblocks,
                       sc->partials.as<Partial>(), sc->hist.as<unsigned long long>(),
                       sc->ledger.as<unsigned long long>());
    hipLaunchKernelGGL(k_finalize, dim3((unsigned)reps), dim3(256), 0, s, sc->partials.as<Partial>(), nblocks, ctr,
                       (const unsigned long long*)stripes);
    EHSS_HIP(hipGetLastError());
    const int64_t host_words = (int64_t)kCounterWords * reps + 2 * bins * reps;
    EHSS_TRY(sc->host_words.reserve(sizeof(unsigned long long) * host_words));
    unsigned long long* hw = sc->host_words.as<unsigned long long>();
    EHSS_HIP(hipMemcpyAsync(hw, ctr, sizeof(unsigned long long) * kCounterWords * reps, hipMemcpyDeviceToHost, s));
    EHSS_HIP(hipMemcpyAsync(hw + (int64_t)kCounterWords * reps, sc->hist.p, sizeof(unsigned long long) * bins * reps,
                            hipMemcpyDeviceToHost, s));
    EHSS_HIP(hipMemcpyAsync(hw + (int64_t)kCounterWords * reps + bins * reps, sc->ledger.p,
                            sizeof(unsigned long long) * bins * reps, hipMemcpyDeviceToHost, s));
    EHSS_HIP(hipEventRecord(sc->ev_done, s));
    sc->pending = true;
    make_ray_map(sobs[0], power, sc->p_map);
    sc->p_power = power;
    sc->p_reps = reps;
    sc->p_rays = n;
    sc->p_cap = cap;
    sc->p_flags = flags;
    sc->p_weight = A.w;
    return EHSS_OK;
}

// Continues the scene's last continuable solve at new_cap (ehss_solve_extend):
// the saved unresolved rays are drained with kStage2Group lanes per ray (they are
// the long, trapped paths), rays unresolved again are saved into the other
// queue, and every replicate's records are reduced again.  Enqueued like a
// solve; solve_collect waits for it.
int extend_enqueue(ehss_scene* sc, int new_cap) {
    EHSS_TRY(ensure_init());
    EHSS_CHECK(sc, "null scene");
    std::lock_guard<std::mutex> guard(sc->lock);
    if (sc->pending) return set_error(EHSS_ERR_BUSY, "a solve is already pending on this scene");
    EHSS_CHECK(sc->cont_valid,
               "no continuable solve on this scene (EHSS_FLAG_CONTINUABLE, and no other solve or rebuild since)");
    EHSS_CHECK(new_cap > sc->cont_cap && new_cap <= (int)kBounceMask - 1, "new cap %d must be in %d..%u", new_cap,
               sc->cont_cap + 1, kBounceMask - 1);
    hipStream_t s = sc->stream;
    const int reps = sc->cont_reps, power = sc->cont_power;
    const int64_t n = 1LL << power;
    const int64_t total = n * reps;
    const int64_t nblocks = (n + kReduceRays - 1) / kReduceRays;
    const int64_t bins = new_cap + 1;
    const int64_t words = (int64_t)kCounterWords * reps + 2;
    const int in = sc->cont_which, out = 1 - in;
    EHSS_TRY(sc->cont_q[out].reserve(sizeof(RayState) * total));
    EHSS_TRY(sc->hist.reserve(sizeof(unsigned long long) * bins * reps));
    EHSS_TRY(sc->ledger.reserve(sizeof(unsigned long long) * bins * reps));
    unsigned long long* ctr = sc->counters.as<unsigned long long>();
    EHSS_TRY(launch_zero(ctr, words, s));
    EHSS_TRY(launch_zero(sc->hist.as<unsigned long long>(), bins * reps, s));
    EHSS_TRY(launch_zero(sc->ledger.as<unsigned long long>(), bins * reps, s));
    uint32_t* cw = sc->cont_words.as<uint32_t>();
    hipLaunchKernelGGL(k_cont_reset, dim3(1), dim3(64), 0, s, cw, in);

    SolveArgs A;
    A.n = n;
    A.power = power;
    A.cap = new_cap;
    A.w = M_PI * sc->radius * sc->radius / (double)n;  // np.pi*radius*radius/n
    A.cx = sc->center[0];
    A.cy = sc->center[1];
    A.cz = sc->center[2];
    A.radius = sc->radius;
    A.info = sc->ray_info.as<uint32_t>();
    A.contrib = sc->ray_contrib.as<double>();
    A.counters = ctr;
    A.coop = coop_from_flags(sc->cont_flags);
    A.group = kStage2Group;
    A.packets = packets_from_flags(sc->cont_flags);
    A.pad_ = 0;
    A.trig = trig_compat();
    A.cont = sc->cont_q[out].as<RayState>();
    A.cont_count = cw + 2 * out;
    A.rep = 0;
    A.pad2_ = 0;
    A.times = nullptr;
    unsigned long long* stripes = nullptr;
    if (reps > 1) {
        const int64_t sw = (int64_t)reps * kDrainStripes * kStripeWords;
        EHSS_TRY(sc->drain_stripes.reserve(sizeof(unsigned long long) * sw));
        stripes = sc->drain_stripes.as<unsigned long long>();
        EHSS_TRY(launch_zero(stripes, sw, s));
    }
    EHSS_HIP(hipEventRecord(sc->ev0, s));
    const size_t lds = drain_lds_bytes(sc->view);
    const dim3 dgrid((unsigned)drain_blocks(total));
    const RayState* q = sc->cont_q[in].as<RayState>();
    if (lds <= kDrainLdsLimit) {
        EHSS_TRY(set_drain_lds_limit());
        hipLaunchKernelGGL(k_drain<1>, dgrid, dim3(256), lds, s, sc->view, A, reps, q, cw + 2 * in, cw + 2 * in + 1,
                           kStage2Group, 0, nullptr, nullptr, stripes);
    } else {
        hipLaunchKernelGGL(k_drain<0>, dgrid, dim3(256), 0, s, sc->view, A, reps, q, cw + 2 * in, cw + 2 * in + 1,
                           kStage2Group, 0, nullptr, nullptr, stripes);
    }
    EHSS_HIP(hipGetLastError());
    EHSS_HIP(hipEventRecord(sc->ev1, s));
    hipLaunchKernelGGL(k_reduce, dim3((unsigned)nblocks, (unsigned)reps), dim3(256), 0, s, A, nblocks,
                       sc->partials.as<Partial>(), sc->hist.as<unsigned long long>(),
                       sc->ledger.as<unsigned long long>());
    hipLaunchKernelGGL(k_finalize, dim3((unsigned)reps), dim3(256), 0, s, sc->partials.as<Partial>(), nblocks, ctr,
                       (const unsigned long long*)stripes);
    EHSS_HIP(hipGetLastError());
    const int64_t host_words = (int64_t)kCounterWords * reps + 2 * bins * reps;
    EHSS_TRY(sc->host_words.reserve(sizeof(unsigned long long) * host_words));
    unsigned long long* hw = sc->host_words.as<unsigned long long>();
    EHSS_HIP(hipMemcpyAsync(hw, ctr, sizeof(unsigned long long) * kCounterWords * reps, hipMemcpyDeviceToHost, s));
    EHSS_HIP(hipMemcpyAsync(hw + (int64_t)kCounterWords * reps, sc->hist.p, sizeof(unsigned long long) * bins * reps,
                            hipMemcpyDeviceToHost, s));
    EHSS_HIP(hipMemcpyAsync(hw + (int64_t)kCounterWords * reps + bins * reps, sc->ledger.p,
                            sizeof(unsigned long long) * bins * reps, hipMemcpyDeviceToHost, s));
    EHSS_HIP(hipEventRecord(sc->ev_done, s));
    sc->pending = true;
    sc->cont_which = out;
    sc->p_power = power;
    sc->p_reps = reps;
    sc->p_rays = n;
    sc->p_cap = new_cap;
    sc->p_flags = sc->cont_flags;
    sc->p_weight = A.w;
    return EHSS_OK;
}

int solve_query(ehss_scene* sc, int* done) {
    EHSS_CHECK(sc && done, "null scene/done");
    std::lock_guard<std::mutex> guard(sc->lock);
    if (!sc->pending) return set_error(EHSS_ERR_INVALID, "no solve pending on this scene");
    const hipError_t e = hipEventQuery(sc->ev_done);
    if (e == hipErrorNotReady) {
        *done = 0;
        return EHSS_OK;
    }
    EHSS_HIP(e);
    *done = 1;
    return EHSS_OK;
}

// Waits for the pending launch and fills one ehss_totals per replicate (and, if
// not NULL, cap + 1 histogram and ledger bins per replicate, replicate-major).
int solve_collect(ehss_scene* sc, ehss_totals* totals, int64_t* histogram, double* ledger_out) {
    EHSS_CHECK(sc && totals, "null scene/totals");
    std::lock_guard<std::mutex> guard(sc->lock);
    if (!sc->pending) return set_error(EHSS_ERR_INVALID, "no solve pending on this scene");
    sc->pending = false;
    EHSS_HIP(hipEventSynchronize(sc->ev_done));
    float ms = 0.f;
    EHSS_HIP(hipEventElapsedTime(&ms, sc->ev0, sc->ev1));
    const unsigned long long* hw = sc->host_words.as<unsigned long long>();
    const int reps = sc->p_reps;
    const int64_t bins = sc->p_cap + 1;
    const int64_t n = sc->p_rays;
    const double w = sc->p_weight;
    const bool pa = (sc->p_flags & EHSS_FLAG_PA) != 0;
    for (int r = 0; r < reps; ++r) {
        const unsigned long long* c = hw + (int64_t)kCounterWords * r;
        Partial fin;
        memcpy(&fin, c + 8, sizeof fin);
        ehss_totals& T = totals[r];
        memset(&T, 0, sizeof T);
        T.weight = w;
        T.omega = fin.s + fin.c;
        T.omega_err = fin.c;
        T.rays = n;
        T.escaped = fin.escaped;
        T.unresolved = fin.unresolved;
        T.pa_hits = pa ? fin.pa_hits : 0;
        T.ehss_hits = fin.ehss_hits;
        T.grazing_rays = fin.grazing;
        T.pa = pa ? w * (double)fin.pa_hits : 0.;
        T.pa_from_ehss = w * (double)fin.ehss_hits;
        T.residual_mass = w * (double)fin.unresolved;
        T.tail_bound = 2 * T.residual_mass;
        T.box_tests = (int64_t)c[0];
        T.sphere_tests = (int64_t)c[1];
        T.fp64_retests = (int64_t)(c[2] + c[5]);
        T.pa_box_tests = (int64_t)c[3];
        T.pa_sphere_tests = (int64_t)c[4];
        T.max_bounces = (int64_t)(int)c[6];
        T.kernel_ms = ms;  // the whole launch
        const unsigned long long* h_hist = hw + (int64_t)kCounterWords * reps + bins * r;
        const unsigned long long* h_led = hw + (int64_t)kCounterWords * reps + bins * reps + bins * r;
        if (histogram)
            for (int64_t k = 0; k < bins; ++k) histogram[bins * r + k] = (int64_t)h_hist[k];
        if (ledger_out)
            for (int64_t k = 0; k < bins; ++k) ledger_out[bins * r + k] = (double)h_led[k] / kLedgerScale * w;
    }
    if (sc->p_flags & EHSS_FLAG_CONTINUABLE) {
        sc->cont_valid = true;
        sc->cont_cap = sc->p_cap;
        sc->cont_reps = reps;
        sc->cont_power = sc->p_power;
        sc->cont_flags = sc->p_flags;
    }
    sc->last_rays = (sc->p_flags & EHSS_FLAG_KEEP_RAYS) ? n : 0;
    sc->last_times = (sc->p_flags & EHSS_FLAG_TIMESTAMPS) ? n : 0;
    sc->last_weight = w;
    memcpy(sc->last_map, sc->p_map, sizeof sc->last_map);
    sc->last_power = sc->p_power;
    return EHSS_OK;
}

}  // namespace

namespace ehss {

int drain_queued(ehss_scene* sc, const SolveArgs& A, const RayState* queue, uint32_t* words, int64_t capacity,
                 unsigned flags, hipStream_t s) {
    const int stage1 = stage1_queries(flags, capacity);
    RayState* spill = nullptr;
    if (stage1 > 0 && A.group != 1) {
        EHSS_TRY(sc->queue2.reserve(sizeof(RayState) * capacity));
        spill = sc->queue2.as<RayState>();
    }
    const size_t lds = drain_lds_bytes(sc->view);
    const dim3 dgrid((unsigned)drain_blocks(capacity));
    if (lds <= kDrainLdsLimit) EHSS_TRY(set_drain_lds_limit());
    auto drain = [&](const RayState* in, const uint32_t* count, uint32_t* head, int group, int max_q, RayState* out,
                     uint32_t* out_count) {
        if (lds <= kDrainLdsLimit)
            hipLaunchKernelGGL(k_drain<1>, dgrid, dim3(256), lds, s, sc->view, A, 1, in, count, head, group, max_q,
                               out, out_count, nullptr);
        else
            hipLaunchKernelGGL(k_drain<0>, dgrid, dim3(256), 0, s, sc->view, A, 1, in, count, head, group, max_q,
                               out, out_count, nullptr);
    };
    drain(queue, words, words + 1, A.group, spill ? stage1 : 0, spill, words + 2);
    if (spill) drain(spill, words + 2, words + 3, kStage2Group, 0, nullptr, nullptr);
    EHSS_HIP(hipGetLastError());
    return EHSS_OK;
}

}  // namespace ehss

extern "C" {

int ehss_cross_map(const ehss_sobol* sobols, int reps, int power, int spatial, uint32_t* out) {
    EHSS_CHECK(sobols && out, "null sobols/out");
    EHSS_CHECK(reps >= 1 && reps <= kMaxReplicates, "replicate count %d out of range 1..%d", reps, kMaxReplicates);
    EHSS_CHECK(power >= 1 && power <= 30, "power %d out of range 1..30", power);
    EHSS_CHECK(spatial >= 0 && spatial <= power, "spatial bits %d out of range 0..%d", spatial, power);
    std::vector<CrossMap> maps((size_t)reps);
    if (!make_cross_maps(sobols, reps, power, maps.data(), spatial > 0 ? spatial : -1))
        return set_error(EHSS_ERR_INVALID, "no spatial partition shared by the replicates");
    memcpy(out, maps.data(), sizeof(CrossMap) * reps);
    return EHSS_OK;
}

int ehss_ray_map(const ehss_sobol* sobol, int power, uint32_t* cols) {
    EHSS_CHECK(sobol && cols, "null sobol/cols");
    EHSS_CHECK(power >= 0 && power <= 30, "power %d out of range 0..30", power);
    make_ray_map(*sobol, power, cols);
    return EHSS_OK;
}

int ehss_solve(ehss_scene* scene, const ehss_sobol* sobol, int power, int cap, int precision, unsigned flags,
               ehss_totals* totals, int64_t* histogram, double* ledger) {
    EHSS_CHECK(totals, "null totals");
    EHSS_TRY(solve_enqueue(scene, 1, sobol, power, cap, precision, flags));
    return solve_collect(scene, totals, histogram, ledger);
}

// Bytes the allocation-free solve carves at its limits (2^EHSS_NOALLOC_MAX_POWER
// rays, cap EHSS_NOALLOC_MAX_CAP): counters, histogram, ledger, partials, the
// scramble and ray map, records, the drain queue; each piece 256-byte aligned.
constexpr size_t noalloc_piece(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kNoallocMaxRays = (size_t)1 << EHSS_NOALLOC_MAX_POWER;
constexpr size_t kNoallocMaxBytes =
    noalloc_piece(8 * (kCounterWords + 2)) + 2 * noalloc_piece(8 * (EHSS_NOALLOC_MAX_CAP + 1)) +
    noalloc_piece(sizeof(Partial) * ((kNoallocMaxRays + kReduceRays - 1) / kReduceRays)) +
    noalloc_piece(sizeof(ehss_sobol)) + noalloc_piece(sizeof(RayMap)) + noalloc_piece(4 * kNoallocMaxRays) +
    noalloc_piece(8 * kNoallocMaxRays) + noalloc_piece(sizeof(RayState) * kNoallocMaxRays);
static_assert(kNoallocMaxBytes <= kNoallocScratchBytes, "the allocation-free work array is too small");

int ehss_solve_noalloc(ehss_scene* sc, const ehss_sobol* sobol, int power, int cap, int precision, unsigned flags,
                       void* stream, ehss_totals* totals, int64_t* histogram, double* ledger) {
    EHSS_TRY(ensure_init());  // the calling thread may not have selected the device
    EHSS_CHECK(sc && sobol && totals, "null scene/sobol/totals");
    std::lock_guard<std::mutex> guard(sc->lock);
    EHSS_CHECK(power >= 0 && power <= EHSS_NOALLOC_MAX_POWER, "power %d out of range 0..%d of the allocation-free solve",
               power, EHSS_NOALLOC_MAX_POWER);
    EHSS_CHECK(cap >= 1 && cap <= EHSS_NOALLOC_MAX_CAP, "cap %d out of range 1..%d of the allocation-free solve", cap,
               EHSS_NOALLOC_MAX_CAP);
    EHSS_CHECK(precision == EHSS_PRECISION_FP64 || precision == EHSS_PRECISION_FP32_FIXUP, "unknown precision %d",
               precision);
    EHSS_CHECK(!(flags & (EHSS_FLAG_KEEP_RAYS | EHSS_FLAG_TIMESTAMPS | EHSS_FLAG_CONTINUABLE)),
               "the allocation-free solve keeps no per-ray records (KEEP_RAYS / TIMESTAMPS / CONTINUABLE)");
    const int64_t n = 1LL << power;
    int variant = (int)((flags >> EHSS_VARIANT_SHIFT) & 0xFu);
    if (precision == EHSS_PRECISION_FP64) variant = 0;
    else if (variant == 0) variant = kDefaultVariant;
    EHSS_CHECK(variant >= 0 && variant <= 4, "unknown kernel variant %d", variant);
    const int group = group_from_flags(flags, n);
    const int stage1 = stage1_queries(flags, n);
    EHSS_CHECK(!(variant == 4 && stage1 > 0 && group != 1),
               "the allocation-free solve has no second drain queue (EHSS_STAGE1 with EHSS_GROUP_LOG2 > 0)");
    unsigned char* base = static_cast<unsigned char*>(noalloc_scratch());
    if (!base) return set_error(EHSS_ERR_HIP, "the allocation-free work array has no address on this device");

    // Carve the work array (256-byte aligned pieces).
    const int64_t bins = cap + 1;
    const int64_t nblocks = (n + kReduceRays - 1) / kReduceRays;
    const int64_t words = (int64_t)kCounterWords + 2;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        unsigned char* p = base + off;
        off += (bytes + 255) & ~(size_t)255;
        return p;
    };
    auto* ctr = reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long) * words));
    auto* hist = reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long) * bins));
    auto* led = reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long) * bins));
    auto* partials = reinterpret_cast<Partial*>(take(sizeof(Partial) * nblocks));
    auto* d_sob = reinterpret_cast<ehss_sobol*>(take(sizeof(ehss_sobol)));
    auto* d_map = reinterpret_cast<RayMap*>(take(sizeof(RayMap)));
    auto* info = reinterpret_cast<uint32_t*>(take(sizeof(uint32_t) * n));
    auto* contrib = reinterpret_cast<double*>(take(sizeof(double) * n));
    RayState* qbuf = variant == 4 ? reinterpret_cast<RayState*>(take(sizeof(RayState) * n)) : nullptr;
    // The largest carve (2^EHSS_NOALLOC_MAX_POWER rays, cap EHSS_NOALLOC_MAX_CAP) fits: checked per call.
    if (off > kNoallocScratchBytes)
        return set_error(EHSS_ERR_INVALID, "allocation-free solve needs %zu bytes of its %zu-byte work array", off,
                         kNoallocScratchBytes);

    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : sc->stream;
    if (s != sc->stream) EHSS_HIP(hipStreamWaitEvent(s, sc->ev_built, 0));  // the scene's last (re)build
    EHSS_TRY(launch_zero(ctr, words, s));
    EHSS_TRY(launch_zero(hist, bins, s));
    EHSS_TRY(launch_zero(led, bins, s));
    RayMap map;
    ray_map_words(*sobol, power, map);
    hipLaunchKernelGGL(k_put_tables, dim3(1), dim3(128), 0, s, *sobol, map, d_sob, d_map);

    SolveArgs A;
    A.n = n;
    A.power = power;
    A.cap = cap;
    A.w = M_PI * sc->radius * sc->radius / (double)n;  // np.pi*radius*radius/n
    A.cx = sc->center[0];
    A.cy = sc->center[1];
    A.cz = sc->center[2];
    A.radius = sc->radius;
    A.info = info;
    A.contrib = contrib;
    A.counters = ctr;
    A.coop = coop_from_flags(flags);
    A.group = group;
    A.packets = packets_from_flags(flags);
    A.pad_ = 0;
    A.trig = trig_compat();
    A.cont = nullptr;
    A.cont_count = nullptr;
    A.rep = 0;
    A.pad2_ = 0;
    A.times = nullptr;
    EHSS_CHECK(A.coop <= kWave, "tail threshold %d out of range 0..64", A.coop);
    const bool pa = (flags & EHSS_FLAG_PA) != 0;
    uint32_t* qcount = reinterpret_cast<uint32_t*>(ctr + kCounterWords);
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    const size_t stage_bytes = staged_scene_bytes(sc->view, true);
    const bool lds_packets = (variant >= 2) && stage_bytes <= kSolveLdsLimit && (flags & EHSS_FLAG_LDS_PACKETS);
    const size_t solve_lds = lds_packets ? stage_bytes : 0;
#define EHSS_LAUNCH_L(M, P, K)                                                                                 \
    hipLaunchKernelGGL((k_solve<M, P, K>), grid, block, solve_lds, s, sc->view, A, d_sob, d_map, qbuf, qcount)
#define EHSS_LAUNCH(M)                          \
    do {                                        \
        if (lds_packets) {                      \
            if (pa) EHSS_LAUNCH_L(M, 1, 1);     \
            else EHSS_LAUNCH_L(M, 0, 1);        \
        } else {                                \
            if (pa) EHSS_LAUNCH_L(M, 1, 0);     \
            else EHSS_LAUNCH_L(M, 0, 0);        \
        }                                       \
    } while (0)
    switch (variant) {
        case 0: EHSS_LAUNCH(0); break;
        case 1: EHSS_LAUNCH(1); break;
        case 2: EHSS_LAUNCH(2); break;
        case 3: EHSS_LAUNCH(3); break;
        default: EHSS_LAUNCH(4); break;
    }
#undef EHSS_LAUNCH
#undef EHSS_LAUNCH_L
    if (variant == 4) {
        const size_t lds = drain_lds_bytes(sc->view);
        const dim3 dgrid((unsigned)drain_blocks(n));
        if (lds <= kDrainLdsLimit) {
            EHSS_TRY(set_drain_lds_limit());
            hipLaunchKernelGGL(k_drain<1>, dgrid, dim3(256), lds, s, sc->view, A, 1, qbuf, qcount, qcount + 1, group,
                               0, nullptr, nullptr, nullptr);
        } else {
            hipLaunchKernelGGL(k_drain<0>, dgrid, dim3(256), 0, s, sc->view, A, 1, qbuf, qcount, qcount + 1, group, 0,
                               nullptr, nullptr, nullptr);
        }
    }
    hipLaunchKernelGGL(k_reduce, dim3((unsigned)nblocks, 1u), dim3(256), 0, s, A, nblocks, partials, hist, led);
    hipLaunchKernelGGL(k_finalize, dim3(1u), dim3(256), 0, s, partials, nblocks, ctr,
                       (const unsigned long long*)nullptr);
    hipLaunchKernelGGL(k_write_totals, dim3(1), dim3(256), 0, s, (const unsigned long long*)ctr,
                       (const unsigned long long*)hist, (const unsigned long long*)led, bins, n, A.w, (int)pa, totals,
                       histogram, ledger);
    EHSS_HIP(hipGetLastError());
    return EHSS_OK;
}

int ehss_solve_many(ehss_scene* scene, int replicates, const ehss_sobol* sobols, int power, int cap, int precision,
                    unsigned flags, ehss_totals* totals, int64_t* histograms, double* ledgers) {
    EHSS_CHECK(totals, "null totals");
    EHSS_TRY(solve_enqueue(scene, replicates, sobols, power, cap, precision, flags));
    return solve_collect(scene, totals, histograms, ledgers);
}

int ehss_solve_extend(ehss_scene* scene, int new_cap, ehss_totals* totals, int64_t* histograms, double* ledgers) {
    EHSS_CHECK(totals, "null totals");
    EHSS_TRY(extend_enqueue(scene, new_cap));
    return solve_collect(scene, totals, histograms, ledgers);
}

int ehss_solve_async(ehss_scene* scene, const ehss_sobol* sobol, int power, int cap, int precision, unsigned flags) {
    return solve_enqueue(scene, 1, sobol, power, cap, precision, flags);
}

int ehss_solve_many_async(ehss_scene* scene, int replicates, const ehss_sobol* sobols, int power, int cap,
                          int precision, unsigned flags) {
    return solve_enqueue(scene, replicates, sobols, power, cap, precision, flags);
}

int ehss_solve_query(ehss_scene* scene, int* done) { return solve_query(scene, done); }

int ehss_solve_wait(ehss_scene* scene, ehss_totals* totals, int64_t* histogram, double* ledger) {
    return solve_collect(scene, totals, histogram, ledger);
}

int ehss_scene_last_ray_times(ehss_scene* scene, int64_t n, int64_t* ticks, double* tick_hz) {
    EHSS_CHECK(scene && ticks, "null scene/ticks");
    std::lock_guard<std::mutex> guard(scene->lock);
    if (scene->pending) return set_error(EHSS_ERR_BUSY, "a solve is pending on this scene");
    EHSS_CHECK(scene->last_times > 0, "no time stamps kept: solve with EHSS_FLAG_TIMESTAMPS first");
    EHSS_CHECK(n == scene->last_times, "record count %lld != %lld rays of the last solve", (long long)n,
               (long long)scene->last_times);
    hipStream_t s = scene->stream;
    std::vector<int64_t> slots((size_t)n);
    EHSS_HIP(hipMemcpyAsync(slots.data(), scene->ray_times.p, sizeof(int64_t) * n, hipMemcpyDeviceToHost, s));
    EHSS_HIP(hipStreamSynchronize(s));
    for (int64_t t = 0; t < n; ++t) ticks[slot_ray(scene->last_map, scene->last_power, (uint32_t)t)] = slots[t];
    if (tick_hz) {
        int dev = 0, khz = 0;
        EHSS_HIP(hipGetDevice(&dev));
        EHSS_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        *tick_hz = 1e3 * (double)khz;
    }
    return EHSS_OK;
}

int ehss_scene_last_rays(ehss_scene* scene, int64_t n, uint32_t* flags, double* contrib, int32_t* bounces) {
    EHSS_CHECK(scene, "null scene");
    std::lock_guard<std::mutex> guard(scene->lock);
    if (scene->pending) return set_error(EHSS_ERR_BUSY, "a solve is pending on this scene");
    EHSS_CHECK(scene->last_rays > 0, "no per-ray records kept: solve with EHSS_FLAG_KEEP_RAYS first");
    EHSS_CHECK(n == scene->last_rays, "record count %lld != %lld rays of the last solve", (long long)n,
               (long long)scene->last_rays);
    hipStream_t s = scene->stream;
    std::vector<uint32_t> info((size_t)n);
    std::vector<double> ang(contrib ? (size_t)n : 0);
    EHSS_HIP(hipMemcpyAsync(info.data(), scene->ray_info.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
    if (contrib)
        EHSS_HIP(hipMemcpyAsync(ang.data(), scene->ray_contrib.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    EHSS_HIP(hipStreamSynchronize(s));
    for (int64_t t = 0; t < n; ++t) {  // slot order -> ray-index order
        const uint32_t i = slot_ray(scene->last_map, scene->last_power, (uint32_t)t);
        if (flags) flags[i] = info[t] >> kFlagShift;
        if (bounces) bounces[i] = (int32_t)(info[t] & kBounceMask);
        if (contrib) contrib[i] = ang[t];
    }
    return EHSS_OK;
}

int ehss_source_rays(ehss_scene* scene, const ehss_sobol* sobol, int power, double* u, double* origins,
                     double* directions) {
    EHSS_TRY(ensure_init());
    EHSS_CHECK(scene && sobol, "null scene/sobol");
    EHSS_CHECK(power >= 0 && power <= 26, "power %d out of range 0..26", power);
    std::lock_guard<std::mutex> guard(scene->lock);
    hipStream_t s = scene->stream;
    const int64_t n = 1LL << power;
    DevBuf bu, bo, bd;
    if (u) EHSS_TRY(bu.alloc(sizeof(double) * 4 * n));
    if (origins) EHSS_TRY(bo.alloc(sizeof(double) * 3 * n));
    if (directions) EHSS_TRY(bd.alloc(sizeof(double) * 3 * n));
    hipLaunchKernelGGL(k_source_rays, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, scene->center[0],
                       scene->center[1], scene->center[2], scene->radius, trig_compat(), *sobol, bu.as<double>(),
                       bo.as<double>(),
                       bd.as<double>());
    EHSS_HIP(hipGetLastError());
    if (u) EHSS_HIP(hipMemcpyAsync(u, bu.p, sizeof(double) * 4 * n, hipMemcpyDeviceToHost, s));
    if (origins) EHSS_HIP(hipMemcpyAsync(origins, bo.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
    if (directions) EHSS_HIP(hipMemcpyAsync(directions, bd.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
    EHSS_HIP(hipStreamSynchronize(s));
    return EHSS_OK;
}

}  // extern "C"
