// multi.hip -- prepared multi-structure launches (ehss_plan_*): many fused solves
// of different scenes, scrambles, ray counts and caps in one launch sequence,
// optionally captured once into a hipGraph and replayed.
//
// Grid layout: the packet kernel runs blocks (x, j) for job j = blockIdx.y and
// its 256-ray block x; blocks past a job's ray count exit at once.  Every ray is
// recorded at its job's slot; the rays still bouncing after the packet queries
// go to their job's own queue segment, and the drain kernels run blocks (x, j)
// over job j's segment, so the waves of a block work on one scene.  (Staging
// that scene in LDS, as the single-scene drain does, measured slower here:
// 28.9 ms against 26.4 ms per launch of 18 globules of 200-5000 atoms at 2^20
// rays each.)  Each job is reduced in its own fixed order by blocks (x, j) of
// the reduction kernel.  A job's totals are therefore bit-identical to
// ehss_solve on its scene alone.  The job table (scene views, arguments,
// scrambles, ray maps, queue segments) lives in device memory and is
// re-uploaded before a launch when a scramble changed, so a captured graph
// follows new scrambles.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "ehss_solve.h"

namespace ehss {

struct PlanJob {
    SceneView view;
    SolveArgs args;
    ehss_sobol sob;
    RayMap map;
    int64_t red_blocks;  // reduction blocks of this job
    Partial* partials;
    unsigned long long* hist;
    unsigned long long* ledger;
    RayState* queue;  // rays still bouncing after the packet queries (n slots)
    RayState* spill;  // rays still bouncing after the first drain stage (n slots, or null)
    uint32_t* qwords;  // [0] queue length, [1] queue head, [2] spill length, [3] spill head
};

namespace {

// kLds: the block first stages its job's filter arrays in LDS (stage_scene).
template <int kMode, int kPA, int kLds>
__global__ void EHSS_SOLVE_KERNEL k_plan_solve(const PlanJob* __restrict__ jobs) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int32_t j = (int32_t)blockIdx.y;
    const PlanJob& J = jobs[j];
    const int64_t n = J.args.n;
    if ((int64_t)blockIdx.x * blockDim.x >= n) return;  // whole block: past this job's rays
    const SceneView S = kLds ? stage_scene<true>(J.view, lds) : J.view;
    const SolveArgs A = J.args;
    uint64_t ctr[6] = {0, 0, 0, 0, 0, 0};
    solve_ray<kMode, kPA, kLds>(S, A, (CSobol)&J.sob, (CMap)&J.map, blockIdx.x * blockDim.x + threadIdx.x, ctr,
                                J.queue, J.qwords, j);
    flush_counters(ctr, A.counters);
}

// Drain stage `stage` (0: the packet kernel's queue, optionally spilling rays
// still bouncing after max_queries into the spill segment; 1: the spill
// segment) of job blockIdx.y, by blocks x.  A wave whose job's segment is
// already taken exits at once.
__global__ void EHSS_PLAN_DRAIN_KERNEL k_plan_drain(const PlanJob* __restrict__ jobs, int32_t stage, int32_t group,
                                                    int32_t max_queries) {
    const PlanJob& J = jobs[blockIdx.y];
    uint32_t* w = J.qwords + 2 * stage;
    const uint32_t len = __builtin_amdgcn_readfirstlane(w[0]);
    if (len == 0) return;
    const uint32_t head =
        __builtin_amdgcn_readfirstlane(__hip_atomic_load(&w[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    if (head >= len) return;
    const SceneView S = J.view;
    const SolveArgs A = J.args;
    if (group == kAutoGroup) group = drain_group_for(len, gridDim.x * blockDim.x);
    uint64_t ctr[6] = {0, 0, 0, 0, 0, 0};
    auto scene_of = [&](int32_t) -> const SceneView& { return S; };
    auto args_of = [&](int32_t) -> const SolveArgs& { return A; };
    auto account = [&](int32_t, uint64_t b, uint64_t sp, uint64_t r) {
        ctr[0] += b;
        ctr[1] += sp;
        ctr[2] += r;
    };
    const RayState* q = stage == 0 ? J.queue : J.spill;
    RayState* out = stage == 0 && max_queries > 0 ? J.spill : nullptr;
    if (group > 1)
        drain_groups(q, len, &w[1], group, scene_of, args_of, account, max_queries, out, out ? J.qwords + 2 : nullptr);
    else
        drain_queue(q, len, &w[1], A.coop, scene_of, args_of, account);
    flush_counters(ctr, A.counters);
}

__global__ void __launch_bounds__(256) k_plan_reduce(const PlanJob* jobs) {
    const PlanJob& J = jobs[blockIdx.y];
    if ((int64_t)blockIdx.x >= J.red_blocks) return;  // whole block
    reduce_block(blockIdx.x, J.args.n, J.args.cap, J.args.w, J.args.info, J.args.contrib, J.partials, J.hist,
                 J.ledger, reinterpret_cast<int*>(J.args.counters + 6));
}

__global__ void __launch_bounds__(256) k_plan_finalize(const PlanJob* jobs) {
    const PlanJob& J = jobs[blockIdx.x];
    finalize_block(J.partials, J.red_blocks, reinterpret_cast<Partial*>(J.args.counters + 8));
}

}  // namespace

}  // namespace ehss

using namespace ehss;

struct ehss_plan {
    int n_jobs = 0, precision = 0, variant = 0, coop = 0, group = 1, stage1 = 0;
    unsigned flags = 0;
    bool pa = false, use_graph = false, captured = false, dirty = true;
    bool holds = false;  // the scenes count this plan (ehss_scene::plans)
    std::vector<PlanJob> jobs;
    std::vector<ehss_scene*> scenes;
    DevBuf jobs_dev, info, contrib, partials, words, bins, queue, queue2;
    int64_t rays = 0, max_blocks = 0, max_red = 0, n_bins = 0, n_words = 0, drain_grid = 0;
    size_t lds_bytes = 0;  // dynamic LDS of the packet kernel (0: scenes read from HBM)
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

constexpr int kPlanWords = kCounterWords + 2;  // per job: counters, then 4 uint32 queue words
// Default drain lanes per ray of a plan (first stage, second stage).
constexpr int kPlanGroup = 4;
constexpr int kPlanStage2Group = kWave;
// Packet kernels: a staged scene of at most this size keeps several blocks per
// CU and fits beside the kernels' static LDS in a workgroup's 64 KB.
constexpr size_t kPlanLdsLimit = 40 * 1024;
static_assert(kPlanLdsLimit + kPacketStaticLds <= 64 * 1024, "packet kernel LDS exceeds 64 KB per workgroup");

void destroy_plan(ehss_plan* p) {
    if (!p) return;
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    if (p->holds)
        for (ehss_scene* sc : p->scenes) sc->plans.fetch_sub(1);
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->ev0) (void)hipEventDestroy(p->ev0);
    if (p->ev1) (void)hipEventDestroy(p->ev1);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

// The launch sequence: zero fills (kernels, so that it can be captured), the
// packet kernel, the drain stage(s), the reductions.
int enqueue_plan(ehss_plan* p) {
    hipStream_t s = p->stream;
    EHSS_TRY(launch_zero(p->words.as<unsigned long long>(), p->n_words, s));
    EHSS_TRY(launch_zero(p->bins.as<unsigned long long>(), 2 * p->n_bins, s));
    const PlanJob* jd = p->jobs_dev.as<PlanJob>();
    const dim3 grid((unsigned)p->max_blocks, (unsigned)p->n_jobs), block(256);
    const size_t lds = p->lds_bytes;
#define EHSS_PLAN_L(M, K)                                                                  \
    do {                                                                                   \
        if (p->pa)                                                                         \
            hipLaunchKernelGGL((k_plan_solve<M, 1, K>), grid, block, lds, s, jd);         \
        else                                                                               \
            hipLaunchKernelGGL((k_plan_solve<M, 0, K>), grid, block, lds, s, jd);         \
    } while (0)
#define EHSS_PLAN(M)                 \
    do {                             \
        if (lds) EHSS_PLAN_L(M, 1);  \
        else EHSS_PLAN_L(M, 0);      \
    } while (0)
    switch (p->variant) {
        case 0: EHSS_PLAN_L(0, 0); break;
        case 1: EHSS_PLAN_L(1, 0); break;
        case 2: EHSS_PLAN(2); break;
        case 3: EHSS_PLAN(3); break;
        default: EHSS_PLAN(4); break;
    }
#undef EHSS_PLAN_L
#undef EHSS_PLAN
    if (p->variant == 4) {
        const dim3 dgrid((unsigned)p->drain_grid, (unsigned)p->n_jobs);
        auto drain = [&](int stage, int group, int max_queries) {
            hipLaunchKernelGGL(k_plan_drain, dgrid, block, 0, s, jd, stage, group, max_queries);
        };
        const bool two = p->stage1 > 0 && p->group != 1;
        drain(0, p->group, two ? p->stage1 : 0);
        if (two) drain(1, kPlanStage2Group, 0);
    }
    hipLaunchKernelGGL(k_plan_reduce, dim3((unsigned)p->max_red, (unsigned)p->n_jobs), block, 0, s, jd);
    hipLaunchKernelGGL(k_plan_finalize, dim3((unsigned)p->n_jobs), block, 0, s, jd);
    EHSS_HIP(hipGetLastError());
    return EHSS_OK;
}

// Points the plan at a job list: validates it, lays out the records, queues,
// partials, words and bins (grow-only buffers), fills the host job table and
// moves the plan's holds from its previous scenes to these.  n_jobs == 0
// empties the plan (holds released; launching is refused until it is filled).
int fill_plan(ehss_plan* p, int n_jobs, ehss_scene* const* scenes, const ehss_sobol* sobols, const int32_t* powers,
              const int32_t* caps) {
    EHSS_CHECK(n_jobs >= 0 && n_jobs <= 65535 && (n_jobs == 0 || (scenes && sobols && powers && caps)),
               "invalid plan description");
    int64_t rays = 0;
    for (int j = 0; j < n_jobs; ++j) {
        EHSS_CHECK(scenes[j], "null scene for job %d", j);
        std::lock_guard<std::mutex> guard(scenes[j]->lock);
        EHSS_CHECK(!scenes[j]->pending, "a solve is pending on the scene of job %d", j);
        EHSS_CHECK(powers[j] >= 0 && powers[j] <= 26, "power of job %d out of range 0..26", j);
        EHSS_CHECK(caps[j] >= 1 && caps[j] <= (int)kBounceMask - 1, "cap of job %d out of range", j);
        rays += 1LL << powers[j];
    }
    EHSS_CHECK(rays <= (1LL << 30), "%lld rays exceed 2^30 per plan", (long long)rays);
    // The previous job list's holds go first (the same scene may appear again).
    if (p->holds)
        for (ehss_scene* sc : p->scenes) sc->plans.fetch_sub(1);
    p->holds = false;
    p->scenes.assign(scenes, scenes + n_jobs);
    p->jobs.clear();
    p->n_jobs = n_jobs;
    p->rays = rays;
    p->dirty = true;
    p->max_blocks = p->max_red = p->n_bins = p->n_words = 0;
    if (n_jobs == 0) return EHSS_OK;
    const unsigned flags = p->flags;
    p->group = group_from_flags(flags, rays);
    // A plan's jobs drain side by side, each on its slice of the grid, and the
    // largest job's long paths set the launch time, so the library default is
    // fixed lanes here: 4 per ray, then whole waves.  18 globules of 200-5,000
    // atoms at 2^20 rays: 14.4 ms per launch with 4 / 64 lanes against 17.4 ms
    // with lanes from each job's queue length (which the tuner can still pick).
    if (p->group == kAutoGroup) p->group = kPlanGroup;
    p->stage1 = stage1_queries(flags, rays);
    std::vector<int64_t> ray_off(n_jobs), red_off(n_jobs), bin_off(n_jobs);
    int64_t red = 0, bins = 0, off = 0;
    for (int j = 0; j < n_jobs; ++j) {
        const int64_t n = 1LL << powers[j];
        const int64_t rb = (n + kReduceRays - 1) / kReduceRays;
        ray_off[j] = off;
        red_off[j] = red;
        bin_off[j] = bins;
        off += n;
        red += rb;
        bins += caps[j] + 1;
        p->max_blocks = std::max(p->max_blocks, (n + 255) / 256);
        p->max_red = std::max(p->max_red, rb);
    }
    EHSS_CHECK(p->max_blocks <= (1LL << 31) - 1, "too many blocks per job");
    p->n_bins = bins;
    p->n_words = (int64_t)kPlanWords * n_jobs;
    EHSS_TRY(p->info.reserve(sizeof(uint32_t) * rays));
    EHSS_TRY(p->contrib.reserve(sizeof(double) * rays));
    EHSS_TRY(p->partials.reserve(sizeof(Partial) * red));
    EHSS_TRY(p->words.reserve(sizeof(unsigned long long) * p->n_words));
    EHSS_TRY(p->bins.reserve(sizeof(unsigned long long) * 2 * bins));
    if (p->variant == 4) {
        EHSS_TRY(p->queue.reserve(sizeof(RayState) * rays));
        if (p->stage1 > 0 && p->group != 1) EHSS_TRY(p->queue2.reserve(sizeof(RayState) * rays));
    }
    EHSS_TRY(p->jobs_dev.reserve(sizeof(PlanJob) * n_jobs));
    unsigned long long* W = p->words.as<unsigned long long>();
    unsigned long long* H = p->bins.as<unsigned long long>();
    p->jobs.resize(n_jobs);
    for (int j = 0; j < n_jobs; ++j) {
        ehss_scene* sc = scenes[j];
        std::lock_guard<std::mutex> guard(sc->lock);
        // The plan's launches read the scene's arrays: they follow its last
        // (re)build, which cannot change while the plan holds the scene.
        if (j == 0 || scenes[j - 1] != sc) EHSS_HIP(hipStreamWaitEvent(p->stream, sc->ev_built, 0));
        PlanJob& J = p->jobs[j];
        memset(&J, 0, sizeof J);
        J.view = sc->view;
        const int64_t n = 1LL << powers[j];
        J.args.n = n;
        J.args.power = powers[j];
        J.args.cap = caps[j];
        J.args.w = M_PI * sc->radius * sc->radius / (double)n;  // np.pi*radius*radius/n
        J.args.cx = sc->center[0];
        J.args.cy = sc->center[1];
        J.args.cz = sc->center[2];
        J.args.radius = sc->radius;
        J.args.info = p->info.as<uint32_t>() + ray_off[j];
        J.args.contrib = p->contrib.as<double>() + ray_off[j];
        J.args.counters = W + (int64_t)kPlanWords * j;
        J.args.coop = p->coop;
        J.args.group = p->group;
        J.args.packets = packets_from_flags(flags);
        J.args.trig = trig_compat();
        J.args.times = nullptr;
        J.sob = sobols[j];
        ray_map_words(J.sob, powers[j], J.map);
        J.red_blocks = (n + kReduceRays - 1) / kReduceRays;
        J.partials = p->partials.as<Partial>() + red_off[j];
        J.hist = H + bin_off[j];
        J.ledger = H + bins + bin_off[j];
        J.qwords = reinterpret_cast<uint32_t*>(W + (int64_t)kPlanWords * j + kCounterWords);
        if (p->variant == 4) {
            J.queue = p->queue.as<RayState>() + ray_off[j];
            if (p->stage1 > 0 && p->group != 1) J.spill = p->queue2.as<RayState>() + ray_off[j];
        }
    }
    int64_t most_rays = 0;
    for (int j = 0; j < n_jobs; ++j) most_rays = std::max<int64_t>(most_rays, 1LL << powers[j]);
    p->drain_grid = drain_blocks(most_rays);
    p->lds_bytes = 0;
    if ((flags & EHSS_FLAG_LDS_PACKETS) && p->variant >= 2) {
        size_t most = 0;
        for (int j = 0; j < n_jobs; ++j) most = std::max(most, staged_scene_bytes(scenes[j]->view, true));
        if (most <= kPlanLdsLimit) p->lds_bytes = most;
    }
    // The plan copied each scene's views (device pointers, counts, disk):
    // the scenes refuse rebuild and destroy until the plan lets them go.
    for (ehss_scene* sc : p->scenes) sc->plans.fetch_add(1);
    p->holds = true;
    return EHSS_OK;
}

int create_plan(int n_jobs, ehss_scene* const* scenes, const ehss_sobol* sobols, const int32_t* powers,
                const int32_t* caps, int precision, unsigned flags, ehss_plan** out) {
    EHSS_TRY(ensure_init());
    EHSS_CHECK(out, "null output handle");
    *out = nullptr;
    EHSS_CHECK(n_jobs >= 1, "a plan needs at least one job");
    EHSS_CHECK(precision == EHSS_PRECISION_FP64 || precision == EHSS_PRECISION_FP32_FIXUP, "unknown precision");
    int variant = (int)((flags >> EHSS_VARIANT_SHIFT) & 0xFu);
    if (precision == EHSS_PRECISION_FP64) variant = 0;
    else if (variant == 0) variant = 4;
    EHSS_CHECK(variant >= 0 && variant <= 4, "unknown kernel variant %d", variant);
    EHSS_CHECK(coop_from_flags(flags) <= kWave, "tail threshold %d out of range 0..64", coop_from_flags(flags));
    ehss_plan* p = new ehss_plan();
    struct Guard {
        ehss_plan*& q;
        ~Guard() { destroy_plan(q); }
    } guard{p};
    p->flags = flags;
    p->precision = precision;
    p->variant = variant;
    p->pa = (flags & EHSS_FLAG_PA) != 0;
    p->use_graph = (flags & EHSS_FLAG_GRAPH) != 0;
    p->coop = coop_from_flags(flags);
    EHSS_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    EHSS_HIP(hipEventCreate(&p->ev0));
    EHSS_HIP(hipEventCreate(&p->ev1));
    EHSS_TRY(fill_plan(p, n_jobs, scenes, sobols, powers, caps));
    *out = p;
    p = nullptr;
    return EHSS_OK;
}

// Re-points a plan at another job list, reusing its stream and (grow-only)
// buffers; waits for the plan's previous launch, drops a captured graph (the
// next launch captures the new sequence).
int retarget_plan(ehss_plan* p, int n_jobs, ehss_scene* const* scenes, const ehss_sobol* sobols,
                  const int32_t* powers, const int32_t* caps) {
    EHSS_CHECK(p, "null plan");
    EHSS_HIP(hipStreamSynchronize(p->stream));
    if (p->exec) EHSS_HIP(hipGraphExecDestroy(p->exec));
    if (p->graph) EHSS_HIP(hipGraphDestroy(p->graph));
    p->exec = nullptr;
    p->graph = nullptr;
    p->captured = false;
    return fill_plan(p, n_jobs, scenes, sobols, powers, caps);
}

int launch_plan(ehss_plan* p) {
    EHSS_CHECK(p, "null plan");
    EHSS_CHECK(p->n_jobs > 0 && p->holds, "the plan has no jobs (retarget it first)");
    hipStream_t s = p->stream;
    if (p->dirty) {
        EHSS_HIP(hipMemcpyAsync(p->jobs_dev.p, p->jobs.data(), sizeof(PlanJob) * p->n_jobs, hipMemcpyHostToDevice,
                                s));
        EHSS_HIP(hipStreamSynchronize(s));  // the host table may change before the next launch
        p->dirty = false;
    }
    EHSS_HIP(hipEventRecord(p->ev0, s));
    if (p->use_graph) {
        if (!p->captured) {
            EHSS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            const int st = enqueue_plan(p);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(s, &g);
            EHSS_TRY(st);
            EHSS_HIP(e);
            p->graph = g;
            EHSS_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
            p->captured = true;
        }
        EHSS_HIP(hipGraphLaunch(p->exec, s));
    } else {
        EHSS_TRY(enqueue_plan(p));
    }
    EHSS_HIP(hipEventRecord(p->ev1, s));
    return EHSS_OK;
}

int wait_plan(ehss_plan* p, ehss_totals* totals, int64_t* histograms, double* ledgers) {
    EHSS_CHECK(p && totals, "null plan/totals");
    hipStream_t s = p->stream;
    std::vector<unsigned long long> w((size_t)p->n_words), b((size_t)(2 * p->n_bins));
    EHSS_HIP(hipMemcpyAsync(w.data(), p->words.p, sizeof(unsigned long long) * p->n_words, hipMemcpyDeviceToHost, s));
    EHSS_HIP(hipMemcpyAsync(b.data(), p->bins.p, sizeof(unsigned long long) * 2 * p->n_bins, hipMemcpyDeviceToHost,
                            s));
    EHSS_HIP(hipStreamSynchronize(s));
    float ms = 0.f;
    EHSS_HIP(hipEventElapsedTime(&ms, p->ev0, p->ev1));
    int64_t off = 0;
    for (int j = 0; j < p->n_jobs; ++j) {
        const PlanJob& J = p->jobs[j];
        const unsigned long long* c = w.data() + (int64_t)kPlanWords * j;
        Partial fin;
        memcpy(&fin, c + 8, sizeof fin);
        ehss_totals& T = totals[j];
        memset(&T, 0, sizeof T);
        const double wgt = J.args.w;
        T.weight = wgt;
        T.omega = fin.s + fin.c;
        T.omega_err = fin.c;
        T.rays = J.args.n;
        T.escaped = fin.escaped;
        T.unresolved = fin.unresolved;
        T.pa_hits = p->pa ? fin.pa_hits : 0;
        T.ehss_hits = fin.ehss_hits;
        T.grazing_rays = fin.grazing;
        T.pa = p->pa ? wgt * (double)fin.pa_hits : 0.;
        T.pa_from_ehss = wgt * (double)fin.ehss_hits;
        T.residual_mass = wgt * (double)fin.unresolved;
        T.tail_bound = 2 * T.residual_mass;
        T.box_tests = (int64_t)c[0];
        T.sphere_tests = (int64_t)c[1];
        T.fp64_retests = (int64_t)(c[2] + c[5]);
        T.pa_box_tests = (int64_t)c[3];
        T.pa_sphere_tests = (int64_t)c[4];
        T.max_bounces = (int64_t)(int)c[6];
        T.kernel_ms = ms;  // the whole plan
        const int64_t nb = J.args.cap + 1;
        for (int64_t k = 0; k < nb; ++k) {
            if (histograms) histograms[off + k] = (int64_t)b[off + k];
            if (ledgers) ledgers[off + k] = (double)b[p->n_bins + off + k] / kLedgerScale * wgt;
        }
        off += nb;
    }
    return EHSS_OK;
}

}  // namespace

extern "C" {

int ehss_plan_create(int n_jobs, ehss_scene* const* scenes, const ehss_sobol* sobols, const int32_t* powers,
                     const int32_t* caps, int precision, unsigned flags, ehss_plan** out) {
    return create_plan(n_jobs, scenes, sobols, powers, caps, precision, flags, out);
}

int ehss_plan_retarget(ehss_plan* plan, int n_jobs, ehss_scene* const* scenes, const ehss_sobol* sobols,
                       const int32_t* powers, const int32_t* caps) {
    EHSS_TRY(ensure_init());
    return retarget_plan(plan, n_jobs, scenes, sobols, powers, caps);
}

int ehss_plan_set_sobol(ehss_plan* plan, int job, const ehss_sobol* sobol) {
    EHSS_CHECK(plan && sobol, "null plan/sobol");
    EHSS_CHECK(job >= 0 && job < plan->n_jobs, "job %d out of range", job);
    PlanJob& J = plan->jobs[job];
    J.sob = *sobol;
    ray_map_words(*sobol, J.args.power, J.map);
    plan->dirty = true;
    return EHSS_OK;
}

int ehss_plan_launch(ehss_plan* plan) {
    EHSS_TRY(ensure_init());
    return launch_plan(plan);
}

int ehss_plan_query(ehss_plan* plan, int* done) {
    EHSS_CHECK(plan && done, "null plan/done");
    const hipError_t e = hipEventQuery(plan->ev1);
    if (e == hipErrorNotReady) {
        *done = 0;
        return EHSS_OK;
    }
    EHSS_HIP(e);
    *done = 1;
    return EHSS_OK;
}

int ehss_plan_wait(ehss_plan* plan, ehss_totals* totals, int64_t* histograms, double* ledgers) {
    return wait_plan(plan, totals, histograms, ledgers);
}

int ehss_plan_destroy(ehss_plan* plan) {
    destroy_plan(plan);
    return EHSS_OK;
}

}  // extern "C"
