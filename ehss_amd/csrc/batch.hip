// batch.hip -- many fused solves (different structures, seeds, ray counts, caps)
// in one persistent launch.
//
// Work unit: a packet of 64 coherent rays of one job (one wave).  A fixed grid of
// waves pulls packets from a global counter until every job is done, so jobs of
// uneven size share the GPU without per-job launches.  Every ray writes its
// record into its own job's ray-index-ordered arrays, and each job is reduced in
// its own fixed order afterwards: a job's totals are bit-identical whether it
// runs alone, in a batch, or in any co-scheduling.  The zero-fill, the persistent
// kernel and the reductions can be captured once into a hipGraph and replayed;
// seeds change between replays through the job table in device memory.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "ehss_solve.h"

namespace ehss {

struct JobDev {
    SceneView view;
    SolveArgs args;
    ehss_sobol sob;
    RayMap map;
    int64_t pkt_begin;  // first 64-ray packet of this job in the batch
    int64_t red_begin;  // first reduction block of this job
    int64_t red_blocks;
    Partial* partials;
    unsigned long long* hist;
    unsigned long long* ledger;
    unsigned long long* words;  // [0..5] counters, [6] max bounces, [8..15] final Partial
};

namespace {

typedef const __attribute__((address_space(4))) uint32_t* CU32;
typedef const __attribute__((address_space(4))) int64_t* CI64;

// Copies a POD record from device memory through the constant address space so
// that wave-uniform fields land in scalar registers.
template <class T>
__device__ __forceinline__ T load_uniform(const T* p) {
    static_assert(sizeof(T) % 4 == 0, "record must be a whole number of dwords");
    T out;
    uint32_t* dst = reinterpret_cast<uint32_t*>(&out);
    const CU32 src = (CU32)p;
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 4); ++k) dst[k] = src[k];
    return out;
}

__device__ __forceinline__ int find_job(const JobDev* jobs, int n_jobs, int64_t key, bool by_packet) {
    int lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const CI64 p = (CI64)(by_packet ? &jobs[mid].pkt_begin : &jobs[mid].red_begin);
        if (*p <= key)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

template <int kMode, int kPA>
__global__ void __launch_bounds__(256) k_batch(const JobDev* jobs, int n_jobs, int64_t total_packets,
                                               unsigned long long* next_packet, RayState* queue,
                                               uint32_t* queue_count) {
    const int lane = threadIdx.x & 63;
    for (;;) {
        unsigned long long pkt = 0;
        if (lane == 0) pkt = atomicAdd(next_packet, 1ull);
        pkt = __shfl(pkt, 0, 64);
        pkt = ((unsigned long long)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(pkt >> 32)) << 32) |
              (unsigned long long)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)pkt);
        if ((int64_t)pkt >= total_packets) break;
        const int j = find_job(jobs, n_jobs, (int64_t)pkt, true);
        const SceneView S = load_uniform(&jobs[j].view);
        const SolveArgs A = load_uniform(&jobs[j].args);
        const CSobol sob = (CSobol)&jobs[j].sob;
        const CMap map = (CMap)&jobs[j].map;
        const int64_t begin = *(CI64)&jobs[j].pkt_begin;
        uint64_t ctr[6] = {0, 0, 0, 0, 0, 0};
        solve_ray<kMode, kPA>(S, A, sob, map, (uint32_t)(((int64_t)pkt - begin) * 64 + lane), ctr, queue,
                              queue_count, j);
        flush_counters(ctr, A.counters);
    }
}

// Drains the batch queue; each queued ray carries its job, whose scene and
// arguments are read from the job table.  Work counters go to the ray's job.
__global__ void __launch_bounds__(256) k_drain_batch(const JobDev* jobs, const RayState* queue,
                                                     const uint32_t* queue_count, uint32_t* queue_head, int32_t coop,
                                                     int32_t group) {
    const uint32_t len = __builtin_amdgcn_readfirstlane(*queue_count);
    auto scene_of = [&](int32_t j) -> const SceneView& { return jobs[j].view; };
    auto args_of = [&](int32_t j) -> const SolveArgs& { return jobs[j].args; };
    auto account = [&](int32_t j, uint64_t b, uint64_t sp, uint64_t r) {
        unsigned long long* c = jobs[j].args.counters;
        if (b) atomicAdd(&c[0], (unsigned long long)b);
        if (sp) atomicAdd(&c[1], (unsigned long long)sp);
        if (r) atomicAdd(&c[2], (unsigned long long)r);
    };
    if (group > 1)
        drain_groups(queue, len, queue_head, group, scene_of, args_of, account);
    else
        drain_queue(queue, len, queue_head, coop, scene_of, args_of, account);
}

__global__ void __launch_bounds__(256) k_reduce_batch(const JobDev* jobs, int n_jobs) {
    const int j = find_job(jobs, n_jobs, (int64_t)blockIdx.x, false);
    const JobDev& J = jobs[j];
    reduce_block((int64_t)blockIdx.x - J.red_begin, J.args.n, J.args.cap, J.args.w, J.args.info, J.args.contrib,
                 J.partials, J.hist, J.ledger, reinterpret_cast<int*>(J.words + 6));
}

__global__ void __launch_bounds__(256) k_finalize_batch(const JobDev* jobs) {
    const JobDev& J = jobs[blockIdx.x];
    finalize_block(J.partials, J.red_blocks, reinterpret_cast<Partial*>(J.words + 8));
}

__global__ void k_fill_zero(unsigned long long* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

}  // namespace

constexpr int kWords = 16;
constexpr bool kBatchEnabled = false;

}  // namespace ehss

using namespace ehss;

struct ehss_batch {
    int n_jobs = 0, precision = 0, variant = 0, blocks = 0, coop = 0, group = 1;
    bool pa = false, use_graph = false, captured = false, dirty = true, keep = false;
    std::vector<JobDev> jobs;
    std::vector<ehss_scene*> scenes;
    DevBuf jobs_dev, info, contrib, partials, words, queue;
    unsigned long long* next = nullptr;  // packet counter: word zero_words - 2 of `words`
    uint32_t* qwords = nullptr;          // queue length and head: last word of `words`
    int64_t rays = 0;
    int64_t total_packets = 0, total_red = 0, zero_words = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

int enqueue_batch(ehss_batch* b) {
    hipStream_t s = b->stream;
    const unsigned zero_blocks = (unsigned)std::min<int64_t>((b->zero_words + 255) / 256, 1024);
    hipLaunchKernelGGL(k_fill_zero, dim3(zero_blocks), dim3(256), 0, s, b->words.as<unsigned long long>(),
                       b->zero_words);
    const JobDev* jd = b->jobs_dev.as<JobDev>();
    unsigned long long* next = b->next;
    RayState* qbuf = b->variant == 4 ? b->queue.as<RayState>() : nullptr;
    const dim3 grid((unsigned)b->blocks), block(256);
#define EHSS_BATCH(M)                                                                                         \
    do {                                                                                                      \
        if (b->pa)                                                                                            \
            hipLaunchKernelGGL((k_batch<M, 1>), grid, block, 0, s, jd, b->n_jobs, b->total_packets, next, qbuf,   \
                               b->qwords);                                                                    \
        else                                                                                                  \
            hipLaunchKernelGGL((k_batch<M, 0>), grid, block, 0, s, jd, b->n_jobs, b->total_packets, next, qbuf,   \
                               b->qwords);                                                                    \
    } while (0)
    switch (b->variant) {
        case 0: EHSS_BATCH(0); break;
        case 1: EHSS_BATCH(1); break;
        case 2: EHSS_BATCH(2); break;
        case 3: EHSS_BATCH(3); break;
        default: EHSS_BATCH(4); break;
    }
#undef EHSS_BATCH
    if (b->variant == 4)
        hipLaunchKernelGGL(k_drain_batch, dim3((unsigned)drain_blocks(b->rays)), dim3(256), 0, s, jd, qbuf, b->qwords,
                           b->qwords + 1, b->coop, b->group);
    hipLaunchKernelGGL(k_reduce_batch, dim3((unsigned)b->total_red), dim3(256), 0, s, jd, b->n_jobs);
    hipLaunchKernelGGL(k_finalize_batch, dim3((unsigned)b->n_jobs), dim3(256), 0, s, jd);
    EHSS_HIP(hipGetLastError());
    return EHSS_OK;
}

void destroy_batch(ehss_batch* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int create_batch(int n_jobs, ehss_scene* const* scenes, const ehss_sobol* sobols, const int32_t* powers,
                 const int32_t* caps, int precision, unsigned flags, int blocks_per_cu, ehss_batch** out) {
    EHSS_TRY(ensure_init());
    EHSS_CHECK(out, "null output handle");
    *out = nullptr;
    EHSS_CHECK(n_jobs >= 1 && scenes && sobols && powers && caps, "invalid batch description");
    EHSS_CHECK(precision == EHSS_PRECISION_FP64 || precision == EHSS_PRECISION_FP32_FIXUP, "unknown precision");
    int variant = (int)((flags >> EHSS_VARIANT_SHIFT) & 0xFu);
    if (precision == EHSS_PRECISION_FP64) variant = 0;
    else if (variant == 0) variant = 4;
    EHSS_CHECK(variant >= 0 && variant <= 4, "unknown kernel variant %d", variant);
    for (int j = 0; j < n_jobs; ++j) {
        EHSS_CHECK(scenes[j], "null scene for job %d", j);
        EHSS_CHECK(powers[j] >= 0 && powers[j] <= 26, "power of job %d out of range 0..26", j);
        EHSS_CHECK(caps[j] >= 1 && caps[j] <= (int)kBounceMask - 1, "cap of job %d out of range", j);
    }
    ehss_batch* b = new ehss_batch();
    struct Guard {
        ehss_batch*& p;
        ~Guard() { destroy_batch(p); }
    } guard{b};
    b->n_jobs = n_jobs;
    b->precision = precision;
    b->variant = variant;
    b->pa = (flags & EHSS_FLAG_PA) != 0;
    b->use_graph = (flags & EHSS_FLAG_GRAPH) != 0;
    b->coop = coop_from_flags(flags);
    int64_t all_rays = 0;
    for (int j = 0; j < n_jobs; ++j) all_rays += 1LL << powers[j];
    b->group = group_from_flags(flags, all_rays);
    EHSS_CHECK(b->coop <= kWave, "tail threshold %d out of range 0..64", b->coop);
    b->scenes.assign(scenes, scenes + n_jobs);
    int64_t rays = 0, bins = 0;
    std::vector<int64_t> ray_off(n_jobs), bin_off(n_jobs), red_off(n_jobs), pkt_off(n_jobs);
    for (int j = 0; j < n_jobs; ++j) {
        const int64_t n = 1LL << powers[j];
        ray_off[j] = rays;
        bin_off[j] = bins;
        pkt_off[j] = b->total_packets;
        red_off[j] = b->total_red;
        rays += n;
        bins += caps[j] + 1;
        b->total_packets += (n + 63) / 64;
        b->total_red += (n + kReduceRays - 1) / kReduceRays;
    }
    EHSS_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    EHSS_HIP(hipEventCreate(&b->ev0));
    EHSS_HIP(hipEventCreate(&b->ev1));
    EHSS_TRY(b->info.alloc(sizeof(uint32_t) * rays));
    EHSS_TRY(b->contrib.alloc(sizeof(double) * rays));
    EHSS_TRY(b->partials.alloc(sizeof(Partial) * b->total_red));
    // words, then histogram and ledger bins in one zero-filled block; the packet
    // counter is its last word.
    b->zero_words = (int64_t)kWords * n_jobs + 2 * bins + 2;
    b->rays = rays;
    if (variant == 4) EHSS_TRY(b->queue.alloc(sizeof(RayState) * rays));
    EHSS_TRY(b->words.alloc(sizeof(unsigned long long) * b->zero_words));
    unsigned long long* W = b->words.as<unsigned long long>();
    unsigned long long* H = W + (int64_t)kWords * n_jobs;
    unsigned long long* LG = H + bins;
    b->next = W + (b->zero_words - 2);
    b->qwords = reinterpret_cast<uint32_t*>(W + (b->zero_words - 1));
    b->jobs.resize(n_jobs);
    for (int j = 0; j < n_jobs; ++j) {
        ehss_scene* sc = scenes[j];
        JobDev& J = b->jobs[j];
        memset(&J, 0, sizeof J);
        J.view = sc->view;
        const int64_t n = 1LL << powers[j];
        J.args.n = n;
        J.args.power = powers[j];
        J.args.cap = caps[j];
        J.args.w = M_PI * sc->radius * sc->radius / (double)n;
        J.args.cx = sc->center[0];
        J.args.cy = sc->center[1];
        J.args.cz = sc->center[2];
        J.args.radius = sc->radius;
        J.args.info = b->info.as<uint32_t>() + ray_off[j];
        J.args.contrib = b->contrib.as<double>() + ray_off[j];
        J.args.counters = W + (int64_t)kWords * j;
        J.args.coop = b->coop;
        J.args.group = b->group;
        J.args.packets = packets_from_flags(flags);
        J.sob = sobols[j];
        make_ray_map(J.sob, powers[j], J.map.col);
        J.pkt_begin = pkt_off[j];
        J.red_begin = red_off[j];
        J.red_blocks = (n + kReduceRays - 1) / kReduceRays;
        J.partials = b->partials.as<Partial>() + red_off[j];
        J.hist = H + bin_off[j];
        J.ledger = LG + bin_off[j];
        J.words = W + (int64_t)kWords * j;
    }
    EHSS_TRY(b->jobs_dev.alloc(sizeof(JobDev) * n_jobs));
    hipDeviceProp_t prop;
    int dev = 0;
    EHSS_HIP(hipGetDevice(&dev));
    EHSS_HIP(hipGetDeviceProperties(&prop, dev));
    const int per_cu = blocks_per_cu > 0 ? blocks_per_cu : 4;
    b->blocks = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)prop.multiProcessorCount * per_cu,
                                                            (b->total_packets + 3) / 4));
    *out = b;
    b = nullptr;
    return EHSS_OK;
}

int launch_batch(ehss_batch* b) {
    EHSS_CHECK(b, "null batch");
    hipStream_t s = b->stream;
    if (b->dirty) {
        EHSS_HIP(hipMemcpyAsync(b->jobs_dev.p, b->jobs.data(), sizeof(JobDev) * b->n_jobs, hipMemcpyHostToDevice, s));
        b->dirty = false;
    }
    EHSS_HIP(hipEventRecord(b->ev0, s));
    if (b->use_graph) {
        if (!b->captured) {
            EHSS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            const int st = enqueue_batch(b);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(s, &g);
            EHSS_TRY(st);
            EHSS_HIP(e);
            b->graph = g;
            EHSS_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
            b->captured = true;
        }
        EHSS_HIP(hipGraphLaunch(b->exec, s));
    } else {
        EHSS_TRY(enqueue_batch(b));
    }
    EHSS_HIP(hipEventRecord(b->ev1, s));
    return EHSS_OK;
}

int wait_batch(ehss_batch* b, ehss_totals* totals, int64_t* histograms, double* ledgers) {
    EHSS_CHECK(b, "null batch");
    hipStream_t s = b->stream;
    const int64_t nw = b->zero_words;
    std::vector<unsigned long long> host((size_t)nw);
    EHSS_HIP(hipMemcpyAsync(host.data(), b->words.p, sizeof(unsigned long long) * nw, hipMemcpyDeviceToHost, s));
    EHSS_HIP(hipStreamSynchronize(s));
    float ms = 0.f;
    EHSS_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    const unsigned long long* H = host.data() + (int64_t)kWords * b->n_jobs;
    int64_t bins = 0;
    for (int j = 0; j < b->n_jobs; ++j) bins += b->jobs[j].args.cap + 1;
    const unsigned long long* LG = H + bins;
    int64_t off = 0;
    for (int j = 0; j < b->n_jobs; ++j) {
        const JobDev& J = b->jobs[j];
        const unsigned long long* w = host.data() + (int64_t)kWords * j;
        Partial fin;
        memcpy(&fin, w + 8, sizeof fin);
        if (totals) {
            ehss_totals& T = totals[j];
            memset(&T, 0, sizeof T);
            T.weight = J.args.w;
            T.omega = fin.s + fin.c;
            T.omega_err = fin.c;
            T.rays = J.args.n;
            T.escaped = fin.escaped;
            T.unresolved = fin.unresolved;
            T.pa_hits = b->pa ? fin.pa_hits : 0;
            T.ehss_hits = fin.ehss_hits;
            T.grazing_rays = fin.grazing;
            T.pa = b->pa ? J.args.w * (double)fin.pa_hits : 0.;
            T.pa_from_ehss = J.args.w * (double)fin.ehss_hits;
            T.residual_mass = J.args.w * (double)fin.unresolved;
            T.tail_bound = 2 * T.residual_mass;
            T.box_tests = (int64_t)w[0];
            T.sphere_tests = (int64_t)w[1];
            T.fp64_retests = (int64_t)(w[2] + w[5]);
            T.pa_box_tests = (int64_t)w[3];
            T.pa_sphere_tests = (int64_t)w[4];
            T.max_bounces = (int64_t)(int)w[6];
            T.kernel_ms = ms;  // whole batch
        }
        const int nb = J.args.cap + 1;
        for (int k = 0; k < nb; ++k) {
            if (histograms) histograms[off + k] = (int64_t)H[off + k];
            if (ledgers) ledgers[off + k] = (double)LG[off + k] / kLedgerScale * J.args.w;
        }
        off += nb;
    }
    return EHSS_OK;
}

}  // namespace

extern "C" {

int ehss_batch_create(int n_jobs, ehss_scene* const* scenes, const ehss_sobol* sobols, const int32_t* powers,
                      const int32_t* caps, int precision, unsigned flags, int blocks_per_cu, ehss_batch** out) {
    // The persistent batch launch did not complete on its first device run and
    // the cause has not been identified; it stays disabled until it is.
    if (out) *out = nullptr;
    if (!kBatchEnabled)
        return set_error(EHSS_ERR_INVALID, "batched launch is disabled: its first device run did not complete "
                                           "and the cause is not yet identified; use ehss_solve per job");
    return create_batch(n_jobs, scenes, sobols, powers, caps, precision, flags, blocks_per_cu, out);
}

int ehss_batch_set_sobol(ehss_batch* batch, int job, const ehss_sobol* sobol) {
    EHSS_CHECK(batch && sobol, "null batch/sobol");
    EHSS_CHECK(job >= 0 && job < batch->n_jobs, "job %d out of range", job);
    batch->jobs[job].sob = *sobol;
    make_ray_map(*sobol, batch->jobs[job].args.power, batch->jobs[job].map.col);
    batch->dirty = true;
    return EHSS_OK;
}

int ehss_batch_launch(ehss_batch* batch) { return launch_batch(batch); }

int ehss_batch_query(ehss_batch* batch, int* done) {
    EHSS_CHECK(batch && done, "null batch/done");
    const hipError_t e = hipEventQuery(batch->ev1);
    if (e == hipSuccess) {
        *done = 1;
        return EHSS_OK;
    }
    if (e == hipErrorNotReady) {
        *done = 0;
        return EHSS_OK;
    }
    EHSS_HIP(e);
    return EHSS_OK;
}

int ehss_batch_wait(ehss_batch* batch, ehss_totals* totals, int64_t* histograms, double* ledgers) {
    return wait_batch(batch, totals, histograms, ledgers);
}

int ehss_batch_destroy(ehss_batch* batch) {
    destroy_batch(batch);
    return EHSS_OK;
}

}  // extern "C"
