/*
 * ehss_hip.h -- C-ABI of libehss_hip.so, the MI355X (gfx950) EHSS/PA engine.
 *
 * Plain C types only: pointers, sizes, status codes.  Every entry point except
 * ehss_trace returns an int status (EHSS_OK == 0); the message of the last
 * failure on the calling thread is available from ehss_last_error().
 *
 * Host arrays are owned by the caller; the library never retains a caller
 * pointer after a call returns.  Device memory is owned by the library
 * (scene handles, request handles, workspaces).
 *
 * Reference interfaces replaced (paths relative to the reference checkout):
 *   ehss_trace               src/native/ehss_cpu.cpp:42-46 (identical signature),
 *                            bound by ehss_cpu_cpp.py:13-16.
 *   ehss_scene_create        PreparedStaticCPUEHSS.__init__ (ehss_static.py:62-101)
 *                            plus SphereIndex.__init__ (pa_spatial.py:157-160).
 *   ehss_scene_trace         CPUEHSSTransport.trace (ehss_cpu.py:118-142).
 *   ehss_scene_pa_query      SphereIndex.query / tree_blocked (pa_spatial.py:133-142,162-164).
 *   ehss_solve               one replicate of solve_frame (ccs_cli.py:196-219):
 *                            external_source (ehss_state.py:42-52) -> index.query ->
 *                            engine.trace -> EHSSResult readout (ehss_state.py:84-103).
 *   ehss_plan_*              no single reference counterpart: the per-frame
 *                            solve_frame loop (ccs_cli.py:305-307) over many
 *                            structures as one launch.
 */
#ifndef EHSS_HIP_H
#define EHSS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EHSS_ABI_VERSION 1

/* Status codes. */
#define EHSS_OK 0
#define EHSS_ERR_INVALID 1   /* bad argument; nothing was launched */
#define EHSS_ERR_HIP 2       /* HIP runtime or kernel failure */
#define EHSS_ERR_NOMEM 3     /* device or host allocation failed */
#define EHSS_ERR_ADMISSION 4 /* request does not fit the configured device budget */
#define EHSS_ERR_BUSY 5      /* request still running (ehss_request_poll) */

/* Precision / kernel selectors. */
#define EHSS_PRECISION_FP64 0      /* plain FP64 kernel, reference operation order */
#define EHSS_PRECISION_FP32_FIXUP 1 /* FP32 filter + FP64 re-test; identical results */

/* ehss_solve flags. */
#define EHSS_FLAG_PA 1u        /* also run the BVH any-hit PA on the same rays */
#define EHSS_FLAG_KEEP_RAYS 2u /* keep per-ray records for ehss_scene_last_rays */
#define EHSS_FLAG_GRAPH 4u     /* plan: capture the launch sequence in a hipGraph and replay it */
#define EHSS_FLAG_TIMESTAMPS 8u /* record each ray's completion time (ehss_scene_last_ray_times) */
#define EHSS_FLAG_LDS_PACKETS 16u /* packet kernels stage the scene in LDS first (measured: not faster than
                                     scalar loads through the constant cache; kept for measurement) */
#define EHSS_FLAG_PER_REPLICATE 32u /* solve_many: never schedule replicates across the lanes of a wave */
#define EHSS_FLAG_CONTINUABLE 64u /* keep the records and the state of every ray that ends unresolved, so that
                                     ehss_solve_extend can continue exactly those rays at a larger cap */
#define EHSS_FLAG_UNSORTED 128u /* solve_many of >= 256 replicates of small nets: keep the shared-partition
                                   schedule instead of ordering the launch's rays by 4-D cell (same results) */
/* Kernel variant of the FP32 filter path in flag bits 8..11 (0 = default):
 * 1 one ray per lane, 2 wave-uniform first query then per-lane bounces,
 * 3 wave-uniform traversal for every bounce, 4 wave-uniform first queries, then
 * the rays still bouncing are compacted into a queue drained by a persistent
 * kernel.  All return identical per-ray results. */
#define EHSS_VARIANT_SHIFT 8
#define EHSS_VARIANT(v) ((unsigned)(v) << EHSS_VARIANT_SHIFT)
/* Tail threshold in flag bits 12..18: a wave left with at most n live rays
 * finishes them one at a time with the whole wave on each query.  Field value 0
 * selects the default; EHSS_COOP(0) turns the tail mode off.  n <= 64. */
#define EHSS_COOP_SHIFT 12
#define EHSS_COOP(n) (((unsigned)(n) + 1u) << EHSS_COOP_SHIFT)
/* Variant 4: lanes per queued ray in the drain, 2^(k-1) for field value k in
 * bits 19..21 (0 = default).  With one lane per ray the drain traverses the tree
 * per lane; with more, each query is split over the ray's lanes. */
#define EHSS_GROUP_SHIFT 19
#define EHSS_GROUP_LOG2(k) (((unsigned)(k) + 1u) << EHSS_GROUP_SHIFT)
/* Variant 4: queries a queued ray gets in the drain before it moves to a second
 * drain with a whole wave per ray, as q + 1 in bits 22..25 (0 = default,
 * EHSS_STAGE1(0) = a single drain).  q <= 14. */
#define EHSS_STAGE1_SHIFT 22
#define EHSS_STAGE1(q) (((unsigned)(q) + 1u) << EHSS_STAGE1_SHIFT)
/* Variant 4: wave-uniform (packet) queries every ray gets before the rays still
 * bouncing are queued, in bits 26..29 (0 = default 3).  1 <= q <= 15. */
#define EHSS_PACKETS_SHIFT 26
#define EHSS_PACKETS(q) ((unsigned)(q) << EHSS_PACKETS_SHIFT)

/* Library / device. */
int ehss_abi_version(void);
int ehss_last_error(char* buf, size_t size);
int ehss_init(int device);
int ehss_device_info(char* name, size_t name_size, int* cu_count, int64_t* total_mem, int* arch_major,
                     int* arch_minor);
/* Device FP64 arithmetic check (no reference counterpart): quotient[i] = a[i] / b[i]
 * and root[i] = sqrt(a[i]) computed by a kernel, for comparison with host IEEE. */
int ehss_selftest_fp64(int64_t n, const double* a, const double* b, double* quotient, double* root);

/* Source-angle cos/sin (no reference counterpart).  The incidence source takes
 * cos and sin of the Sobol angles x = fl(2pi * q 2^-30), q a 30-bit Sobol word
 * (ehss_state.py:48, inside_out_pa.py:47), which the reference evaluates with
 * NumPy, i.e. the host libm.  The device computes them correctly rounded and
 * applies a compatibility table (trig_compat.bin beside the library, or the file
 * named by $EHSS_TRIG_COMPAT) listing every q where the host libm rounds the
 * other way; it is read at initialisation.
 * ehss_trig_compat_info: *entries = angles listed, or -1 when no table was found
 * (cos/sin then stay correctly rounded); libm: the libm the table was generated
 * from; path: the file looked for.
 * ehss_sobol_sincos: s[i], c[i] = sin, cos of the angle of word q[i] < 2^30, as
 * the device source computes them. */
int ehss_trig_compat_info(int64_t* entries, char* libm, size_t libm_size, char* path, size_t path_size);
int ehss_sobol_sincos(int64_t n, const uint32_t* q, double* s, double* c);

/* ---------------------------------------------------------------------------
 * Reference seam.  Exact signature and array contract of ehss_trace at
 * src/native/ehss_cpu.cpp:42-46: all buffers are caller-owned host arrays;
 * p, d, counts are updated in place; escaped/unresolved must be zeroed by the
 * caller; history (nr x cap int32, -1 filled) is written only when record != 0;
 * counters[0..1] receive box and sphere tests.  Runs the plain FP64 kernel on
 * the current device.  `threads` is accepted for ABI compatibility and echoed in
 * *observed_threads; on failure *observed_threads = -1 and counters = -1, and
 * ehss_last_error() holds the reason.
 * ------------------------------------------------------------------------- */
void ehss_trace(int64_t nr, int64_t nn, int cap, int threads, int record, double* p, double* d,
                const int64_t* lasts, int64_t* counts, uint8_t* escaped, uint8_t* unresolved, int32_t* history,
                const double* lo, const double* hi, const int64_t* end, const int64_t* leaf,
                const int64_t* offsets, const int64_t* ids, const double* world, const double* radii,
                int64_t* counters, int* observed_threads);

/* ehss_trace with a status return and a precision selector. */
int ehss_trace_status(int64_t nr, int64_t nn, int cap, int precision, int record, double* p, double* d,
                      const int64_t* lasts, int64_t* counts, uint8_t* escaped, uint8_t* unresolved,
                      int32_t* history, const double* lo, const double* hi, const int64_t* end,
                      const int64_t* leaf, const int64_t* offsets, const int64_t* ids, const double* world,
                      const double* radii, int64_t* counters);

/* ---------------------------------------------------------------------------
 * Device-resident scenes: spheres + static EHSS tree + PA tree, built on the GPU.
 * ------------------------------------------------------------------------- */
typedef struct ehss_scene ehss_scene;

typedef struct {
    int64_t atoms, groups, nodes, pa_nodes;
    int32_t leaf_size, max_depth;
    double center[3]; /* centers.mean(axis=0) (ehss_state.py:46) */
    double radius;    /* max(|c_i - center| + r_i) * (1 + 1e-9) (ehss_state.py:47) */
    double scale;     /* FP32 filter scale: bound on |coordinate - center| */
    int64_t device_bytes;
    int32_t build_path; /* EHSS_BUILD_SHAPED or EHSS_BUILD_GENERAL: how the last (re)build ran */
    int32_t packet_nodes; /* nodes of the tree the packet queries walk: the static tree with
                           * subtrees of about 10 atoms merged into leaves (== nodes when none are) */
} ehss_scene_info;

/* Tree builds.  Without caller groups, with leaf_size <= 32 and no two atoms
 * sharing their key triple ((c - r) + (c + r)) / 2, no segment of the median
 * split can have zero span, so the tree's shape is a function of (n, leaf_size):
 * the build is enqueued on the scene's stream with every count known and
 * returns without synchronising (EHSS_BUILD_SHAPED).  Otherwise the level-by-
 * level build reads its counts back (EHSS_BUILD_GENERAL).  Both give the same
 * trees.  Setting the environment variable EHSS_BUILD_PATH=general forces the
 * latter. */
#define EHSS_BUILD_GENERAL 0
#define EHSS_BUILD_SHAPED 1
/* Group, static-tree node and PA node counts and the depth of the deepest
 * segment of the median-split tree of n atoms with leaf_size, when no segment
 * has zero span (the shaped case above): out[0..3]. */
int ehss_tree_counts(int64_t n, int leaf_size, int64_t* out);

/* centers: n x 3 (row-major), radii: n; leaf_size >= 1.  Builds on the device. */
int ehss_scene_create(int64_t n, const double* centers, const double* radii, int leaf_size, ehss_scene** out);
/* Same with caller-chosen groups (one id per atom, contiguous 0..g-1, all
 * nonempty), as PreparedStaticCPUEHSS(spheres, groups=...) (ehss_static.py:64-77);
 * the PA tree is still the median-split tree of leaf_size. */
int ehss_scene_create_grouped(int64_t n, const double* centers, const double* radii, const int64_t* groups,
                              int leaf_size, ehss_scene** out);
/* Replaces the scene's structure by another (a new conformer), as
 * PreparedCPUEHSS.update(spheres) does for the reference's prepared geometry
 * (ehss_cpu_workload.py:35-46): the trees and filter arrays are rebuilt on the
 * scene's stream into the scene's existing device arrays, which only grow, so a
 * structure no larger than the largest one the scene has held allocates and
 * frees nothing.  Fails with EHSS_ERR_BUSY while a solve is pending or while a
 * prepared plan (ehss_plan_create) holds the scene; per-ray records of the
 * previous structure are dropped.  Same validation as ehss_scene_create. */
int ehss_scene_rebuild(ehss_scene* scene, int64_t n, const double* centers, const double* radii, int leaf_size);
/* Fails with EHSS_ERR_BUSY (and destroys nothing) while a plan holds the scene. */
int ehss_scene_destroy(ehss_scene* scene);
int ehss_scene_get_info(const ehss_scene* scene, ehss_scene_info* info);
/* Copies the static tree back (arrays sized from ehss_scene_get_info):
 * lo, hi: nodes x 3; end, leaf_group: nodes; offsets: groups + 1; ids, groups: atoms. */
int ehss_scene_export_tree(const ehss_scene* scene, double* lo, double* hi, int64_t* end, int64_t* leaf_group,
                           int64_t* offsets, int64_t* ids, int64_t* groups);
/* PA tree (pa_spatial.py:17-28 layout): lo, hi: pa_nodes x 3; end, first, count: pa_nodes; ids: atoms. */
int ehss_scene_export_pa_tree(const ehss_scene* scene, double* lo, double* hi, int64_t* end, int64_t* first,
                              int64_t* count, int64_t* ids);

/* Traces caller rays (host arrays, ehss_trace contract) against a scene. */
int ehss_scene_trace(ehss_scene* scene, int64_t nr, int cap, int precision, int record, double* p, double* d,
                     const int64_t* lasts, int64_t* counts, uint8_t* escaped, uint8_t* unresolved,
                     int32_t* history, int64_t* counters);
/* Any-hit PA on caller rays (SphereIndex.query contract): hits[nr] (0/1),
 * counters[0..1] = box tests, sphere tests on the PA tree. */
int ehss_scene_pa_query(ehss_scene* scene, int64_t nr, int precision, const double* origins,
                        const double* directions, const int64_t* owners, uint8_t* hits, int64_t* counters);

/* ---------------------------------------------------------------------------
 * Fused solve: scrambled Sobol -> impact-disk source -> [PA any-hit] -> trace ->
 * FP64 compensated (1 - cos chi) reduction, all on the device.
 * ------------------------------------------------------------------------- */

/* Scrambled direction numbers and digital shift of a 4-D, 30-bit Sobol net
 * (SciPy qmc.Sobol LMS + shift); sv[dim][j] is the number for index bit j. */
typedef struct {
    uint32_t sv[4][32];
    uint32_t shift[4];
} ehss_sobol;

typedef struct {
    double omega;          /* sum w (1 - u.v) over escaped rays, A^2 */
    double omega_err;      /* compensation residual of the FP64 sum (diagnostic) */
    double pa;             /* BVH any-hit PA: w * #hits (EHSS_FLAG_PA) */
    double pa_from_ehss;   /* w * #(bounces > 0 or unresolved) (ehss_state.py:101-103) */
    double residual_mass;  /* w * #unresolved (ehss_state.py:93-95) */
    double tail_bound;     /* 2 * residual_mass (ehss_state.py:97-99) */
    double weight;         /* w = pi R^2 / n */
    int64_t rays, escaped, unresolved, pa_hits, ehss_hits;
    int64_t box_tests, sphere_tests, pa_box_tests, pa_sphere_tests;
    int64_t fp64_retests;  /* FP32 path: exact FP64 re-evaluations */
    int64_t grazing_rays;  /* rays flagged near a hit/miss or nearest-atom decision */
    int64_t max_bounces;
    double kernel_ms;      /* device time of the trace kernel(s) (HIP events) */
} ehss_totals;

/* Host-only: the scrambled net of scipy.stats.qmc.Sobol(4, scramble=True,
 * seed) (_qmc.py:1808-1823; the ehss_state.py:44 call), built from NumPy's
 * default_rng(seed) stream (SeedSequence + PCG64, Lemire bounded draws) as SciPy
 * does.  Replaces the host scramble of the reference's source per replicate. */
int ehss_scramble(uint64_t seed, ehss_sobol* out);
/* ehss_scramble for n seeds on the host's cores (up to 16 threads): out[i] is
 * the net of seeds[i].  One CCS estimate needs one scramble, so batched
 * estimates make many per launch. */
int ehss_scramble_many(int64_t n, const uint64_t* seeds, ehss_sobol* out);

/* Host-only: the wave-coherent ray map used by ehss_solve for 2^power rays
 * (cols[32]; no reference counterpart).  Thread t = 64 w + l traces the ray with
 * Gray-code index g = XOR of cols[k] over set bits k of (w | l << (power - 6)),
 * so the 64 rays of a wave share one elementary cell of the scrambled net. */
int ehss_ray_map(const ehss_sobol* sobol, int power, uint32_t* cols);

/* Host-only: the cross-replicate schedule ehss_solve_many uses (no reference
 * counterpart): a spatial partition of `spatial` bits shared by the replicates
 * (0: the finest one with at least power - 6 bits).  out: per replicate 68 words --
 * inv2[32] (columns of T2^-1), rows1[32] (rows of the replicate's own T1),
 * cell_shift, 3 pad.  Index u = c | j << spatial (cell c, sub-index j < 2^(power
 * - spatial)) is Gray index g = T2_r^-1 (u ^ cell_shift_r), whose point lies in
 * cell c; it is recorded at the slot of g under the replicate's own map.  A
 * launch of many replicates of small nets takes one point (index c) of each of
 * 64 replicates per wave; a launch of 2^k <= 64 replicates takes 2^(6 - k)
 * points of each, from a partition of power - 6 + k bits.  EHSS_ERR_INVALID if
 * the replicates share no such partition (the launch then uses one replicate per
 * block row). */
int ehss_cross_map(const ehss_sobol* sobols, int reps, int power, int spatial, uint32_t* out);

/* Synchronous solve of 2^power rays.  histogram (cap+1 int64) and ledger
 * (cap+1 double) may be NULL. */
int ehss_solve(ehss_scene* scene, const ehss_sobol* sobol, int power, int cap, int precision, unsigned flags,
               ehss_totals* totals, int64_t* histogram, double* ledger);

/* Allocation-free solve.  No counterpart in the reference: it is for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  One replicate of ehss_solve -- same
 * kernels, same results -- with every work buffer (records, drain queue,
 * partials, histogram, counters, scramble tables) in one static device array of
 * EHSS_NOALLOC_SCRATCH_MB MiB that the library holds on each device until the
 * process ends.  It calls no hipMalloc/hipFree/hipHostMalloc and no allocating
 * library routine, records no memset, and does not synchronise: the launches
 * are enqueued on `stream` (a hipStream_t; NULL = the scene's stream) and the
 * results are written by a kernel into caller memory the device can write
 * (device memory, or mapped pinned host memory): totals (one ehss_totals,
 * kernel_ms = 0), histogram (cap + 1 int64) and ledger (cap + 1 double), both
 * may be NULL.  Limits: power <= EHSS_NOALLOC_MAX_POWER, cap <=
 * EHSS_NOALLOC_MAX_CAP; flags as ehss_solve except EHSS_FLAG_KEEP_RAYS,
 * EHSS_FLAG_TIMESTAMPS and a two-stage drain (EHSS_STAGE1(q > 0) with more than
 * one lane per ray), which are refused.  Every call on a device shares that
 * device's array and starts by clearing what must start at zero: calls that
 * overlap on one device (on different streams, or from different threads) race
 * and give wrong results.  The scene must not be rebuilt or destroyed before
 * the stream has passed the call. */
#define EHSS_NOALLOC_MAX_POWER 16
#define EHSS_NOALLOC_MAX_CAP 4096
#define EHSS_NOALLOC_SCRATCH_MB 9
int ehss_solve_noalloc(ehss_scene* scene, const ehss_sobol* sobol, int power, int cap, int precision, unsigned flags,
                       void* stream, ehss_totals* totals, int64_t* histogram, double* ledger);

/* Continues the scene's last solve (ehss_solve / ehss_solve_many with
 * EHSS_FLAG_CONTINUABLE, or a previous extension) at a larger collision cap:
 * only the rays that ended unresolved are traced further, from their saved
 * state, and every replicate is then reduced again in its fixed order.  The
 * results -- totals, cap + 1 histogram and ledger bins per replicate -- are
 * bit-identical to a full retrace of the same rays at new_cap (the work
 * counters are those of the continuation).  This is the cap ladder of the
 * publication protocol (manuscript.md:61: the same rays retraced at 1,024,
 * 2,048 and 4,096) without retracing the resolved paths.  Synchronous;
 * new_cap > the previous cap.  Invalidated by any other solve on the scene or a
 * rebuild. */
int ehss_solve_extend(ehss_scene* scene, int new_cap, ehss_totals* totals, int64_t* histograms, double* ledgers);

/* `replicates` independent solves of the scene in one launch (one scramble each,
 * the replicates of ccs_cli.py:186-227): totals gets one entry per replicate,
 * histograms / ledgers (may be NULL) cap + 1 bins per replicate, replicate-major.
 * Each replicate's totals are bit-identical to ehss_solve with its scramble.
 * kernel_ms is the device time of the whole launch.  1 <= replicates <= 16384
 * and replicates x 2^power <= 2^30.  Per-ray records (EHSS_FLAG_KEEP_RAYS) are
 * those of the first replicate. */
int ehss_solve_many(ehss_scene* scene, int replicates, const ehss_sobol* sobols, int power, int cap, int precision,
                    unsigned flags, ehss_totals* totals, int64_t* histograms, double* ledgers);

/* The same solves split for overlapped use (no reference counterpart; used by the
 * request loop): the _async calls enqueue on the scene's stream and return;
 * ehss_solve_query sets *done without blocking; ehss_solve_wait blocks and fills
 * what ehss_solve / ehss_solve_many fill (one totals entry per replicate of the
 * pending launch).  One launch may be pending per scene (EHSS_ERR_BUSY
 * otherwise); launches on different scenes run concurrently. */
int ehss_solve_async(ehss_scene* scene, const ehss_sobol* sobol, int power, int cap, int precision, unsigned flags);
int ehss_solve_many_async(ehss_scene* scene, int replicates, const ehss_sobol* sobols, int power, int cap,
                          int precision, unsigned flags);
int ehss_solve_query(ehss_scene* scene, int* done);
int ehss_solve_wait(ehss_scene* scene, ehss_totals* totals, int64_t* histogram, double* ledger);

/* Per-ray records of the last ehss_solve on this scene made with
 * EHSS_FLAG_KEEP_RAYS, in ray-index order: info bits (see below), angle
 * (1 - u.v clipped to [0,2], 0 unless escaped), bounces. */
#define EHSS_RAY_ESCAPED 0x1u
#define EHSS_RAY_UNRESOLVED 0x2u
#define EHSS_RAY_PA_HIT 0x4u
#define EHSS_RAY_EHSS_HIT 0x8u
#define EHSS_RAY_GRAZING 0x10u
int ehss_scene_last_rays(ehss_scene* scene, int64_t n, uint32_t* flags, double* angle, int32_t* bounces);

/* Per-ray completion time stamps of the last solve run with EHSS_FLAG_TIMESTAMPS
 * (device wall clock ticks, ray-index order) and the clock rate in Hz.  A
 * diagnostic of where a solve spends its time; no reference counterpart. */
int ehss_scene_last_ray_times(ehss_scene* scene, int64_t n, int64_t* ticks, double* tick_hz);

/* Device-generated source rays of a Sobol net (for parity tests of the source
 * mapping): origins/directions n x 3, u n x 4 (any may be NULL). */
int ehss_source_rays(ehss_scene* scene, const ehss_sobol* sobol, int power, double* u, double* origins,
                     double* directions);

/* ---------------------------------------------------------------------------
 * Alternative incidence sources, generated, traced and reduced on the device.
 *
 *   EHSS_SOURCE_BOX        CompactBoundary(kind='box').source (ehss_compact_source.py:56-70):
 *                          directions uniform on S2, a face chosen in proportion to
 *                          its projected area, a uniform point on it; weight =
 *                          projected area / n.  5-D net.
 *   EHSS_SOURCE_ELLIPSOID  CompactBoundary(kind='ellipsoid').source (ehss_compact_source.py:71-84):
 *                          a uniform point of the ellipsoid's projection; weight =
 *                          projected area / n.  4-D net (dimensions 0-3 of the table).
 *   EHSS_SOURCE_REVERSE_PA reverse_pa_source (ehss_state.py:55-69): sample_source's
 *                          area-selected surface points and cosine hemisphere
 *                          (inside_out_pa.py:72-90, 5-D net), the inside-out any-hit
 *                          test against every other sphere (inside_out_pa.py:115-118),
 *                          then the unblocked states reflected once (bounce count 1,
 *                          last atom = owner), weight = area / (4 n).
 * The scrambled net is SciPy's qmc.Sobol(dims, scramble=True, seed) as direction
 * numbers and shift (dims 5, or 4 for the ellipsoid: sv[4] and shift[4] unused).
 * Ray i of the net is point i of random_base2(power).  Parameters computed from
 * the structure with NumPy in the reference (fit, area, the searchsorted table)
 * are inputs, so their bits are the caller's.
 * ------------------------------------------------------------------------- */
typedef struct {
    uint32_t sv[5][32];
    uint32_t shift[5];
} ehss_sobol5;

#define EHSS_SOURCE_BOX 1
#define EHSS_SOURCE_ELLIPSOID 2
#define EHSS_SOURCE_REVERSE_PA 3

typedef struct {
    int32_t kind;
    int32_t pad;
    double center[3];    /* box / ellipsoid: CompactBoundary.center */
    double rotation[9];  /* row-major CompactBoundary.rotation (local = d @ rotation) */
    double axes[3];      /* half-widths (box) or semi-axes (ellipsoid) */
    double area;         /* reverse PA: Spheres.area = 4 pi r.r (inside_out_pa.py:42) */
    const double* cdf;   /* reverse PA: host array, n atoms: cumsum(r^2) / sum(r^2) (inside_out_pa.py:79) */
} ehss_source;

/* Traces the source's 2^power states against the scene and reduces them as
 * EHSSResult does (ehss_state.py:84-103) with per-ray weights: omega = sum w
 * (1 - u.v) over escaped rays, residual_mass = sum w over unresolved rays,
 * pa_from_ehss = sum w over rays with bounces > 0 or unresolved (fixed-order
 * compensated sums).  weight = the sum of all weights.  Reverse PA: pa =
 * area * mean(unblocked) / 4 (inside_out_pa.py:66-67), pa_hits = blocked
 * samples (not traced; their records are zero).  Optional per-ray outputs in ray
 * order: info (bounces | flags << 20, as ehss_scene_last_rays) and contrib
 * (weighted); histogram: cap + 1 bounce counts.  The states are traced by the
 * fused solve's queue drain (FP32 filter, FP64 fix-up: per-ray results equal
 * the FP64 trace's except on rays flagged grazing).  Synchronous. */
int ehss_solve_source(ehss_scene* scene, const ehss_source* source, const ehss_sobol5* sobol, int power, int cap,
                      ehss_totals* totals, int64_t* histogram, uint32_t* info, double* contrib);
/* The source's states as generated on the device (parity tests): origins,
 * directions (outgoing), incoming: n x 3; weights, owners (last atom, -1 for
 * external states), bounces (initial count), blocked (reverse PA; 0 otherwise):
 * n each, ray order; blocked states have weight 0.  Any output may be NULL. */
int ehss_source_states(ehss_scene* scene, const ehss_source* source, const ehss_sobol5* sobol, int power,
                       double* origins, double* directions, double* incoming, double* weights, int64_t* owners,
                       int64_t* bounces, uint8_t* blocked);

/* ---------------------------------------------------------------------------
 * Finite-grid PA: the CLI's adaptive method (ccs_cli.py:142-160; the union
 * count of pa_sparse_blocks.py:125-174 / pa_dyadic_bitmask.py:86-97).
 * ------------------------------------------------------------------------- */
/* For each of n_dirs directions: frames[6 d .. 6 d + 5] = the unit vectors u, v
 * of the projection plane (pa_walsh_geometry.py:173-176, computed by the
 * caller); centered = atom centres minus their mean.  The atoms are projected on
 * the device, the grid half-width is max(|x| + r, |y| + r)(1 + 1e-10), and
 * counts[d] gets the number of grid x grid midpoint pixels covered by the union
 * of the projected disks, spacing[d] (may be NULL) the pixel size: PA of the
 * direction = counts[d] spacing[d]^2.  grid: power of two in 8..8192. */
int ehss_pa_grid(int64_t n_atoms, const double* centered, const double* radii, int n_dirs, const double* frames,
                 int grid, int64_t* counts, double* spacing);
/* The covered-pixel count for caller-projected disks (xy: n x 2) on the grid
 * with lower corner lower[2] and pixel size spacing. */
int ehss_pa_grid_count(int64_t n_atoms, const double* xy, const double* radii, const double* lower, double spacing,
                       int grid, int64_t* count);

/* ---------------------------------------------------------------------------
 * Prepared multi-structure plans: many (scene, scramble, 2^power rays, cap)
 * jobs run as one launch sequence on the plan's own stream -- a 2-D packet grid
 * (blockIdx.y = job), one shared drain queue for the rays still bouncing after
 * the packet queries, and per-job fixed-order reductions -- optionally captured
 * once into a hipGraph (EHSS_FLAG_GRAPH) and replayed.  Each job's totals,
 * histogram and ledger equal those of ehss_solve on the same scene, scramble,
 * power, cap, precision and flags.  Replaces, for a set of structures, the
 * per-frame loop of the CLI (ccs_cli.py:305-307, one solve_frame per structure,
 * ccs_cli.py:170).  The plan copies each job's scene views at creation, so it
 * holds its scenes: ehss_scene_rebuild and ehss_scene_destroy fail with
 * EHSS_ERR_BUSY until ehss_plan_destroy.  Scenes must have no solve pending
 * while a plan launch runs.
 * ------------------------------------------------------------------------- */
typedef struct ehss_plan ehss_plan;

/* flags: EHSS_FLAG_PA, EHSS_FLAG_GRAPH, EHSS_VARIANT(v), EHSS_COOP(n),
 * EHSS_GROUP_LOG2(k), EHSS_STAGE1(q), EHSS_PACKETS(q); powers 0..26 each, at most 2^30 rays and
 * 65535 jobs per plan. */
int ehss_plan_create(int n_jobs, ehss_scene* const* scenes, const ehss_sobol* sobols, const int32_t* powers,
                     const int32_t* caps, int precision, unsigned flags, ehss_plan** out);
/* Re-points the plan at another job list (same precision and flags), reusing
 * its stream and grow-only device buffers: waits for the plan's last launch,
 * releases the scenes it held, holds the new ones, and drops a captured graph
 * (the next launch captures again).  n_jobs == 0 only releases the scenes.
 * For continuous batching: a server retargets a few plans to the replicates of
 * the requests ready at each launch. */
int ehss_plan_retarget(ehss_plan* plan, int n_jobs, ehss_scene* const* scenes, const ehss_sobol* sobols,
                       const int32_t* powers, const int32_t* caps);
/* Replaces the scramble of one job (used by the next launch, graph or not). */
int ehss_plan_set_sobol(ehss_plan* plan, int job, const ehss_sobol* sobol);
/* Asynchronous: enqueues the plan on its stream and returns. */
int ehss_plan_launch(ehss_plan* plan);
/* *done = 1 once the last launch has finished (never blocks). */
int ehss_plan_query(ehss_plan* plan, int* done);
/* Blocks until done; totals: n_jobs entries; histograms / ledgers (may be NULL):
 * concatenated (cap_j + 1)-bin arrays in job order.  kernel_ms is the device
 * time of the whole plan. */
int ehss_plan_wait(ehss_plan* plan, ehss_totals* totals, int64_t* histograms, double* ledgers);
int ehss_plan_destroy(ehss_plan* plan);

#ifdef __cplusplus
}
#endif

#endif /* EHSS_HIP_H */
