// ehss_internal.h -- host-side declarations shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "ehss_device.h"
#include "ehss_trig.h"
#include "../../include/ehss_hip.h"

namespace ehss {

// ---------------------------------------------------------------------------
// Errors: every entry point returns a status; the message is thread-local.
// ---------------------------------------------------------------------------
int set_error(int code, const char* fmt, ...);
int hip_error(hipError_t e, const char* what, const char* file, int line);

#define EHSS_HIP(expr)                                                       \
    do {                                                                     \
        hipError_t e_ = (expr);                                              \
        if (e_ != hipSuccess) return ::ehss::hip_error(e_, #expr, __FILE__, __LINE__); \
    } while (0)

#define EHSS_CHECK(cond, ...)                                         \
    do {                                                              \
        if (!(cond)) return ::ehss::set_error(EHSS_ERR_INVALID, __VA_ARGS__); \
    } while (0)

#define EHSS_TRY(expr)            \
    do {                          \
        int s_ = (expr);          \
        if (s_ != EHSS_OK) return s_; \
    } while (0)

int ensure_init();           // selects the device, loads the trig table, creates the library stream
hipStream_t lib_stream();
// Host-libm compatibility table of the Sobol angles on the device (null
// pointers when no table file was found: cos/sin then stay correctly rounded).
ehss_trig::CompatTable trig_compat();
// Size of the allocation-free solve's static work array (ehss_solve_noalloc),
// and its device address on the current device (null before ensure_init).
constexpr size_t kNoallocScratchBytes = (size_t)EHSS_NOALLOC_SCRATCH_MB << 20;
void* noalloc_scratch();

// Owning device buffer.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    int alloc(size_t n);     // exact size; releases the previous allocation
    int reserve(size_t n);   // grow-only
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Pinned host buffer (hipHostMalloc), grow-only; device-to-host copies into it
// are asynchronous.
struct HostBuf {
    void* p = nullptr;
    size_t bytes = 0;
    HostBuf() = default;
    HostBuf(const HostBuf&) = delete;
    HostBuf& operator=(const HostBuf&) = delete;
    ~HostBuf() { release(); }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
    int reserve(size_t n);
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// ---------------------------------------------------------------------------
// Device view of a scene (everything the kernels read).
// ---------------------------------------------------------------------------
struct SceneView {
    int32_t n_atoms = 0, n_groups = 0, n_nodes = 0, pa_nodes = 0;
    const Node64* nodes = nullptr;    // static tree, FP64 (reference-identical)
    const Node32* nodes32 = nullptr;  // static tree, FP32 filter boxes
    const int32_t* offsets = nullptr; // group -> first slot
    const int32_t* ids = nullptr;     // slot -> atom
    const double4* sph = nullptr;     // atom -> {x, y, z, r}, FP64
    const float4* sph32 = nullptr;    // slot -> {x - Cx, y - Cy, z - Cz, r}, FP32
    const double4* sphs = nullptr;    // slot -> sph[ids[slot]] (FP64, slot order)
    const Node32* leaves32 = nullptr; // group -> FP32 filter box of its leaf; end = first slot, leaf = slots
    const NodePA* pa = nullptr;       // PA tree
    const int32_t* pa_ids = nullptr;  // PA id slots
    const int32_t* frontier = nullptr; // roots of <= kFrontier disjoint subtrees covering every leaf
    int32_t n_frontier = 0;
    // Packet tree (the static tree cut at packet_cut groups, FP32 boxes, the
    // nodes32 encoding): the tree the packet queries walk; its leaf j holds
    // slots pk_offsets[j] .. pk_offsets[j + 1] - 1.  Equal to nodes32 / offsets
    // when the cut is 1.
    const Node32* pk_nodes32 = nullptr;
    const int32_t* pk_offsets = nullptr;
    int32_t pk_nodes = 0;
    double cx = 0, cy = 0, cz = 0;    // filter origin (scene centre)
    double scale = 1;                 // bound on |coordinates - centre| and radii
    float m_box = 0, m_p = 0, m_disc = 0;  // FP32 filter margins
};

// Filter margins for a scale s (FP32 unit roundoff u = 2^-24):
// box inflation and projection margin 64 u s, discriminant margin 1024 u s^2.
void set_filter_margins(SceneView& v);

// Builds the FP32 filter arrays (nodes32, sph32) from the FP64 ones, on the device.
// Also writes the slot-ordered FP64 spheres and, when leaves32 is not null, the
// per-group leaf boxes.
int launch_make_filter(const SceneView& v, Node32* nodes32, float4* sph32, double4* sphs, Node32* leaves32,
                       hipStream_t s);

// Packet tree of a scene whose nodes32 and offsets are built (cut > 1):
// cut_nodes(n_groups, cut) nodes and cut_leaves(n_groups, cut) + 1 offsets.
int launch_packet_tree(const SceneView& v, int32_t cut, Node32* pk_nodes32, int32_t* pk_offsets, hipStream_t s);

// Per-ray trace of caller rays (AoS p, d as in ehss_trace).
struct TraceIO {
    int64_t nr = 0;
    int32_t cap = 0;
    int32_t record = 0;
    double* p = nullptr;
    double* d = nullptr;
    const int64_t* lasts = nullptr;
    int64_t* counts = nullptr;
    uint8_t* escaped = nullptr;
    uint8_t* unresolved = nullptr;
    int32_t* history = nullptr;
    unsigned long long* counters = nullptr;  // [boxes, spheres, fp64 retests]
};
int launch_trace(const SceneView& v, const TraceIO& io, int precision, hipStream_t s);

struct PAIO {
    int64_t nr = 0;
    const double* o = nullptr;
    const double* d = nullptr;
    const int64_t* owners = nullptr;
    uint8_t* hits = nullptr;
    unsigned long long* counters = nullptr;  // [boxes, tests]
};
int launch_pa_query(const SceneView& v, const PAIO& io, hipStream_t s);

// Uploads caller rays, traces them against `view`, downloads the outputs.
int trace_host_rays(const SceneView& view, int64_t nr, int cap, int precision, int record, double* p, double* d,
                    const int64_t* lasts, int64_t* counts, uint8_t* escaped, uint8_t* unresolved,
                    int32_t* history, int64_t* counters_out, hipStream_t s);

struct BuiltScene {
    int32_t n_groups = 0, n_nodes = 0, pa_nodes = 0, max_depth = 0;
};
// Device work arrays of a tree build, kept by the scene (grow-only) so that a
// rebuild for a new conformer allocates and frees nothing.
struct BuildWork {
    DevBuf keys, ids2, pseg, pseg2, skey, rank, seg, ctl, flag, ndepth, nright, leaf_node, pabox;
};
int build_scene_device(int32_t n, int32_t leaf_size, const double* d_centers, const double* d_radii, double4* d_sph,
                       int32_t* d_groups, int32_t* d_offsets, int32_t* d_eids, Node64* d_nodes, NodePA* d_pa,
                       int32_t* d_pa_ids, const int32_t* d_custom_groups, int32_t custom_ng, BuiltScene* out,
                       BuildWork& work, hipStream_t s);

// Shape of the median-split segment tree (ehss_static.py:16-50) of n atoms
// when no segment has zero key span: a segment of size s at depth d splits iff
// s > leaf and d < kMaxSplitDepth, into sizes s / 2 and s - s / 2.  Every
// segment at depth d then has size q[d] = n >> d or q[d] + 1 (halving q or
// q + 1 gives q / 2 or q / 2 + 1), so leaves[d][t], the number of leaves under
// a segment of size q[d] + t at depth d, gives the size of every subtree.
constexpr int kShapeDepths = 32;
constexpr int kMaxSplitDepth = 24;
struct TreeShape {
    int32_t n, leaf, groups, max_depth;
    int32_t q[kShapeDepths];
    int32_t leaves[kShapeDepths][2];
};
TreeShape make_tree_shape(int32_t n, int32_t leaf);
// True when two atoms have the same key triple ((c - r) + (c + r)) / 2: only
// then can a segment of two or more atoms have zero span on every axis.
bool has_duplicate_keys(int64_t n, const double* centers, const double* radii);
// Largest leaf size the shaped build takes (its leaves are ranked in one thread).
constexpr int32_t kShapedMaxLeaf = 32;
// The build for structures without duplicate key triples: every count is
// known from `shape`, so it enqueues its kernels on s and returns without
// synchronising.  Same outputs as build_scene_device without caller groups.
int build_scene_shaped(const TreeShape& shape, const double* d_centers, const double* d_radii, double rmax,
                       double4* d_sph, int32_t* d_groups, int32_t* d_offsets, int32_t* d_eids, Node64* d_nodes,
                       NodePA* d_pa, int32_t* d_pa_ids, BuildWork& work, hipStream_t s);

// Roots of the cooperative-query subtrees of the static tree over ng groups
// (its topology is implicit in ng: ehss_cpu_geometry.py:19-38): starting from
// the root, every internal node is replaced by its two children while the list
// stays within kFrontier entries.
std::vector<int32_t> frontier_of(int32_t ng);
// Writes a frontier list (at most kFrontier entries) to device memory on s.
int launch_put_frontier(const std::vector<int32_t>& f, int32_t* dst, hipStream_t s);

// Impact disk of ehss_state.py:46-47 computed on the host in NumPy's order:
// center = sequential sum / n, radius = max(sqrt((dx^2 + dy^2) + dz^2) + r) * (1 + 1e-9).
void enclosing_disk(int64_t n, const double* centers, const double* radii, double center[3], double* radius);

// Per-ray records of a solve (ray-index order).
struct RayRecords {
    uint32_t* info = nullptr;   // bounces (low 20 bits) | flags << 20
    double* contrib = nullptr;  // w * angle for escaped rays, else 0
};

constexpr uint32_t kBounceMask = 0xFFFFFu;

int64_t drain_blocks(int64_t n);
int launch_zero(unsigned long long* p, int64_t n, hipStream_t s);

// Wave-coherent ray map columns for a scrambled net of 2^m points (host).
void make_ray_map(const ehss_sobol& sob, int m, uint32_t cols[32]);
constexpr int kFlagShift = 20;

}  // namespace ehss

// Device-resident scene (opaque handle of the C-ABI).
struct ehss_scene {
    int32_t n = 0, ng = 0, nn = 0, pa_nn = 0, leaf_size = 4, max_depth = 0;
    double center[3] = {0, 0, 0};
    double radius = 0, scale = 1;
    ehss::DevBuf centers, radii, sph, groups, offsets, eids, nodes, nodes32, sph32, sphs, leaves32, pa, pa_ids,
        frontier, pk_nodes32, pk_offsets;
    ehss::SceneView view;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_done = nullptr;
    // recorded after each (re)build: launches on other streams wait for it
    hipEvent_t ev_built = nullptr;
    int build_path = 0;     // EHSS_BUILD_SHAPED or EHSS_BUILD_GENERAL
    int64_t cap_atoms = 0;  // largest structure the grow-only arrays were sized for
    // Serialises the entry points that use this scene's buffers: calls on one
    // scene from several threads wait for each other instead of racing (a
    // rebuild may reallocate arrays another call's kernels still read).
    mutable std::mutex lock;
    // prepared plans holding this scene's views (ehss_plan_create counts one per
    // job): rebuild and destroy are refused while nonzero
    std::atomic<int> plans{0};
    // pending asynchronous solve (at most one per scene)
    bool pending = false;
    int p_reps = 1;
    int64_t p_rays = 0;
    int p_cap = 0;
    unsigned p_flags = 0;
    double p_weight = 0;
    ehss::HostBuf host_words;  // counters, histograms and ledgers of the pending launch
    ehss::HostBuf host_table;  // scrambles and ray maps of the pending launch (staging)
    // solve workspace (grow-only)
    ehss::DevBuf ray_info, ray_contrib, partials, hist, ledger, counters, queue, queue2, ray_times, tables,
        drain_stripes, custom_groups, sort_bins, sort_order;
    ehss::DevBuf src_weights, src_cdf;  // alternative-source solves: per-ray weights, searchsorted table
    ehss::BuildWork build_work;  // tree-build work arrays, reused by ehss_scene_rebuild
    // continuation of the last solve (EHSS_FLAG_CONTINUABLE): two queues of
    // saved unresolved rays (ping-pong) and their words {count, head} x 2
    ehss::DevBuf cont_q[2], cont_words;
    bool cont_valid = false;
    int cont_which = 0, cont_cap = 0, cont_reps = 0, cont_power = 0;
    unsigned cont_flags = 0;
    int64_t last_rays = 0, last_times = 0;
    uint32_t last_map[32] = {0};  // ray map of the last solve (slot -> ray index)
    int last_power = 0;
    uint32_t p_map[32] = {0};
    int p_power = 0;
    double last_weight = 0;
    int64_t device_bytes() const;
};
