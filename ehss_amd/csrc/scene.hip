// scene.hip -- device-resident scenes: creation (on-device tree build), export,
// and tracing / PA queries of caller rays against them.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "ehss_internal.h"

namespace ehss {

void enclosing_disk(int64_t n, const double* centers, const double* radii, double center[3], double* radius) {
    double s[3] = {0, 0, 0};
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) s[k] += centers[3 * i + k];  // NumPy axis-0 add.reduce: sequential
    for (int k = 0; k < 3; ++k) center[k] = s[k] / (double)n;
    double best = -INFINITY;
    for (int64_t i = 0; i < n; ++i) {
        const double dx = centers[3 * i] - center[0], dy = centers[3 * i + 1] - center[1],
                     dz = centers[3 * i + 2] - center[2];
        const double v = sqrt((dx * dx + dy * dy) + dz * dz) + radii[i];
        best = std::max(best, v);  // exact, so the order of the maximum does not matter
    }
    *radius = best * (1. + 1e-9);
}

}  // namespace ehss

int64_t ehss_scene::device_bytes() const {
    const ehss::DevBuf* bufs[] = {&centers, &radii, &sph,   &groups, &offsets, &eids,    &nodes,   &nodes32,
                                  &sph32,   &pa,    &pa_ids, &ray_info, &ray_contrib, &partials, &hist,  &ledger,
                                  &counters, &queue, &frontier, &sphs, &leaves32, &ray_times, &tables, &queue2,
                                  &drain_stripes, &src_weights, &src_cdf, &pk_nodes32, &pk_offsets};
    int64_t total = 0;
    for (const ehss::DevBuf* b : bufs) total += (int64_t)b->bytes;
    return total;
}

using namespace ehss;

namespace {

// Roots of the subtrees one wave splits a cooperative query over (at most
// kFrontier): a function of the group count alone, since the static tree's
// topology is implicit in it.
int put_frontier(ehss_scene* sc, hipStream_t s) {
    const std::vector<int32_t> f = frontier_of(sc->ng);
    EHSS_TRY(sc->frontier.reserve(sizeof(int32_t) * kFrontier));
    EHSS_TRY(launch_put_frontier(f, sc->frontier.as<int32_t>(), s));
    sc->view.frontier = sc->frontier.as<int32_t>();
    sc->view.n_frontier = (int32_t)f.size();
    return EHSS_OK;
}

// Validates a structure and its optional groups (ehss_static.py:65-77); cg
// receives the groups as int32 and *custom_ng their count (0 without groups).
int check_structure(int64_t n, const double* centers, const double* radii, const int64_t* groups, int leaf_size,
                    std::vector<int32_t>& cg, int32_t* custom_ng) {
    EHSS_CHECK(n >= 1 && n < (1LL << 28), "atom count %lld out of range", (long long)n);
    EHSS_CHECK(centers && radii, "null centers/radii");
    EHSS_CHECK(leaf_size >= 1, "Positive integer leaf size required");
    for (int64_t i = 0; i < n; ++i) {
        EHSS_CHECK(isfinite(centers[3 * i]) && isfinite(centers[3 * i + 1]) && isfinite(centers[3 * i + 2]) &&
                       isfinite(radii[i]) && radii[i] > 0,
                   "Finite centers and positive finite radii required (atom %lld)", (long long)i);
    }
    *custom_ng = 0;
    cg.clear();
    if (groups) {
        // ehss_static.py:71-77: one integer group per atom, contiguous nonempty ids.
        int64_t gmax = -1;
        for (int64_t i = 0; i < n; ++i) {
            EHSS_CHECK(groups[i] >= 0 && groups[i] < n, "Contiguous nonempty groups required");
            gmax = std::max(gmax, groups[i]);
        }
        std::vector<char> seen((size_t)gmax + 1, 0);
        for (int64_t i = 0; i < n; ++i) seen[groups[i]] = 1;
        for (int64_t g = 0; g <= gmax; ++g) EHSS_CHECK(seen[g], "Contiguous nonempty groups required");
        cg.resize((size_t)n);
        for (int64_t i = 0; i < n; ++i) cg[i] = (int32_t)groups[i];
        *custom_ng = (int32_t)(gmax + 1);
    }
    return EHSS_OK;
}

// Builds (or rebuilds) the scene's trees and filter arrays for a structure on
// the scene's stream.  Every device array is grow-only, so a rebuild for a
// structure no larger than the scene's largest so far allocates nothing.
int build_into(ehss_scene* sc, int64_t n, const double* centers, const double* radii, const std::vector<int32_t>& cg,
               int32_t custom_ng, int leaf_size) {
    hipStream_t s = sc->stream;
    sc->cont_valid = false;  // the saved rays belong to the previous structure
    sc->n = (int32_t)n;
    sc->leaf_size = leaf_size;
    enclosing_disk(n, centers, radii, sc->center, &sc->radius);
    double scale = 1.5 * sc->radius;
    for (int64_t i = 0; i < n; ++i) {
        const double e = std::max({fabs(centers[3 * i] - sc->center[0]), fabs(centers[3 * i + 1] - sc->center[1]),
                                   fabs(centers[3 * i + 2] - sc->center[2])}) + radii[i];
        scale = std::max(scale, e);
    }
    sc->scale = scale;
    // A grow-only array that grows is freed and reallocated: let the previous
    // build's kernels (not synchronised) finish with it first.
    if (n > sc->cap_atoms) EHSS_HIP(hipStreamSynchronize(s));
    EHSS_TRY(sc->centers.reserve(sizeof(double) * 3 * n));
    EHSS_TRY(sc->radii.reserve(sizeof(double) * n));
    EHSS_TRY(sc->sph.reserve(sizeof(double4) * n));
    EHSS_TRY(sc->groups.reserve(sizeof(int32_t) * n));
    EHSS_TRY(sc->offsets.reserve(sizeof(int32_t) * (n + 1)));
    EHSS_TRY(sc->eids.reserve(sizeof(int32_t) * n));
    EHSS_TRY(sc->nodes.reserve(sizeof(Node64) * (2 * n)));
    EHSS_TRY(sc->pa.reserve(sizeof(NodePA) * (2 * n + 2)));
    EHSS_TRY(sc->pa_ids.reserve(sizeof(int32_t) * n));
    EHSS_HIP(hipMemcpyAsync(sc->centers.p, centers, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
    EHSS_HIP(hipMemcpyAsync(sc->radii.p, radii, sizeof(double) * n, hipMemcpyHostToDevice, s));
    if (custom_ng) {
        EHSS_TRY(sc->custom_groups.reserve(sizeof(int32_t) * n));
        EHSS_HIP(hipMemcpyAsync(sc->custom_groups.p, cg.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
    }
    // Without caller groups or two atoms sharing a key triple, the tree's shape
    // is known here and the build runs without synchronising (build_shaped.hip);
    // otherwise the level-by-level build reads its counts back.  EHSS_BUILD_PATH
    // = general forces the latter (for comparisons).
    const char* forced = getenv("EHSS_BUILD_PATH");
    const bool shaped = !custom_ng && leaf_size <= kShapedMaxLeaf && !(forced && !strcmp(forced, "general")) &&
                        !has_duplicate_keys(n, centers, radii);
    if (shaped) {
        const TreeShape shape = make_tree_shape((int32_t)n, leaf_size);
        double rmax = 0;
        for (int64_t i = 0; i < n; ++i) rmax = std::max(rmax, radii[i]);
        EHSS_TRY(build_scene_shaped(shape, sc->centers.as<double>(), sc->radii.as<double>(), rmax,
                                    sc->sph.as<double4>(), sc->groups.as<int32_t>(), sc->offsets.as<int32_t>(),
                                    sc->eids.as<int32_t>(), sc->nodes.as<Node64>(), sc->pa.as<NodePA>(),
                                    sc->pa_ids.as<int32_t>(), sc->build_work, s));
        sc->ng = shape.groups;
        sc->nn = 2 * shape.groups - 1;
        sc->pa_nn = 2 * shape.groups - 1;
        sc->max_depth = shape.max_depth;
    } else {
        BuiltScene built;
        EHSS_TRY(build_scene_device((int32_t)n, leaf_size, sc->centers.as<double>(), sc->radii.as<double>(),
                                    sc->sph.as<double4>(), sc->groups.as<int32_t>(), sc->offsets.as<int32_t>(),
                                    sc->eids.as<int32_t>(), sc->nodes.as<Node64>(), sc->pa.as<NodePA>(),
                                    sc->pa_ids.as<int32_t>(), custom_ng ? sc->custom_groups.as<int32_t>() : nullptr,
                                    custom_ng, &built, sc->build_work, s));
        sc->ng = built.n_groups;
        sc->nn = built.n_nodes;
        sc->pa_nn = built.pa_nodes;
        sc->max_depth = built.max_depth;
    }
    sc->build_path = shaped ? EHSS_BUILD_SHAPED : EHSS_BUILD_GENERAL;
    SceneView& v = sc->view;
    v = SceneView();
    v.n_atoms = sc->n;
    v.n_groups = sc->ng;
    v.n_nodes = sc->nn;
    v.pa_nodes = sc->pa_nn;
    v.nodes = sc->nodes.as<Node64>();
    v.offsets = sc->offsets.as<int32_t>();
    v.ids = sc->eids.as<int32_t>();
    v.sph = sc->sph.as<double4>();
    v.pa = sc->pa.as<NodePA>();
    v.pa_ids = sc->pa_ids.as<int32_t>();
    v.cx = sc->center[0];
    v.cy = sc->center[1];
    v.cz = sc->center[2];
    v.scale = sc->scale;
    set_filter_margins(v);
    EHSS_TRY(sc->nodes32.reserve(sizeof(Node32) * sc->nn));
    EHSS_TRY(sc->sph32.reserve(sizeof(float4) * n));
    EHSS_TRY(sc->sphs.reserve(sizeof(double4) * n));
    EHSS_TRY(sc->leaves32.reserve(sizeof(Node32) * sc->ng));
    EHSS_TRY(launch_make_filter(v, sc->nodes32.as<Node32>(), sc->sph32.as<float4>(), sc->sphs.as<double4>(),
                                sc->leaves32.as<Node32>(), s));
    v.nodes32 = sc->nodes32.as<Node32>();
    v.sph32 = sc->sph32.as<float4>();
    v.sphs = sc->sphs.as<double4>();
    v.leaves32 = sc->leaves32.as<Node32>();
    const int32_t cut = packet_cut(v.n_atoms, v.n_groups);
    if (cut > 1) {
        v.pk_nodes = cut_nodes(v.n_groups, cut);
        EHSS_TRY(sc->pk_nodes32.reserve(sizeof(Node32) * v.pk_nodes));
        EHSS_TRY(sc->pk_offsets.reserve(sizeof(int32_t) * (cut_leaves(v.n_groups, cut) + 1)));
        EHSS_TRY(launch_packet_tree(v, cut, sc->pk_nodes32.as<Node32>(), sc->pk_offsets.as<int32_t>(), s));
        v.pk_nodes32 = sc->pk_nodes32.as<Node32>();
        v.pk_offsets = sc->pk_offsets.as<int32_t>();
    } else {
        v.pk_nodes = v.n_nodes;
        v.pk_nodes32 = v.nodes32;
        v.pk_offsets = v.offsets;
    }
    EHSS_TRY(put_frontier(sc, s));
    // Work on other streams that reads the scene (plans, caller streams) waits for this.
    EHSS_HIP(hipEventRecord(sc->ev_built, s));
    sc->cap_atoms = std::max(sc->cap_atoms, n);
    // Per-ray records of an earlier structure no longer apply.
    sc->last_rays = 0;
    sc->last_times = 0;
    return EHSS_OK;
}

int create_scene(int64_t n, const double* centers, const double* radii, const int64_t* groups, int leaf_size,
                 ehss_scene** out) {
    EHSS_TRY(ensure_init());
    EHSS_CHECK(out, "null output handle");
    *out = nullptr;
    std::vector<int32_t> cg;
    int32_t custom_ng = 0;
    EHSS_TRY(check_structure(n, centers, radii, groups, leaf_size, cg, &custom_ng));
    ehss_scene* sc = new ehss_scene();
    struct Guard {
        ehss_scene*& p;
        ~Guard() {
            if (p) {
                if (p->ev0) (void)hipEventDestroy(p->ev0);
                if (p->ev1) (void)hipEventDestroy(p->ev1);
                if (p->ev_done) (void)hipEventDestroy(p->ev_done);
                if (p->ev_built) (void)hipEventDestroy(p->ev_built);
                if (p->stream) (void)hipStreamDestroy(p->stream);
                delete p;
            }
        }
    } guard{sc};
    EHSS_HIP(hipStreamCreateWithFlags(&sc->stream, hipStreamNonBlocking));
    EHSS_HIP(hipEventCreate(&sc->ev0));
    EHSS_HIP(hipEventCreate(&sc->ev1));
    EHSS_HIP(hipEventCreateWithFlags(&sc->ev_done, hipEventDisableTiming));
    EHSS_HIP(hipEventCreateWithFlags(&sc->ev_built, hipEventDisableTiming));
    EHSS_TRY(build_into(sc, n, centers, radii, cg, custom_ng, leaf_size));
    *out = sc;
    sc = nullptr;  // release the guard
    return EHSS_OK;
}

template <class T>
int download_as_i64(const ehss::DevBuf& b, int64_t count, int64_t* dst, hipStream_t s) {
    if (!dst || count == 0) return EHSS_OK;
    std::vector<T> tmp((size_t)count);
    EHSS_HIP(hipMemcpyAsync(tmp.data(), b.p, sizeof(T) * count, hipMemcpyDeviceToHost, s));
    EHSS_HIP(hipStreamSynchronize(s));
    for (int64_t i = 0; i < count; ++i) dst[i] = (int64_t)tmp[i];
    return EHSS_OK;
}

}  // namespace

extern "C" {

int ehss_scene_create(int64_t n, const double* centers, const double* radii, int leaf_size, ehss_scene** out) {
    return create_scene(n, centers, radii, nullptr, leaf_size, out);
}

int ehss_scene_create_grouped(int64_t n, const double* centers, const double* radii, const int64_t* groups,
                              int leaf_size, ehss_scene** out) {
    EHSS_CHECK(groups, "null groups");
    return create_scene(n, centers, radii, groups, leaf_size, out);
}

int ehss_scene_rebuild(ehss_scene* scene, int64_t n, const double* centers, const double* radii, int leaf_size) {
    EHSS_TRY(ensure_init());  // the calling thread (a serving worker) may not have selected the device
    EHSS_CHECK(scene, "null scene");
    std::lock_guard<std::mutex> guard(scene->lock);
    if (scene->pending) return set_error(EHSS_ERR_BUSY, "a solve is pending on this scene");
    if (const int held = scene->plans.load())
        return set_error(EHSS_ERR_BUSY, "the scene is held by %d plan job(s); destroy the plan(s) first", held);
    std::vector<int32_t> cg;
    int32_t custom_ng = 0;
    EHSS_TRY(check_structure(n, centers, radii, nullptr, leaf_size, cg, &custom_ng));
    EHSS_HIP(hipStreamSynchronize(scene->stream));
    return build_into(scene, n, centers, radii, cg, custom_ng, leaf_size);
}

int ehss_scene_destroy(ehss_scene* scene) {
    if (!scene) return EHSS_OK;
    {
        std::lock_guard<std::mutex> guard(scene->lock);  // waits for a call in progress
        if (const int held = scene->plans.load())
            return set_error(EHSS_ERR_BUSY, "the scene is held by %d plan job(s); destroy the plan(s) first", held);
    }
    if (scene->stream) (void)hipStreamSynchronize(scene->stream);
    if (scene->ev0) (void)hipEventDestroy(scene->ev0);
    if (scene->ev1) (void)hipEventDestroy(scene->ev1);
    if (scene->ev_done) (void)hipEventDestroy(scene->ev_done);
    if (scene->ev_built) (void)hipEventDestroy(scene->ev_built);
    if (scene->stream) (void)hipStreamDestroy(scene->stream);
    delete scene;
    return EHSS_OK;
}

int ehss_scene_get_info(const ehss_scene* scene, ehss_scene_info* info) {
    EHSS_CHECK(scene && info, "null scene/info");
    std::lock_guard<std::mutex> guard(scene->lock);
    info->atoms = scene->n;
    info->groups = scene->ng;
    info->nodes = scene->nn;
    info->pa_nodes = scene->pa_nn;
    info->leaf_size = scene->leaf_size;
    info->max_depth = scene->max_depth;
    for (int k = 0; k < 3; ++k) info->center[k] = scene->center[k];
    info->radius = scene->radius;
    info->scale = scene->scale;
    info->device_bytes = scene->device_bytes();
    info->build_path = scene->build_path;
    info->packet_nodes = scene->view.pk_nodes;
    return EHSS_OK;
}

int ehss_scene_export_tree(const ehss_scene* scene, double* lo, double* hi, int64_t* end, int64_t* leaf_group,
                           int64_t* offsets, int64_t* ids, int64_t* groups) {
    EHSS_CHECK(scene, "null scene");
    std::lock_guard<std::mutex> guard(scene->lock);
    hipStream_t s = scene->stream;
    std::vector<Node64> nodes((size_t)scene->nn);
    EHSS_HIP(hipMemcpyAsync(nodes.data(), scene->nodes.p, sizeof(Node64) * scene->nn, hipMemcpyDeviceToHost, s));
    EHSS_HIP(hipStreamSynchronize(s));
    for (int32_t k = 0; k < scene->nn; ++k) {
        for (int c = 0; c < 3; ++c) {
            if (lo) lo[3 * k + c] = nodes[k].lo[c];
            if (hi) hi[3 * k + c] = nodes[k].hi[c];
        }
        if (end) end[k] = nodes[k].end;
        if (leaf_group) leaf_group[k] = nodes[k].leaf;
    }
    EHSS_TRY(download_as_i64<int32_t>(scene->offsets, scene->ng + 1, offsets, s));
    EHSS_TRY(download_as_i64<int32_t>(scene->eids, scene->n, ids, s));
    EHSS_TRY(download_as_i64<int32_t>(scene->groups, scene->n, groups, s));
    return EHSS_OK;
}

int ehss_scene_export_pa_tree(const ehss_scene* scene, double* lo, double* hi, int64_t* end, int64_t* first,
                              int64_t* count, int64_t* ids) {
    EHSS_CHECK(scene, "null scene");
    std::lock_guard<std::mutex> guard(scene->lock);
    hipStream_t s = scene->stream;
    std::vector<NodePA> nodes((size_t)scene->pa_nn);
    EHSS_HIP(hipMemcpyAsync(nodes.data(), scene->pa.p, sizeof(NodePA) * scene->pa_nn, hipMemcpyDeviceToHost, s));
    EHSS_HIP(hipStreamSynchronize(s));
    for (int32_t k = 0; k < scene->pa_nn; ++k) {
        for (int c = 0; c < 3; ++c) {
            if (lo) lo[3 * k + c] = nodes[k].lo[c];
            if (hi) hi[3 * k + c] = nodes[k].hi[c];
        }
        if (end) end[k] = nodes[k].end;
        if (first) first[k] = nodes[k].first;
        if (count) count[k] = nodes[k].count;
    }
    EHSS_TRY(download_as_i64<int32_t>(scene->pa_ids, scene->n, ids, s));
    return EHSS_OK;
}

int ehss_scene_trace(ehss_scene* scene, int64_t nr, int cap, int precision, int record, double* p, double* d,
                     const int64_t* lasts, int64_t* counts, uint8_t* escaped, uint8_t* unresolved,
                     int32_t* history, int64_t* counters) {
    EHSS_TRY(ensure_init());
    EHSS_CHECK(scene, "null scene");
    std::lock_guard<std::mutex> guard(scene->lock);
    EHSS_CHECK(nr == 0 || (p && d && lasts && counts && escaped && unresolved), "null ray buffer");
    for (int64_t r = 0; r < nr; ++r)
        EHSS_CHECK(lasts[r] < scene->n, "Invalid initial physical identity (ray %lld)", (long long)r);
    SceneView v = scene->view;
    if (precision == EHSS_PRECISION_FP32_FIXUP) {
        // The filter margins assume ray origins within the scene scale; widen
        // the sphere margins for far rays, and use the FP64 kernel when the box
        // inflation baked into the FP32 tree would no longer cover them.
        double far = 0;
        for (int64_t r = 0; r < nr; ++r)
            far = std::max({far, fabs(p[3 * r] - v.cx), fabs(p[3 * r + 1] - v.cy), fabs(p[3 * r + 2] - v.cz)});
        if (!(far <= 2 * v.scale)) {
            precision = EHSS_PRECISION_FP64;
        } else if (far > v.scale) {
            const float keep_box = v.m_box;
            v.scale = far;
            set_filter_margins(v);
            v.m_box = keep_box;
        }
    }
    int64_t k[3] = {0, 0, 0};
    EHSS_TRY(trace_host_rays(v, nr, cap, precision, record, p, d, lasts, counts, escaped, unresolved, history, k,
                             scene->stream));
    if (counters) {
        counters[0] = k[0];
        counters[1] = k[1];
        if (precision == EHSS_PRECISION_FP32_FIXUP) counters[2] = k[2];
    }
    return EHSS_OK;
}

int ehss_scene_pa_query(ehss_scene* scene, int64_t nr, int precision, const double* origins,
                        const double* directions, const int64_t* owners, uint8_t* hits, int64_t* counters) {
    EHSS_TRY(ensure_init());
    EHSS_CHECK(scene, "null scene");
    std::lock_guard<std::mutex> guard(scene->lock);
    EHSS_CHECK(nr >= 0 && (nr == 0 || (origins && directions && hits)), "null ray buffer");
    EHSS_CHECK(precision == EHSS_PRECISION_FP64 || precision == EHSS_PRECISION_FP32_FIXUP, "unknown precision");
    hipStream_t s = scene->stream;
    DevBuf bo, bd, bw, bh, bk;
    EHSS_TRY(bk.alloc(sizeof(unsigned long long) * 2));
    EHSS_HIP(hipMemsetAsync(bk.p, 0, sizeof(unsigned long long) * 2, s));
    if (nr > 0) {
        const size_t v3 = sizeof(double) * 3 * nr;
        EHSS_TRY(bo.alloc(v3));
        EHSS_TRY(bd.alloc(v3));
        EHSS_TRY(bh.alloc(nr));
        EHSS_HIP(hipMemcpyAsync(bo.p, origins, v3, hipMemcpyHostToDevice, s));
        EHSS_HIP(hipMemcpyAsync(bd.p, directions, v3, hipMemcpyHostToDevice, s));
        PAIO io;
        io.nr = nr;
        io.o = bo.as<double>();
        io.d = bd.as<double>();
        if (owners) {
            EHSS_TRY(bw.alloc(sizeof(int64_t) * nr));
            EHSS_HIP(hipMemcpyAsync(bw.p, owners, sizeof(int64_t) * nr, hipMemcpyHostToDevice, s));
            io.owners = bw.as<int64_t>();
        }
        io.hits = bh.as<uint8_t>();
        io.counters = bk.as<unsigned long long>();
        EHSS_TRY(launch_pa_query(scene->view, io, s));
        EHSS_HIP(hipMemcpyAsync(hits, bh.p, nr, hipMemcpyDeviceToHost, s));
    }
    unsigned long long k[2];
    EHSS_HIP(hipMemcpyAsync(k, bk.p, sizeof k, hipMemcpyDeviceToHost, s));
    EHSS_HIP(hipStreamSynchronize(s));
    if (counters) {
        counters[0] = (int64_t)k[0];
        counters[1] = (int64_t)k[1];
    }
    return EHSS_OK;
}

}  // extern "C"
