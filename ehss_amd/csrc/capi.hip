// capi.hip -- C-ABI entry points of libehss_hip.so: errors, device set-up, and
// the reference's ehss_trace seam (src/native/ehss_cpu.cpp:42-46).
#include <dlfcn.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "ehss_internal.h"

namespace ehss {

namespace {
thread_local std::string g_error;
std::mutex g_init_mutex;
int g_device = -1;
hipStream_t g_stream = nullptr;

// Compatibility table (ehss_trig.h): one device allocation, kept until the
// process ends.
void* g_trig_mem = nullptr;
ehss_trig::CompatTable g_trig{nullptr, nullptr, 0.0};
int64_t g_trig_entries = -1;
std::string g_trig_libm, g_trig_path;

struct TrigFileHeader {
    char magic[8];
    uint32_t key_bits, shift, buckets, entries, not_sure, libm_ulp2;
    double lookup_from;
    char libm[24];
};
static_assert(sizeof(TrigFileHeader) == 64, "trig table header is 64 bytes");

std::string library_dir() {
    Dl_info info;
    if (dladdr((void*)&ensure_init, &info) && info.dli_fname) {
        std::string p = info.dli_fname;
        const size_t cut = p.rfind('/');
        if (cut != std::string::npos) return p.substr(0, cut);
    }
    return ".";
}

// Reads the table written by ehss_trig_compat_generate (trig_host.cpp) from
// $EHSS_TRIG_COMPAT or trig_compat.bin beside this library.  A missing file
// leaves the table off; a malformed one is an error.  Caller holds the init lock.
int load_trig_compat() {
    const char* env = getenv("EHSS_TRIG_COMPAT");
    g_trig_path = env ? env : library_dir() + "/trig_compat.bin";
    FILE* f = g_trig_path.empty() ? nullptr : fopen(g_trig_path.c_str(), "rb");
    if (!f) return EHSS_OK;
    TrigFileHeader h;
    std::vector<uint32_t> off;
    std::vector<uint16_t> ent;
    bool ok = fread(&h, sizeof h, 1, f) == 1 && memcmp(h.magic, "EHSSTRG2", 8) == 0 && h.key_bits == 30 &&
              h.shift == (uint32_t)ehss_trig::kCompatShift && h.buckets == ehss_trig::kCompatBuckets &&
              h.libm_ulp2 == 0 && h.lookup_from > 0.0 && h.lookup_from <= 0.5;
    if (ok) {
        off.resize((size_t)h.buckets + 1);
        ent.resize(h.entries);
        ok = fread(off.data(), sizeof(uint32_t), off.size(), f) == off.size() &&
             (ent.empty() || fread(ent.data(), sizeof(uint16_t), ent.size(), f) == ent.size()) &&
             fgetc(f) == EOF;
    }
    fclose(f);
    for (size_t b = 0; ok && b < h.buckets; ++b) {
        ok = off[b] <= off[b + 1] && off[b + 1] <= h.entries;
        for (uint32_t k = off[b]; ok && k + 1 < off[b + 1]; ++k) ok = (ent[k] >> 4) < (ent[k + 1] >> 4);
    }
    ok = ok && off[0] == 0 && off[h.buckets] == h.entries;
    if (!ok) return set_error(EHSS_ERR_INVALID, "malformed trig compatibility table %s", g_trig_path.c_str());
    const size_t off_bytes = (sizeof(uint32_t) * off.size() + 255) & ~(size_t)255;
    const size_t bytes = off_bytes + sizeof(uint16_t) * ent.size();
    EHSS_HIP(hipMalloc(&g_trig_mem, bytes));
    EHSS_HIP(hipMemcpy(g_trig_mem, off.data(), sizeof(uint32_t) * off.size(), hipMemcpyHostToDevice));
    if (!ent.empty())
        EHSS_HIP(hipMemcpy((char*)g_trig_mem + off_bytes, ent.data(), sizeof(uint16_t) * ent.size(),
                           hipMemcpyHostToDevice));
    g_trig.offsets = (const uint32_t*)g_trig_mem;
    g_trig.entries = (const uint16_t*)((char*)g_trig_mem + off_bytes);
    g_trig.lookup_from = h.lookup_from;
    g_trig_entries = h.entries;
    h.libm[sizeof h.libm - 1] = 0;
    g_trig_libm = h.libm;
    return EHSS_OK;
}

// Work memory of the allocation-free solve (ehss_solve_noalloc): a static device
// array, one copy per device, held until the process ends.  Host code never
// passes `scratch` itself to a kernel (its host-side name is not a device
// address): ensure_init looks up each device's address once with
// hipGetSymbolAddress and noalloc_scratch() hands that out.
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];
constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices] = {};

__global__ void k_sobol_sincos(int64_t n, const uint32_t* q, ehss_trig::CompatTable t, double* s, double* c) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = 6.283185307179586 * ((double)q[i] * 9.313225746154785e-10);
    ehss_trig::sincos_sobol(x, q[i], t, &s[i], &c[i]);
}
}  // namespace

int set_error(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

int hip_error(hipError_t e, const char* what, const char* file, int line) {
    const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? EHSS_ERR_NOMEM : EHSS_ERR_HIP;
    return set_error(code, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
}

int DevBuf::alloc(size_t n) {
    release();
    if (n == 0) return EHSS_OK;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
        p = nullptr;
        return hip_error(e, "hipMalloc", __FILE__, __LINE__);
    }
    bytes = n;
    return EHSS_OK;
}

int DevBuf::reserve(size_t n) {
    if (n <= bytes && p) return EHSS_OK;
    return alloc(n);
}

int HostBuf::reserve(size_t n) {
    if (n <= bytes && p) return EHSS_OK;
    release();
    if (n == 0) return EHSS_OK;
    hipError_t e = hipHostMalloc(&p, n, hipHostMallocDefault);
    if (e != hipSuccess) {
        p = nullptr;
        return hip_error(e, "hipHostMalloc", __FILE__, __LINE__);
    }
    bytes = n;
    return EHSS_OK;
}

int ensure_init() {
    std::lock_guard<std::mutex> lock(g_init_mutex);
    if (g_stream) return hipSetDevice(g_device) == hipSuccess ? EHSS_OK : set_error(EHSS_ERR_HIP, "hipSetDevice");
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0)
        return set_error(EHSS_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(e));
    if (g_device < 0) g_device = 0;
    EHSS_HIP(hipSetDevice(g_device));
    if (!g_trig_mem) EHSS_TRY(load_trig_compat());
    EHSS_CHECK(g_device < kMaxDevices, "device %d out of range", g_device);
    if (!g_scratch[g_device]) {
        void* p = nullptr;
        EHSS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[g_device] = p;
    }
    EHSS_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    return EHSS_OK;
}

hipStream_t lib_stream() { return g_stream; }

ehss_trig::CompatTable trig_compat() { return g_trig; }

void* noalloc_scratch() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    return g_scratch[dev];
}

namespace {

// Host-side description of a caller tree (ehss_trace arrays) on the device.
struct CompatScene {
    DevBuf nodes, nodes32, offsets, ids, sph, sph32, sphs;
    SceneView view;
};

int upload_compat(CompatScene& cs, int64_t nn, const double* lo, const double* hi, const int64_t* end,
                  const int64_t* leaf, const int64_t* offsets, const int64_t* ids, const double* world,
                  const double* radii, bool fast, const double* p, int64_t nr, hipStream_t s) {
    EHSS_CHECK(nn >= 1 && nn < INT32_MAX, "node count %lld out of range", (long long)nn);
    int64_t ng = 0;
    for (int64_t k = 0; k < nn; ++k) {
        EHSS_CHECK(end[k] > k && end[k] <= nn, "end[%lld]=%lld is not a preorder skip pointer", (long long)k,
                   (long long)end[k]);
        EHSS_CHECK(leaf[k] >= -1 && leaf[k] < INT32_MAX, "leaf[%lld] out of range", (long long)k);
        ng = std::max(ng, leaf[k] + 1);
    }
    EHSS_CHECK(offsets[0] == 0, "offsets[0] must be 0");
    for (int64_t g = 0; g < ng; ++g)
        EHSS_CHECK(offsets[g + 1] >= offsets[g], "offsets must be nondecreasing");
    const int64_t slots = offsets[ng];
    EHSS_CHECK(slots < INT32_MAX, "too many atoms");
    int64_t natoms = 0;
    for (int64_t k = 0; k < slots; ++k) {
        EHSS_CHECK(ids[k] >= 0 && ids[k] < INT32_MAX, "ids[%lld] out of range", (long long)k);
        natoms = std::max(natoms, ids[k] + 1);
    }
    std::vector<Node64> hn((size_t)nn);
    for (int64_t k = 0; k < nn; ++k) {
        for (int c = 0; c < 3; ++c) {
            hn[k].lo[c] = lo[3 * k + c];
            hn[k].hi[c] = hi[3 * k + c];
        }
        hn[k].end = (int32_t)end[k];
        hn[k].leaf = (int32_t)leaf[k];
        hn[k].pad[0] = hn[k].pad[1] = 0;
    }
    std::vector<int32_t> hoff((size_t)ng + 1), hids((size_t)std::max<int64_t>(slots, 1));
    for (int64_t g = 0; g <= ng; ++g) hoff[g] = (int32_t)offsets[g];
    for (int64_t k = 0; k < slots; ++k) hids[k] = (int32_t)ids[k];
    std::vector<double4> hs((size_t)std::max<int64_t>(natoms, 1));
    double cx = 0, cy = 0, cz = 0;
    for (int64_t a = 0; a < natoms; ++a) {
        hs[a] = make_double4(world[3 * a], world[3 * a + 1], world[3 * a + 2], radii[a]);
        cx += world[3 * a];
        cy += world[3 * a + 1];
        cz += world[3 * a + 2];
    }
    SceneView& v = cs.view;
    v.n_atoms = (int32_t)slots;
    v.n_groups = (int32_t)ng;
    v.n_nodes = (int32_t)nn;
    if (natoms > 0) {
        cx /= natoms;
        cy /= natoms;
        cz /= natoms;
    }
    v.cx = cx;
    v.cy = cy;
    v.cz = cz;
    double scale = 1.0;
    for (int64_t a = 0; a < natoms; ++a) {
        const double e = std::max({fabs(hs[a].x - cx), fabs(hs[a].y - cy), fabs(hs[a].z - cz)}) + hs[a].w;
        if (isfinite(e)) scale = std::max(scale, e);
    }
    for (int64_t r = 0; r < nr; ++r) {
        const double e = std::max({fabs(p[3 * r] - cx), fabs(p[3 * r + 1] - cy), fabs(p[3 * r + 2] - cz)});
        if (isfinite(e)) scale = std::max(scale, e);
    }
    v.scale = scale;
    set_filter_margins(v);
    EHSS_TRY(cs.nodes.alloc(sizeof(Node64) * nn));
    EHSS_TRY(cs.offsets.alloc(sizeof(int32_t) * hoff.size()));
    EHSS_TRY(cs.ids.alloc(sizeof(int32_t) * hids.size()));
    EHSS_TRY(cs.sph.alloc(sizeof(double4) * hs.size()));
    EHSS_HIP(hipMemcpyAsync(cs.nodes.p, hn.data(), sizeof(Node64) * nn, hipMemcpyHostToDevice, s));
    EHSS_HIP(hipMemcpyAsync(cs.offsets.p, hoff.data(), sizeof(int32_t) * hoff.size(), hipMemcpyHostToDevice, s));
    EHSS_HIP(hipMemcpyAsync(cs.ids.p, hids.data(), sizeof(int32_t) * hids.size(), hipMemcpyHostToDevice, s));
    EHSS_HIP(hipMemcpyAsync(cs.sph.p, hs.data(), sizeof(double4) * hs.size(), hipMemcpyHostToDevice, s));
    v.nodes = cs.nodes.as<Node64>();
    v.offsets = cs.offsets.as<int32_t>();
    v.ids = cs.ids.as<int32_t>();
    v.sph = cs.sph.as<double4>();
    if (fast) {
        EHSS_TRY(cs.nodes32.alloc(sizeof(Node32) * nn));
        EHSS_TRY(cs.sph32.alloc(sizeof(float4) * std::max<int64_t>(slots, 1)));
        EHSS_TRY(cs.sphs.alloc(sizeof(double4) * std::max<int64_t>(slots, 1)));
        EHSS_TRY(launch_make_filter(v, cs.nodes32.as<Node32>(), cs.sph32.as<float4>(), cs.sphs.as<double4>(),
                                    nullptr, s));
        v.nodes32 = cs.nodes32.as<Node32>();
        v.pk_nodes32 = v.nodes32;
        v.pk_offsets = v.offsets;
        v.pk_nodes = v.n_nodes;
        v.sph32 = cs.sph32.as<float4>();
        v.sphs = cs.sphs.as<double4>();
    }
    // Hold the staging vectors until the copies are done.
    EHSS_HIP(hipStreamSynchronize(s));
    return EHSS_OK;
}

}  // namespace

// Uploads caller rays, runs the trace on `view`, downloads the outputs.
int trace_host_rays(const SceneView& view, int64_t nr, int cap, int precision, int record, double* p, double* d,
                    const int64_t* lasts, int64_t* counts, uint8_t* escaped, uint8_t* unresolved,
                    int32_t* history, int64_t* counters_out, hipStream_t s) {
    EHSS_CHECK(nr >= 0, "negative ray count");
    EHSS_CHECK(cap >= 1, "cap must be >= 1");
    EHSS_CHECK(precision == EHSS_PRECISION_FP64 || precision == EHSS_PRECISION_FP32_FIXUP,
               "unknown precision %d", precision);
    EHSS_CHECK(!record || history, "history buffer required when record != 0");
    for (int64_t r = 0; r < nr; ++r) {
        EHSS_CHECK(lasts[r] >= -1 && lasts[r] < INT32_MAX, "last atom of ray %lld out of range", (long long)r);
        EHSS_CHECK(counts[r] >= 0, "negative initial bounce count for ray %lld", (long long)r);
    }
    DevBuf bp, bd, bl, bc, be, bu, bh, bk;
    const size_t v3 = sizeof(double) * 3 * (size_t)nr;
    const size_t hist = record ? sizeof(int32_t) * (size_t)nr * (size_t)cap : 0;
    EHSS_TRY(bk.alloc(sizeof(unsigned long long) * 4));
    EHSS_HIP(hipMemsetAsync(bk.p, 0, sizeof(unsigned long long) * 4, s));
    if (nr > 0) {
        EHSS_TRY(bp.alloc(v3));
        EHSS_TRY(bd.alloc(v3));
        EHSS_TRY(bl.alloc(sizeof(int64_t) * nr));
        EHSS_TRY(bc.alloc(sizeof(int64_t) * nr));
        EHSS_TRY(be.alloc(nr));
        EHSS_TRY(bu.alloc(nr));
        if (hist) EHSS_TRY(bh.alloc(hist));
        EHSS_HIP(hipMemcpyAsync(bp.p, p, v3, hipMemcpyHostToDevice, s));
        EHSS_HIP(hipMemcpyAsync(bd.p, d, v3, hipMemcpyHostToDevice, s));
        EHSS_HIP(hipMemcpyAsync(bl.p, lasts, sizeof(int64_t) * nr, hipMemcpyHostToDevice, s));
        EHSS_HIP(hipMemcpyAsync(bc.p, counts, sizeof(int64_t) * nr, hipMemcpyHostToDevice, s));
        if (hist) EHSS_HIP(hipMemcpyAsync(bh.p, history, hist, hipMemcpyHostToDevice, s));
        TraceIO io;
        io.nr = nr;
        io.cap = cap;
        io.record = record;
        io.p = bp.as<double>();
        io.d = bd.as<double>();
        io.lasts = bl.as<int64_t>();
        io.counts = bc.as<int64_t>();
        io.escaped = be.as<uint8_t>();
        io.unresolved = bu.as<uint8_t>();
        io.history = hist ? bh.as<int32_t>() : nullptr;
        io.counters = bk.as<unsigned long long>();
        EHSS_TRY(launch_trace(view, io, precision, s));
        EHSS_HIP(hipMemcpyAsync(p, bp.p, v3, hipMemcpyDeviceToHost, s));
        EHSS_HIP(hipMemcpyAsync(d, bd.p, v3, hipMemcpyDeviceToHost, s));
        EHSS_HIP(hipMemcpyAsync(counts, bc.p, sizeof(int64_t) * nr, hipMemcpyDeviceToHost, s));
        EHSS_HIP(hipMemcpyAsync(escaped, be.p, nr, hipMemcpyDeviceToHost, s));
        EHSS_HIP(hipMemcpyAsync(unresolved, bu.p, nr, hipMemcpyDeviceToHost, s));
        if (hist) EHSS_HIP(hipMemcpyAsync(history, bh.p, hist, hipMemcpyDeviceToHost, s));
    }
    unsigned long long k[4] = {0, 0, 0, 0};
    EHSS_HIP(hipMemcpyAsync(k, bk.p, sizeof k, hipMemcpyDeviceToHost, s));
    EHSS_HIP(hipStreamSynchronize(s));
    if (counters_out) {
        counters_out[0] = (int64_t)k[0];
        counters_out[1] = (int64_t)k[1];
        if (precision == EHSS_PRECISION_FP32_FIXUP) counters_out[2] = (int64_t)k[2];
    }
    return EHSS_OK;
}

int compat_trace(int64_t nr, int64_t nn, int cap, int precision, int record, double* p, double* d,
                 const int64_t* lasts, int64_t* counts, uint8_t* escaped, uint8_t* unresolved, int32_t* history,
                 const double* lo, const double* hi, const int64_t* end, const int64_t* leaf,
                 const int64_t* offsets, const int64_t* ids, const double* world, const double* radii,
                 int64_t* counters) {
    EHSS_TRY(ensure_init());
    EHSS_CHECK(nr == 0 || (p && d && lasts && counts && escaped && unresolved), "null ray buffer");
    EHSS_CHECK(lo && hi && end && leaf && offsets && ids && world && radii, "null tree buffer");
    hipStream_t s = lib_stream();
    CompatScene cs;
    EHSS_TRY(upload_compat(cs, nn, lo, hi, end, leaf, offsets, ids, world, radii,
                           precision == EHSS_PRECISION_FP32_FIXUP, p, nr, s));
    int64_t k[3] = {0, 0, 0};
    EHSS_TRY(trace_host_rays(cs.view, nr, cap, precision, record, p, d, lasts, counts, escaped, unresolved,
                             history, k, s));
    if (counters) {
        counters[0] = k[0];
        counters[1] = k[1];
    }
    return EHSS_OK;
}

// Device IEEE arithmetic check: out_div[i] = a[i] / b[i], out_sqrt[i] = sqrt(a[i]).
__global__ void k_selftest(int64_t n, const double* a, const double* b, double* q, double* r) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        q[i] = a[i] / b[i];
        r[i] = sqrt(a[i]);
    }
}

}  // namespace ehss

using namespace ehss;

extern "C" {

int ehss_abi_version(void) { return EHSS_ABI_VERSION; }

int ehss_last_error(char* buf, size_t size) {
    if (!buf || size == 0) return EHSS_ERR_INVALID;
    snprintf(buf, size, "%s", g_error.c_str());
    return EHSS_OK;
}

int ehss_init(int device) {
    {
        std::lock_guard<std::mutex> lock(g_init_mutex);
        if (g_stream && device != g_device)
            return set_error(EHSS_ERR_INVALID, "library already bound to device %d", g_device);
        g_device = device;
    }
    return ensure_init();
}

int ehss_device_info(char* name, size_t name_size, int* cu_count, int64_t* total_mem, int* arch_major,
                     int* arch_minor) {
    EHSS_TRY(ensure_init());
    hipDeviceProp_t prop;
    EHSS_HIP(hipGetDeviceProperties(&prop, g_device));
    // The runtime may leave the marketing name empty; the architecture string
    // and the CU count name the part then.
    if (name && name_size) {
        if (prop.name[0])
            snprintf(name, name_size, "%s (%s)", prop.name, prop.gcnArchName);
        else
            snprintf(name, name_size, "%s, %d CUs", prop.gcnArchName, prop.multiProcessorCount);
    }
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (total_mem) *total_mem = (int64_t)prop.totalGlobalMem;
    if (arch_major) *arch_major = prop.major;
    if (arch_minor) *arch_minor = prop.minor;
    return EHSS_OK;
}

int ehss_trig_compat_info(int64_t* entries, char* libm, size_t libm_size, char* path, size_t path_size) {
    EHSS_TRY(ensure_init());
    if (entries) *entries = g_trig_entries;
    if (libm && libm_size) snprintf(libm, libm_size, "%s", g_trig_libm.c_str());
    if (path && path_size) snprintf(path, path_size, "%s", g_trig_path.c_str());
    return EHSS_OK;
}

int ehss_sobol_sincos(int64_t n, const uint32_t* q, double* s, double* c) {
    EHSS_TRY(ensure_init());
    EHSS_CHECK(n >= 0 && q && s && c, "invalid ehss_sobol_sincos arguments");
    for (int64_t i = 0; i < n; ++i) EHSS_CHECK(q[i] < (1u << 30), "q[%lld] is not a 30-bit Sobol word", (long long)i);
    if (n == 0) return EHSS_OK;
    hipStream_t st = lib_stream();
    DevBuf dq, ds, dc;
    EHSS_TRY(dq.alloc(sizeof(uint32_t) * (size_t)n));
    EHSS_TRY(ds.alloc(sizeof(double) * (size_t)n));
    EHSS_TRY(dc.alloc(sizeof(double) * (size_t)n));
    EHSS_HIP(hipMemcpyAsync(dq.p, q, sizeof(uint32_t) * (size_t)n, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_sobol_sincos, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, dq.as<uint32_t>(),
                       g_trig, ds.as<double>(), dc.as<double>());
    EHSS_HIP(hipGetLastError());
    EHSS_HIP(hipMemcpyAsync(s, ds.p, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, st));
    EHSS_HIP(hipMemcpyAsync(c, dc.p, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, st));
    EHSS_HIP(hipStreamSynchronize(st));
    return EHSS_OK;
}

void ehss_trace(int64_t nr, int64_t nn, int cap, int threads, int record, double* p, double* d,
                const int64_t* lasts, int64_t* counts, uint8_t* escaped, uint8_t* unresolved, int32_t* history,
                const double* lo, const double* hi, const int64_t* end, const int64_t* leaf,
                const int64_t* offsets, const int64_t* ids, const double* world, const double* radii,
                int64_t* counters, int* observed_threads) {
    int64_t k[2] = {0, 0};
    const int s = compat_trace(nr, nn, cap, EHSS_PRECISION_FP64, record, p, d, lasts, counts, escaped, unresolved,
                               history, lo, hi, end, leaf, offsets, ids, world, radii, k);
    if (counters) {
        counters[0] = s == EHSS_OK ? k[0] : -1;
        counters[1] = s == EHSS_OK ? k[1] : -1;
    }
    if (observed_threads) *observed_threads = s == EHSS_OK ? threads : -1;
}

int ehss_trace_status(int64_t nr, int64_t nn, int cap, int precision, int record, double* p, double* d,
                      const int64_t* lasts, int64_t* counts, uint8_t* escaped, uint8_t* unresolved,
                      int32_t* history, const double* lo, const double* hi, const int64_t* end,
                      const int64_t* leaf, const int64_t* offsets, const int64_t* ids, const double* world,
                      const double* radii, int64_t* counters) {
    return compat_trace(nr, nn, cap, precision, record, p, d, lasts, counts, escaped, unresolved, history, lo, hi,
                        end, leaf, offsets, ids, world, radii, counters);
}

int ehss_selftest_fp64(int64_t n, const double* a, const double* b, double* quotient, double* root) {
    EHSS_TRY(ensure_init());
    EHSS_CHECK(n >= 0 && a && b && quotient && root, "invalid self-test arguments");
    if (n == 0) return EHSS_OK;
    hipStream_t s = lib_stream();
    DevBuf da, db, dq, dr;
    const size_t bytes = sizeof(double) * (size_t)n;
    EHSS_TRY(da.alloc(bytes));
    EHSS_TRY(db.alloc(bytes));
    EHSS_TRY(dq.alloc(bytes));
    EHSS_TRY(dr.alloc(bytes));
    EHSS_HIP(hipMemcpyAsync(da.p, a, bytes, hipMemcpyHostToDevice, s));
    EHSS_HIP(hipMemcpyAsync(db.p, b, bytes, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_selftest, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, da.as<double>(),
                       db.as<double>(), dq.as<double>(), dr.as<double>());
    EHSS_HIP(hipGetLastError());
    EHSS_HIP(hipMemcpyAsync(quotient, dq.p, bytes, hipMemcpyDeviceToHost, s));
    EHSS_HIP(hipMemcpyAsync(root, dr.p, bytes, hipMemcpyDeviceToHost, s));
    EHSS_HIP(hipStreamSynchronize(s));
    return EHSS_OK;
}

}  // extern "C"
