// pagrid.hip -- finite-grid PA, the CLI's adaptive method (ccs_cli.py:142-160),
// on the device.
//
// For every direction the atoms are projected onto the plane orthogonal to it
// (pa_walsh_geometry.py:171-180); the PA of that direction is the number of
// n x n midpoint-grid pixels covered by the union of the projected disks times
// the pixel area (pa_sparse_blocks.py:166-174 counts exactly this set).  A
// pixel of row y is covered by disk a when
//     (lower0 + (x + .5) spacing - cx)^2 <= r^2 - (lower1 + (y + .5) spacing - cy)^2,
// evaluated per row as an interval [lo, hi) with the rounding and refinement of
// row_span (pa_dyadic_bitmask.py:55-66) over the candidate rows of row_bounds
// (pa_dyadic_bitmask.py:68-72), so the count is the same integer.
//
// Device formulation: one workgroup per (direction, band of kBandRows rows).
// The band's pixels are a bitmask in LDS; every disk ORs its row intervals into
// it (word atomics, order-free), then the set bits are counted.
#include <math.h>
#include <string.h>

#include <vector>

#include "ehss_internal.h"

namespace ehss {

namespace {

constexpr int kBandRows = 16;
constexpr int kMaxGrid = 8192;
constexpr int kWordsPerRow = kMaxGrid / 32;

__device__ __forceinline__ int64_t clamp_n(int64_t v, int64_t n) { return v < 0 ? 0 : (v > n ? n : v); }

__device__ __forceinline__ double sq(double v) { return v * v; }

// pa_dyadic_bitmask.py:55-66.
__device__ __forceinline__ void row_span_dev(double cx, double cy, double r, double lower0, double lower1,
                                             double spacing, int64_t y, int64_t n, int64_t* lo_out,
                                             int64_t* hi_out) {
    const double dy = lower1 + ((double)y + .5) * spacing - cy;
    const double rem = r * r - dy * dy;
    *lo_out = *hi_out = 0;
    if (rem < 0) return;
    const double half = sqrt(rem);
    int64_t lo = clamp_n((int64_t)ceil((cx - half - lower0) / spacing - .5), n);
    int64_t hi = clamp_n((int64_t)floor((cx + half - lower0) / spacing - .5) + 1, n);
    while (lo > 0 && sq(lower0 + ((double)lo - .5) * spacing - cx) <= rem) --lo;
    while (lo < hi && sq(lower0 + ((double)lo + .5) * spacing - cx) > rem) ++lo;
    while (hi < n && sq(lower0 + ((double)hi + .5) * spacing - cx) <= rem) ++hi;
    while (hi > lo && sq(lower0 + ((double)hi - .5) * spacing - cx) > rem) --hi;
    *lo_out = lo;
    *hi_out = hi;
}

// pa_dyadic_bitmask.py:68-72.
__device__ __forceinline__ void row_bounds_dev(double cy, double r, double lower1, double spacing, int64_t n,
                                               int64_t* lo_out, int64_t* hi_out) {
    *lo_out = clamp_n((int64_t)ceil((cy - r - lower1) / spacing - .5) - 1, n);
    *hi_out = clamp_n((int64_t)floor((cy + r - lower1) / spacing - .5) + 2, n);
}

// Projection of atom a for direction d: x = c.u, y = c.v (centred coordinates),
// and the enclosing half-width max(|x| + r, |y| + r) as an order-free integer
// maximum of the (positive) double's bits.
__global__ void k_project(int64_t na, const double* centered, const double* radii, int n_dirs, const double* frames,
                          double* xy, unsigned long long* extent_bits) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= na * n_dirs) return;
    const int64_t d = i / na, a = i % na;
    const double* f = frames + 6 * d;
    const double c0 = centered[3 * a], c1 = centered[3 * a + 1], c2 = centered[3 * a + 2];
    const double x = (c0 * f[0] + c1 * f[1]) + c2 * f[2];
    const double y = (c0 * f[3] + c1 * f[4]) + c2 * f[5];
    xy[2 * i] = x;
    xy[2 * i + 1] = y;
    const double e = fmax(fabs(x) + radii[a], fabs(y) + radii[a]);
    atomicMax(&extent_bits[d], (unsigned long long)__double_as_longlong(e));
}

// Grid frame of direction d (pa_walsh_geometry.py:178-180).
__global__ void k_frame(int n_dirs, int64_t n, const unsigned long long* extent_bits, double* lower,
                        double* spacing) {
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= n_dirs) return;
    const double extent = __longlong_as_double((long long)extent_bits[d]) * (1. + 1e-10);
    lower[2 * d] = -extent;
    lower[2 * d + 1] = -extent;
    spacing[d] = 2 * extent / (double)n;
}

// Covered pixels of band blockIdx.x of direction blockIdx.y.
__global__ void __launch_bounds__(256) k_count(int64_t na, const double* xy, const double* radii, int64_t n,
                                               const double* lower, const double* spacing,
                                               unsigned long long* counts) {
    __shared__ uint32_t mask[kBandRows * kWordsPerRow];
    __shared__ unsigned long long part[256 / kWave];
    const int d = blockIdx.y;
    const int64_t y0 = (int64_t)blockIdx.x * kBandRows;
    const int64_t y1 = y0 + kBandRows < n ? y0 + kBandRows : n;
    const int words = (int)((n + 31) / 32);
    for (int k = threadIdx.x; k < kBandRows * words; k += blockDim.x) mask[k] = 0u;
    __syncthreads();
    const double l0 = lower[2 * d], l1 = lower[2 * d + 1], sp = spacing[d];
    const double* P = xy + 2 * na * d;
    for (int64_t a = threadIdx.x; a < na; a += blockDim.x) {
        const double cx = P[2 * a], cy = P[2 * a + 1], r = radii[a];
        int64_t first, stop;
        row_bounds_dev(cy, r, l1, sp, n, &first, &stop);
        if (first < y0) first = y0;
        if (stop > y1) stop = y1;
        for (int64_t y = first; y < stop; ++y) {
            int64_t lo, hi;
            row_span_dev(cx, cy, r, l0, l1, sp, y, n, &lo, &hi);
            if (hi <= lo) continue;
            uint32_t* row = mask + (y - y0) * words;
            for (int64_t w = lo >> 5; w <= (hi - 1) >> 5; ++w) {
                const int64_t a0 = lo > (w << 5) ? lo - (w << 5) : 0;
                const int64_t a1 = hi < ((w + 1) << 5) ? hi - (w << 5) : 32;
                const uint32_t bits = (a1 >= 32 ? 0xFFFFFFFFu : ((1u << a1) - 1u)) & ~((1u << a0) - 1u);
                atomicOr(&row[w], bits);
            }
        }
    }
    __syncthreads();
    unsigned long long c = 0;
    for (int k = threadIdx.x; k < (int)(y1 - y0) * words; k += blockDim.x) c += (unsigned long long)__popc(mask[k]);
    for (int off = kWave / 2; off > 0; off >>= 1) c += __shfl_xor(c, off, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) part[threadIdx.x / kWave] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int k = 0; k < (int)(blockDim.x / kWave); ++k) t += part[k];
        if (t) atomicAdd(&counts[d], t);
    }
}

__global__ void k_zero_u64(unsigned long long* p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

int check_grid(int grid) {
    EHSS_CHECK(grid >= 8 && grid <= kMaxGrid && (grid & (grid - 1)) == 0,
               "grid size %d must be a power of two in 8..%d", grid, kMaxGrid);
    return EHSS_OK;
}

}  // namespace

}  // namespace ehss

using namespace ehss;

extern "C" {

int ehss_pa_grid(int64_t n_atoms, const double* centered, const double* radii, int n_dirs, const double* frames,
                 int grid, int64_t* counts, double* spacing_out) {
    EHSS_TRY(ensure_init());
    EHSS_CHECK(n_atoms >= 1 && centered && radii && frames && counts, "invalid grid PA arguments");
    EHSS_CHECK(n_dirs >= 1 && n_dirs <= (1 << 16), "direction count %d out of range 1..65536", n_dirs);
    EHSS_TRY(check_grid(grid));
    for (int64_t a = 0; a < n_atoms; ++a)
        EHSS_CHECK(isfinite(centered[3 * a]) && isfinite(centered[3 * a + 1]) && isfinite(centered[3 * a + 2]) &&
                       isfinite(radii[a]) && radii[a] >= 0,
                   "Expected finite geometry, nonnegative radii and positive spacing");
    hipStream_t s = lib_stream();
    DevBuf dc, dr, df, dxy, dext, dlow, dsp, dcnt;
    EHSS_TRY(dc.alloc(sizeof(double) * 3 * n_atoms));
    EHSS_TRY(dr.alloc(sizeof(double) * n_atoms));
    EHSS_TRY(df.alloc(sizeof(double) * 6 * n_dirs));
    EHSS_TRY(dxy.alloc(sizeof(double) * 2 * n_atoms * n_dirs));
    EHSS_TRY(dext.alloc(sizeof(unsigned long long) * n_dirs));
    EHSS_TRY(dlow.alloc(sizeof(double) * 2 * n_dirs));
    EHSS_TRY(dsp.alloc(sizeof(double) * n_dirs));
    EHSS_TRY(dcnt.alloc(sizeof(unsigned long long) * n_dirs));
    EHSS_HIP(hipMemcpyAsync(dc.p, centered, sizeof(double) * 3 * n_atoms, hipMemcpyHostToDevice, s));
    EHSS_HIP(hipMemcpyAsync(dr.p, radii, sizeof(double) * n_atoms, hipMemcpyHostToDevice, s));
    EHSS_HIP(hipMemcpyAsync(df.p, frames, sizeof(double) * 6 * n_dirs, hipMemcpyHostToDevice, s));
    const unsigned zb = (unsigned)((n_dirs + 255) / 256);
    hipLaunchKernelGGL(k_zero_u64, dim3(zb), dim3(256), 0, s, dext.as<unsigned long long>(), (int64_t)n_dirs);
    hipLaunchKernelGGL(k_zero_u64, dim3(zb), dim3(256), 0, s, dcnt.as<unsigned long long>(), (int64_t)n_dirs);
    const int64_t total = n_atoms * n_dirs;
    hipLaunchKernelGGL(k_project, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, n_atoms, dc.as<double>(),
                       dr.as<double>(), n_dirs, df.as<double>(), dxy.as<double>(), dext.as<unsigned long long>());
    hipLaunchKernelGGL(k_frame, dim3(zb), dim3(256), 0, s, n_dirs, (int64_t)grid, dext.as<unsigned long long>(),
                       dlow.as<double>(), dsp.as<double>());
    hipLaunchKernelGGL(k_count, dim3((unsigned)((grid + kBandRows - 1) / kBandRows), (unsigned)n_dirs), dim3(256), 0,
                       s, n_atoms, dxy.as<double>(), dr.as<double>(), (int64_t)grid, dlow.as<double>(),
                       dsp.as<double>(), dcnt.as<unsigned long long>());
    EHSS_HIP(hipGetLastError());
    std::vector<unsigned long long> hc((size_t)n_dirs);
    EHSS_HIP(hipMemcpyAsync(hc.data(), dcnt.p, sizeof(unsigned long long) * n_dirs, hipMemcpyDeviceToHost, s));
    if (spacing_out)
        EHSS_HIP(hipMemcpyAsync(spacing_out, dsp.p, sizeof(double) * n_dirs, hipMemcpyDeviceToHost, s));
    EHSS_HIP(hipStreamSynchronize(s));
    for (int d = 0; d < n_dirs; ++d) counts[d] = (int64_t)hc[d];
    return EHSS_OK;
}

int ehss_pa_grid_count(int64_t n_atoms, const double* xy, const double* radii, const double* lower, double spacing,
                       int grid, int64_t* count) {
    EHSS_TRY(ensure_init());
    EHSS_CHECK(n_atoms >= 1 && xy && radii && lower && count, "invalid grid count arguments");
    EHSS_TRY(check_grid(grid));
    EHSS_CHECK(isfinite(lower[0]) && isfinite(lower[1]) && isfinite(spacing) && spacing > 0,
               "Expected finite geometry, nonnegative radii and positive spacing");
    for (int64_t a = 0; a < n_atoms; ++a)
        EHSS_CHECK(isfinite(xy[2 * a]) && isfinite(xy[2 * a + 1]) && isfinite(radii[a]) && radii[a] >= 0,
                   "Expected finite geometry, nonnegative radii and positive spacing");
    hipStream_t s = lib_stream();
    DevBuf dxy, dr, dlow, dsp, dcnt;
    EHSS_TRY(dxy.alloc(sizeof(double) * 2 * n_atoms));
    EHSS_TRY(dr.alloc(sizeof(double) * n_atoms));
    EHSS_TRY(dlow.alloc(sizeof(double) * 2));
    EHSS_TRY(dsp.alloc(sizeof(double)));
    EHSS_TRY(dcnt.alloc(sizeof(unsigned long long)));
    EHSS_HIP(hipMemcpyAsync(dxy.p, xy, sizeof(double) * 2 * n_atoms, hipMemcpyHostToDevice, s));
    EHSS_HIP(hipMemcpyAsync(dr.p, radii, sizeof(double) * n_atoms, hipMemcpyHostToDevice, s));
    EHSS_HIP(hipMemcpyAsync(dlow.p, lower, sizeof(double) * 2, hipMemcpyHostToDevice, s));
    EHSS_HIP(hipMemcpyAsync(dsp.p, &spacing, sizeof(double), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(256), 0, s, dcnt.as<unsigned long long>(), (int64_t)1);
    hipLaunchKernelGGL(k_count, dim3((unsigned)((grid + kBandRows - 1) / kBandRows), 1), dim3(256), 0, s, n_atoms,
                       dxy.as<double>(), dr.as<double>(), (int64_t)grid, dlow.as<double>(), dsp.as<double>(),
                       dcnt.as<unsigned long long>());
    EHSS_HIP(hipGetLastError());
    unsigned long long c = 0;
    EHSS_HIP(hipMemcpyAsync(&c, dcnt.p, sizeof c, hipMemcpyDeviceToHost, s));
    EHSS_HIP(hipStreamSynchronize(s));
    *count = (int64_t)c;
    return EHSS_OK;
}

}  // extern "C"
