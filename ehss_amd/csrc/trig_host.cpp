// trig_host.cpp -- host build of ehss_trig.h (libehss_trighost.so).
//
// Two uses:
//  * ehss_trig_compat_generate writes the libm compatibility table that the
//    device source generation applies: for every 30-bit Sobol word q it compares
//    the host libm's cos/sin of x = fl(2pi * q 2^-30) (the angles of
//    ehss_state.py:48 and inside_out_pa.py:47) with the correctly rounded value
//    of ehss_trig.h and records each q where they differ, with the direction of
//    the one-ulp step.  Called by the package build.
//  * ehss_trig_host_* evaluate the same functions on the host, so tests can
//    compare the device's bits with the host build of the same source.
//
// Built with g++ -O2 -ffp-contract=off (no FMA contraction); fma() calls are
// exact, so every result equals the device build's.
#include <gnu/libc-version.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "ehss_trig.h"

using namespace ehss_trig;

namespace {

struct FileHeader {
    char magic[8];        // "EHSSTRG2"
    uint32_t key_bits;    // 30
    uint32_t shift;       // kCompatShift
    uint32_t buckets;     // kCompatBuckets
    uint32_t entries;
    uint32_t not_sure;    // angles whose rounding test failed (listed in the table when they differ)
    uint32_t libm_ulp2;   // angles where libm is two or more ulps away (must be 0)
    double lookup_from;   // CompatTable::lookup_from
    char libm[24];        // "glibc " + gnu_get_libc_version()
};
static_assert(sizeof(FileHeader) == 64, "64-byte header");

constexpr double kTwoPi = 6.283185307179586;  // 2 * np.pi

// NumPy calls cos and sin separately; called directly, the compiler may merge
// the pair into one sincos(), whose last bits differ in rare cases.
double (*volatile libm_cos)(double) = cos;
double (*volatile libm_sin)(double) = sin;
constexpr double kScale = 9.313225746154785e-10;  // 2^-30

}  // namespace

extern "C" {

// Host build of sincos_cr: s, c correctly rounded; sure[i] = rounding test passed.
void ehss_trig_host_sincos(int64_t n, const double* x, double* s, double* c, int8_t* sure) {
    for (int64_t i = 0; i < n; ++i) {
        int ok;
        sincos_cr(x[i], &s[i], &c[i], &ok);
        if (sure) sure[i] = (int8_t)ok;
    }
}

// Host build of sincos_sobol on words q (table arrays may be null).
void ehss_trig_host_sobol(int64_t n, const uint32_t* q, const uint32_t* offsets, const uint16_t* entries,
                          double lookup_from, double* s, double* c) {
    const CompatTable t{offsets, entries, lookup_from};
    for (int64_t i = 0; i < n; ++i) {
        const double x = kTwoPi * ((double)q[i] * kScale);
        sincos_sobol(x, q[i], t, &s[i], &c[i]);
    }
}

// Writes the compatibility table to `path`.  stats (8 int64): entries, cos
// differences, sin differences, not-sure angles, libm >= 2 ulps away, buckets,
// angles the device looks up (near_midpoint for cos or sin).
// Returns 0 on success, -1 on an I/O error, -2 if libm is ever two or more ulps
// from the correctly rounded value (the table format cannot express that).
int ehss_trig_compat_generate(const char* path, int threads, int64_t* stats) {
    const int kChunk = 256;  // buckets per work item
    const int chunks = (int)(kCompatBuckets / kChunk);
    std::vector<std::vector<uint16_t>> out(chunks);
    std::vector<uint32_t> counts(kCompatBuckets, 0);
    int64_t ncos = 0, nsin = 0, unsure = 0, far = 0;
    // Largest |lo| / ulp(hi) deficit below one half over the listed angles:
    // every listed angle has |lo| / ulp(hi) >= 0.5 - gap (exact arithmetic).
    double gap = 0.0;
    if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 4) reduction(+ : ncos, nsin, unsure, far) \
    reduction(max : gap)
    for (int ch = 0; ch < chunks; ++ch) {
        for (uint32_t b = (uint32_t)ch * kChunk; b < (uint32_t)(ch + 1) * kChunk; ++b) {
            for (uint32_t k = 0; k < (1u << kCompatShift); ++k) {
                const uint32_t q = (b << kCompatShift) | k;
                const double x = kTwoPi * ((double)q * kScale);
                const SinCos v = sincos_dd(x);
                const double s = v.s.hi, c = v.c.hi;
                unsure += !(rounding_sure(v.s) && rounding_sure(v.c));
                const double lc = libm_cos(x), ls = libm_sin(x);
                uint32_t f = 0;
                if (lc != c) {
                    if (lc == ulp_step(c, true)) f |= 3u;
                    else if (lc == ulp_step(c, false)) f |= 1u;
                    else ++far;
                    ++ncos;
                }
                if (ls != s) {
                    if (ls == ulp_step(s, true)) f |= 12u;
                    else if (ls == ulp_step(s, false)) f |= 4u;
                    else ++far;
                    ++nsin;
                }
                if (f & 1u) gap = fmax(gap, 0.5 - fabs(v.c.lo) / ulp_of(v.c.hi));
                if (f & 4u) gap = fmax(gap, 0.5 - fabs(v.s.lo) / ulp_of(v.s.hi));
                if (f) {
                    out[ch].push_back((uint16_t)((k << 4) | f));
                    ++counts[b];
                }
            }
        }
    }
    std::vector<uint32_t> offsets(kCompatBuckets + 1, 0);
    for (uint32_t b = 0; b < kCompatBuckets; ++b) offsets[b + 1] = offsets[b] + counts[b];
    const uint32_t total = offsets[kCompatBuckets];
    const double lookup_from = 0.5 - gap;
    // How many angles the device looks up with this threshold.
    int64_t looked = 0;
#pragma omp parallel for num_threads(threads) schedule(static) reduction(+ : looked)
    for (int64_t q = 0; q < (1LL << 30); ++q) {
        const SinCos v = sincos_dd(kTwoPi * ((double)q * kScale));
        looked += near_midpoint(v.c, lookup_from) || near_midpoint(v.s, lookup_from);
    }
    if (stats) {
        stats[0] = total;
        stats[1] = ncos;
        stats[2] = nsin;
        stats[3] = unsure;
        stats[4] = far;
        stats[5] = kCompatBuckets;
        stats[6] = looked;
    }
    if (far) return -2;
    FileHeader h;
    memset(&h, 0, sizeof h);
    memcpy(h.magic, "EHSSTRG2", 8);
    h.key_bits = 30;
    h.shift = kCompatShift;
    h.buckets = kCompatBuckets;
    h.entries = total;
    h.not_sure = (uint32_t)unsure;
    h.libm_ulp2 = 0;
    h.lookup_from = lookup_from;
    snprintf(h.libm, sizeof h.libm, "glibc %s", gnu_get_libc_version());
    FILE* f = fopen(path, "wb");
    if (!f) return -1;
    bool okw = fwrite(&h, sizeof h, 1, f) == 1;
    okw = okw && fwrite(offsets.data(), sizeof(uint32_t), offsets.size(), f) == offsets.size();
    for (const auto& v : out) okw = okw && (v.empty() || fwrite(v.data(), sizeof(uint16_t), v.size(), f) == v.size());
    okw = (fclose(f) == 0) && okw;
    return okw ? 0 : -1;
}

// The host libm's version string, as recorded in the table header.
void ehss_trig_libm_version(char* buf, size_t size) { snprintf(buf, size, "glibc %s", gnu_get_libc_version()); }
}
