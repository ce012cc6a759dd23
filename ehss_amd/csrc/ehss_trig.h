// ehss_trig.h -- correctly rounded sin and cos of a double, plus the host-libm
// compatibility table of the Sobol angles, for host and device builds.
//
// Why: the reference computes its incidence source with NumPy on the host
// (ehss_state.py:45-49, inside_out_pa.py:45-49), whose cos/sin are the host
// libm's (glibc 2.35 here).  The long reflection paths are chaotic: a one-ulp
// change of a source ray can change where it escapes.  To give identical rays
// on identical Sobol points, the device must reproduce the host's bits.
//
// How:
//  1. ehss_sincos_dd evaluates sin and cos of a double x (|x| < 16) to about
//     2^-95 relative in double-double: Cody-Waite reduction by pi/2 in four
//     48-bit pieces (exact products), a table of sin/cos(j/512) as
//     double-doubles, and Taylor polynomials in the remainder s (|s| <= 2^-10),
//     the leading terms in double-double.  Rounding the result gives the
//     correctly rounded value unless the rounding test fails (never observed on
//     the 2^30 Sobol angles; the compatibility table would cover it anyway).
//  2. The Sobol angles are x = fl(2pi * q 2^-30) for the 30-bit Sobol word q
//     (ehss_state.py:48, inside_out_pa.py:47).  The host libm is not correctly
//     rounded on about 0.15% of them.  A table, generated at build time by
//     evaluating the host libm on all 2^30 angles (trig_host.cpp), lists every q
//     whose libm result differs from step 1, with the direction of the one-ulp
//     step.  ehss_sincos_sobol applies it.  The libm can only round the other
//     way when the exact value lies very close to a rounding midpoint, so the
//     table also records the threshold `lookup_from`: every listed angle has
//     |lo| >= lookup_from * ulp(hi) for its double-double result, and only
//     angles meeting that test (about 3 % of them) are looked up.
//
// The arithmetic uses only IEEE +, -, *, fma, rint and exact scalings, so the
// host build (g++ -ffp-contract=off) and the device build give the same bits.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define EHSS_TRIG_HD __host__ __device__ __forceinline__
#define EHSS_TRIG_CONST static __constant__ const
#else
#define EHSS_TRIG_HD static inline
#define EHSS_TRIG_CONST static const
#endif

#include "ehss_trig_table.h"

namespace ehss_trig {

struct DD {
    double hi, lo;
};

EHSS_TRIG_HD DD two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    return DD{s, (a - (s - bb)) + (b - bb)};
}

// |a| >= |b| (or a == 0).
EHSS_TRIG_HD DD fast_two_sum(double a, double b) {
    const double s = a + b;
    return DD{s, b - (s - a)};
}

EHSS_TRIG_HD DD two_prod(double a, double b) {
    const double p = a * b;
    return DD{p, fma(a, b, -p)};
}

// (a.hi + a.lo) * (b.hi + b.lo), dropping lo * lo.
EHSS_TRIG_HD DD dd_mul(DD a, DD b) {
    DD p = two_prod(a.hi, b.hi);
    const double t = a.hi * b.lo + a.lo * b.hi;
    return fast_two_sum(p.hi, p.lo + t);
}

EHSS_TRIG_HD DD dd_add(DD a, DD b) {
    DD s = two_sum(a.hi, b.hi);
    return fast_two_sum(s.hi, s.lo + (a.lo + b.lo));
}

// Rounding test: the exact value lies within 2^-94 |h| of h + l (the
// evaluation error bound with margin); RN of it is h unless that interval
// reaches a rounding boundary of h.
EHSS_TRIG_HD bool rounding_sure(DD v) {
    const double e = fabs(v.hi) * 0x1p-93;
    return (v.hi + (v.lo + e) == v.hi) && (v.hi + (v.lo - e) == v.hi);
}

struct SinCos {
    DD s, c;
};

// sin and cos of x, |x| < 16, as double-doubles (relative error < 2^-95).
EHSS_TRIG_HD SinCos sincos_dd(double x) {
    // Quadrant n = rint(x 2/pi); r = x - n pi/2 with exact piece products.
    const double nf = rint(x * kTwoOverPi);
    const int n = (int)nf;
    const double r1 = x - nf * kPio2[0];  // exact (Sterbenz for n != 0)
    DD a = two_sum(r1, -nf * kPio2[1]);
    DD b = two_sum(a.hi, -nf * kPio2[2]);
    const double rl = (a.lo + b.lo) - nf * kPio2[3];
    DD r = fast_two_sum(b.hi, rl);
    // Table point a_j = j/512 nearest r; s = r - a_j, |s| <= 2^-10.
    const double jf = rint(r.hi * 512.0);
    const int j = (int)jf;
    const double sh = r.hi - jf * 0x1p-9;  // exact
    const double sl = r.lo;
    // z = s^2 as a double-double.
    DD z = two_prod(sh, sh);
    z.lo += 2.0 * sh * sl;
    // sin(s) = s - s z/6 + s z^2 (1/120 - z/5040 + z^2/362880)
    DD sz = dd_mul(DD{sh, sl}, z);
    DD t3 = dd_mul(sz, DD{kInv6Hi, kInv6Lo});
    const double zh = z.hi;
    const double t5 = (sh * zh) * (zh * (kInvFact5 + zh * (-kInvFact7 + zh * kInvFact9)));
    DD ss = two_sum(sh, -t3.hi);
    ss = fast_two_sum(ss.hi, ss.lo + ((sl - t3.lo) + t5));
    // cos(s) = 1 - z/2 + z^2 (1/24 - z/720 + z^2/40320)
    const double t4 = (zh * zh) * (kInvFact4 + zh * (-kInvFact6 + zh * kInvFact8));
    DD cs = fast_two_sum(1.0, -0.5 * zh);
    cs = fast_two_sum(cs.hi, cs.lo + (-0.5 * z.lo + t4));
    // sin/cos(a_j + s) from the table (sin(-a) = -sin a).
    const int aj = j < 0 ? -j : j;
    const double sgn = j < 0 ? -1.0 : 1.0;
    const DD sa{sgn * kTab[aj][0], sgn * kTab[aj][1]};
    const DD ca{kTab[aj][2], kTab[aj][3]};
    DD sinr = dd_add(dd_mul(sa, cs), dd_mul(ca, ss));
    DD cosr = dd_add(dd_mul(ca, cs), dd_mul(DD{-sa.hi, -sa.lo}, ss));
    // Quadrant rotation (exact).
    SinCos out;
    switch (n & 3) {
        case 0: out.s = sinr; out.c = cosr; break;
        case 1: out.s = cosr; out.c = DD{-sinr.hi, -sinr.lo}; break;
        case 2: out.s = DD{-sinr.hi, -sinr.lo}; out.c = DD{-cosr.hi, -cosr.lo}; break;
        default: out.s = DD{-cosr.hi, -cosr.lo}; out.c = sinr; break;
    }
    return out;
}

// Correctly rounded cos and sin of x (|x| < 16); *sure = 0 if either rounding
// test failed.
EHSS_TRIG_HD void sincos_cr(double x, double* s, double* c, int* sure) {
    const SinCos v = sincos_dd(x);
    *s = v.s.hi;
    *c = v.c.hi;
    *sure = rounding_sure(v.s) && rounding_sure(v.c);
}

// Compatibility table with the host libm on the Sobol angles fl(2pi q 2^-30):
// buckets of q >> kCompatShift; entry = (q & low mask) << 4 | flags, sorted.
// Flags: 1 cos differs, 2 cos is one ulp up (else down), 4 sin differs, 8 sin up.
constexpr int kCompatShift = 11;
constexpr uint32_t kCompatBuckets = 1u << (30 - kCompatShift);
struct CompatTable {
    const uint32_t* offsets;  // kCompatBuckets + 1
    const uint16_t* entries;
    double lookup_from;       // look up only when |lo| >= lookup_from * ulp(hi) for cos or sin
};

// Spacing of the doubles at |h| (frexp's 2^(e - 53)); exact.
EHSS_TRIG_HD double ulp_of(double h) {
    int e;
    frexp(h, &e);
    return ldexp(1.0, e - 53);
}

// |lo| / ulp(hi) >= from, evaluated exactly (ulp_of is a power of two).
EHSS_TRIG_HD bool near_midpoint(DD v, double from) { return fabs(v.lo) >= from * ulp_of(v.hi); }

EHSS_TRIG_HD double ulp_step(double v, bool up) {
    return nextafter(v, up ? INFINITY : -INFINITY);
}

EHSS_TRIG_HD uint32_t compat_flags(CompatTable t, uint32_t q) {
    if (!t.offsets) return 0;
    const uint32_t b = q >> kCompatShift;
    const uint32_t key = q & ((1u << kCompatShift) - 1u);
    uint32_t lo = t.offsets[b];
    const uint32_t hi = t.offsets[b + 1];
    for (; lo < hi; ++lo) {
        const uint32_t e = t.entries[lo];
        const uint32_t k = e >> 4;
        if (k >= key) return k == key ? (e & 15u) : 0u;
    }
    return 0;
}

// cos and sin of the Sobol angle x = fl(2pi * q 2^-30) as the host libm
// returns them (correctly rounded when the table is absent).
EHSS_TRIG_HD void sincos_sobol(double x, uint32_t q, CompatTable t, double* s, double* c) {
    const SinCos v = sincos_dd(x);
    *s = v.s.hi;
    *c = v.c.hi;
    if (!t.offsets || !(near_midpoint(v.c, t.lookup_from) || near_midpoint(v.s, t.lookup_from))) return;
    const uint32_t f = compat_flags(t, q);
    if (f & 1u) *c = ulp_step(*c, (f & 2u) != 0);
    if (f & 4u) *s = ulp_step(*s, (f & 8u) != 0);
}

}  // namespace ehss_trig
