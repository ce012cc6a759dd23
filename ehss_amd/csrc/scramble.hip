// scramble.hip -- host construction of the scrambled 4-D Sobol net of
// scipy.stats.qmc.Sobol(4, scramble=True, seed) (_qmc.py:1808-1823): the
// digital shift and the lower-triangular (LMS) scramble matrices are drawn from
// numpy.random.default_rng(seed), then applied to the direction numbers.
//
// The generator chain is NumPy's, restated from its published algorithms:
//   SeedSequence(seed)      entropy pooling with the hashmix/mix constants of
//                           numpy/random/bit_generator.pyx (pool of 4 words);
//   generate_state(4, u64)  8 mixed 32-bit words, paired little-endian;
//   PCG64                   128-bit LCG (O'Neill's PCG "XSL RR 128/64"), seeded
//                           as pcg_setseq_128_srandom_r(state, seq);
//   next_uint32             the low half of a 64-bit output, then the high half;
//   integers(0, 2, uint32)  Lemire's bounded draw: (x * 2) >> 32, no rejection
//                           for a range of 2.
// The Python restatement in ehss_amd.engine (checked against SciPy) is the
// reference these bits are tested against.
#include <string.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <thread>
#include <vector>

#include "ehss_internal.h"

namespace ehss {
namespace {

typedef unsigned __int128 u128;

// SeedSequence constants (bit_generator.pyx).
constexpr uint32_t kInitA = 0x43b0d7e5u, kMultA = 0x931e8875u, kInitB = 0x8b51f9ddu, kMultB = 0x58f38dedu;
constexpr uint32_t kMixMultL = 0xca01f9ddu, kMixMultR = 0x4973f715u;
constexpr int kXShift = 16;
constexpr int kPool = 4;

struct SeedSeq {
    uint32_t pool[kPool];

    explicit SeedSeq(uint64_t seed) {
        // An integer entropy becomes its 32-bit words, least significant first
        // (a single 0 word for 0).
        uint32_t ent[2];
        int ne = 0;
        do {
            ent[ne++] = (uint32_t)seed;
            seed >>= 32;
        } while (seed && ne < 2);
        uint32_t hash = kInitA;
        auto hashmix = [&](uint32_t v) {
            v ^= hash;
            hash *= kMultA;
            v *= hash;
            v ^= v >> kXShift;
            return v;
        };
        auto mix = [](uint32_t x, uint32_t y) {
            uint32_t r = kMixMultL * x - kMixMultR * y;
            r ^= r >> kXShift;
            return r;
        };
        for (int i = 0; i < kPool; ++i) pool[i] = hashmix(i < ne ? ent[i] : 0u);
        for (int s = 0; s < kPool; ++s)
            for (int d = 0; d < kPool; ++d)
                if (s != d) pool[d] = mix(pool[d], hashmix(pool[s]));
        for (int s = kPool; s < ne; ++s)
            for (int d = 0; d < kPool; ++d) pool[d] = mix(pool[d], hashmix(ent[s]));
    }

    // generate_state(n64, uint64): 2 n64 words cycled from the pool.
    void state64(uint64_t* out, int n64) const {
        uint32_t hash = kInitB;
        for (int i = 0; i < 2 * n64; ++i) {
            uint32_t v = pool[i % kPool];
            v ^= hash;
            hash *= kMultB;
            v *= hash;
            v ^= v >> kXShift;
            if (i & 1)
                out[i / 2] |= (uint64_t)v << 32;
            else
                out[i / 2] = v;
        }
    }
};

struct Pcg64 {
    u128 state = 0, inc = 0;
    bool has32 = false;
    uint32_t buf32 = 0;
    static constexpr u128 kMult = ((u128)2549297995355413924ULL << 64) | (u128)4865540595714422341ULL;

    Pcg64(u128 initstate, u128 initseq) {
        inc = (initseq << 1) | 1u;
        step();
        state += initstate;
        step();
    }
    void step() { state = state * kMult + inc; }
    uint64_t next64() {
        step();
        const uint64_t x = (uint64_t)(state >> 64) ^ (uint64_t)state;
        const unsigned rot = (unsigned)(state >> 122);
        return (x >> rot) | (x << ((64u - rot) & 63u));
    }
    uint32_t next32() {
        if (has32) {
            has32 = false;
            return buf32;
        }
        const uint64_t v = next64();
        has32 = true;
        buf32 = (uint32_t)(v >> 32);
        return (uint32_t)v;
    }
    // integers(0, 2) as uint32 (Lemire): the top bit of a 32-bit draw.
    uint32_t bit() { return (uint32_t)(((uint64_t)next32() * 2u) >> 32); }
};

constexpr int kBits = 30;
constexpr int kDim = 4;

// Direction numbers v[d][j] = m_j << (bits - 1 - j) of the first four Sobol
// dimensions (Joe-Kuo primitive polynomials; dimension 0 is the identity).
void direction_numbers(uint32_t v[kDim][kBits]) {
    static const int s_[kDim - 1] = {1, 2, 3}, a_[kDim - 1] = {0, 1, 1};
    static const int init[kDim - 1][3] = {{1, 0, 0}, {1, 3, 0}, {1, 3, 1}};
    for (int j = 0; j < kBits; ++j) v[0][j] = 1u << (kBits - 1 - j);
    for (int d = 1; d < kDim; ++d) {
        const int s = s_[d - 1], a = a_[d - 1];
        uint64_t m[kBits];
        for (int j = 0; j < s; ++j) m[j] = (uint64_t)init[d - 1][j];
        for (int j = s; j < kBits; ++j) {
            uint64_t x = m[j - s] ^ (m[j - s] << s);
            for (int k = 1; k < s; ++k)
                if ((a >> (s - 1 - k)) & 1) x ^= m[j - k] << k;
            m[j] = x;
        }
        for (int j = 0; j < kBits; ++j) v[d][j] = (uint32_t)(m[j] << (kBits - 1 - j));
    }
}

}  // namespace
}  // namespace ehss

using namespace ehss;

namespace {

void scramble_one(uint64_t seed, ehss_sobol* out) {
    static uint32_t v[kDim][kBits];
    static std::once_flag once;
    std::call_once(once, [] { direction_numbers(v); });
    const SeedSeq ss(seed);
    uint64_t st[4] = {0, 0, 0, 0};
    ss.state64(st, 4);
    Pcg64 rng(((u128)st[0] << 64) | st[1], ((u128)st[2] << 64) | st[3]);
    memset(out, 0, sizeof *out);
    // Shift bits first (d x bits, bit j weighs 2^j), then the LMS matrices
    // (d x bits x bits, C order), of which only the strict lower triangle is
    // used: the diagonal is one.
    for (int d = 0; d < kDim; ++d) {
        uint32_t sh = 0;
        for (int j = 0; j < kBits; ++j) sh |= rng.bit() << j;
        out->shift[d] = sh;
    }
    for (int d = 0; d < kDim; ++d) {
        uint32_t lrow[kBits];  // lrow[p]: bit q set when L[p][q] = 1 (q <= p)
        for (int p = 0; p < kBits; ++p) {
            uint32_t row = 0;
            for (int q = 0; q < kBits; ++q) {
                const uint32_t b = rng.bit();
                if (q < p) row |= b << q;
            }
            lrow[p] = row | (1u << p);
        }
        // sv[d][j]: MSB-first output bit p = parity(L[p, :] . (MSB-first bits of v[d][j])).
        for (int j = 0; j < kBits; ++j) {
            uint32_t vb = 0;  // bit k = MSB-first bit k of v[d][j]
            for (int k = 0; k < kBits; ++k) vb |= ((v[d][j] >> (kBits - 1 - k)) & 1u) << k;
            uint32_t w = 0;
            for (int p = 0; p < kBits; ++p) w |= (uint32_t)(__builtin_popcount(lrow[p] & vb) & 1) << (kBits - 1 - p);
            out->sv[d][j] = w;
        }
    }
}

}  // namespace

extern "C" int ehss_scramble(uint64_t seed, ehss_sobol* out) {
    EHSS_CHECK(out, "null output");
    scramble_one(seed, out);
    return EHSS_OK;
}

// Many scrambles on up to 16 host threads (chunks of 16 seeds).
extern "C" int ehss_scramble_many(int64_t n, const uint64_t* seeds, ehss_sobol* out) {
    EHSS_CHECK(n >= 0 && (n == 0 || (seeds && out)), "invalid ehss_scramble_many arguments");
    const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
    const int threads = (int)std::min<int64_t>({(int64_t)hw, 16, n / 32});
    std::atomic<int64_t> next{0};
    auto work = [&] {
        for (int64_t i; (i = next.fetch_add(16)) < n;)
            for (int64_t k = i; k < std::min(n, i + 16); ++k) scramble_one(seeds[k], &out[k]);
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
    return EHSS_OK;
}
