// CPU model of the trial-division lattice kernel: the device's per-prime
// routine td_lattice_prime (prime-finder_amd/csrc/td_lattice.h) compiled for
// the host with g++ and run prime by prime, marks OR-ed into a host bitmap.
// Test infrastructure (tests/test_td_lattice_model.py builds and loads it).
#include "td_lattice.h"

#include <stdint.h>

namespace {
struct HostMark {
    uint32_t *bits;
    void operator()(uint32_t bit) const { bits[bit >> 5] |= 1u << (bit & 31u); }
};
}

extern "C" int lattice_sieve(const uint32_t *primes, uint32_t n, long long a, long long b,
                             long long c, uint32_t K, uint32_t J, uint32_t *bits)
{
    const int DJ = J >= 1000u ? 4 : J >= 100u ? 3 : J >= 10u ? 2 : 1;
    if (K < 1u || J < 1u || K > 9999u || J > 9999u || a - c < DJ) return -1;
    HostMark mark{bits};
    const uint32_t wtop = td_top_window(K);
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t p = primes[i];
        if (p <= 9999u || p >= 0x80000000u || (p & 1u) == 0u) continue;
        switch (DJ) {
        case 1: td_lattice_prime<1>(p, a, b, c, K, J, wtop, mark); break;
        case 2: td_lattice_prime<2>(p, a, b, c, K, J, wtop, mark); break;
        case 3: td_lattice_prime<3>(p, a, b, c, K, J, wtop, mark); break;
        default: td_lattice_prime<4>(p, a, b, c, K, J, wtop, mark); break;
        }
    }
    return 0;
}

// The return-time rule of td_gaps against brute force: for every v in
// [0, T), the first n >= 1 with (v + n sig) mod p < T and its landing
// residue.  p an odd prime, 1 <= sig < p, 2 <= T < p.  Returns the number of
// v where the rule and the brute force differ (-1: td_gaps failed).
extern "C" int gaps_rule_mismatches(uint32_t p, uint32_t sig, uint32_t T)
{
    TdGaps g;
    if (!td_gaps(sig, T, p, g)) return -1;
    if ((uint64_t)g.alpha + g.beta < T) return -2;
    int bad = 0;
    for (uint32_t v = 0; v < T; v++) {
        uint64_t n = 1;
        uint64_t w = ((uint64_t)v + sig) % p;
        while (w >= T) {
            n++;
            w = (w + sig) % p;
        }
        uint64_t rn, rw;
        if (v < T - g.alpha) {
            rn = g.a;
            rw = v + g.alpha;
        } else if (v >= g.beta) {
            rn = g.b;
            rw = v - g.beta;
        } else {
            rn = (uint64_t)g.a + g.b;
            rw = (uint64_t)v + g.alpha - g.beta;
        }
        if (rn != n || rw != w) bad++;
    }
    return bad;
}

// The windowed first-hit query td_first_hit against brute force: for n
// start values c (c0, c0 + 1, ...), the first t >= 0 with (c + t sig) mod p
// < T; when that t is below W the query must return it and its residue,
// otherwise any t >= W (the early exit) or none.  p an odd prime, 1 <= sig
// < p, 1 <= T < p, W >= 1.  Returns the number of mismatches.
extern "C" int first_hit_mismatches(uint32_t p, uint32_t sig, uint32_t T, uint32_t W,
                                    uint32_t c0, uint32_t n)
{
    const PfMont m = pf_mont_init(p);
    // sig^-1 in Montgomery form: sig^(p-2) R mod p
    const uint32_t sinvM = pf_mont_pow(pf_mont_in(sig, m), (uint64_t)p - 2u, m);
    int bad = 0;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t c = (uint32_t)(((uint64_t)c0 + i) % p);
        uint64_t t = 0, w = c;
        while (w >= T) {
            t++;
            w = (w + sig) % p;
        }
        uint32_t y = 0;
        const uint32_t got = td_first_hit(c, sig, T, W, sinvM, m, y);
        if (t < W) {
            if (got != t || y != w) bad++;
        } else if (got < W) {
            bad++;
        }
        // the windows just above and at the first hit: the early exit must
        // not fire while the hit is inside, and no t below W may come back
        if (t < 0xFFFFFFFFu) {
            const uint32_t in = td_first_hit(c, sig, T, (uint32_t)t + 1u, sinvM, m, y);
            if (in != t || y != w) bad++;
            if (t > 0 && td_first_hit(c, sig, T, (uint32_t)t, sinvM, m, y) < t) bad++;
        }
    }
    return bad;
}

// td_top_window(K) against 1 + the largest reversal (D_k digits) over the
// top class's k <= K; returns the first K in [1, 9999] where they differ
// (0: none).
extern "C" uint32_t top_window_mismatch(void)
{
    for (uint32_t K = 1; K <= 9999u; K++) {
        const int DK = (K >= 1000u) ? 4 : (K >= 100u) ? 3 : (K >= 10u) ? 2 : 1;
        uint32_t best = 0;
        for (uint32_t k = td_pow10(DK - 1); k <= K; k++) {
            const uint32_t t = td_reverse(k, DK);
            best = t > best ? t : best;
        }
        if (td_top_window(K) != best + 1u) return K;
    }
    return 0;
}

// The lazily reduced shared-squaring powers (pf_mont_pow_multi<NE, true>,
// products without the final subtraction, p < 2^30) against the fully
// reduced ones and against plain square-and-multiply in 64-bit arithmetic,
// for n random exponent triples below 2^20 (seeded xorshift).  Returns the
// number of mismatches.
extern "C" int lazy_pow_mismatches(uint32_t p, uint32_t seed, uint32_t n)
{
    const PfMont m = pf_mont_init(p);
    uint64_t st = 0x9E3779B97F4A7C15ull ^ seed;
    auto next = [&]() {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        return st;
    };
    int bad = 0;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t base = (uint32_t)(next() % (p - 1u)) + 1u;
        const uint64_t e[3] = {next() % (1u << 20), next() % (1u << 20), i % 7u};
        uint32_t lazy[3], full[3];
        const uint32_t bM = pf_mont_in(base, m);
        pf_mont_pow_multi<3, true>(bM, e, lazy, m);
        pf_mont_pow_multi<3>(bM, e, full, m);
        for (int k = 0; k < 3; k++) {
            uint64_t r = 1, b = base, x = e[k];
            while (x) {
                if (x & 1u) r = r * b % p;
                b = b * b % p;
                x >>= 1;
            }
            if (lazy[k] != full[k] || pf_mont_out(lazy[k], m) != (uint32_t)r) bad++;
        }
    }
    return bad;
}
