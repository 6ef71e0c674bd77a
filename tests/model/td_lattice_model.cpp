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
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t p = primes[i];
        if (p <= 9999u || p >= 0x80000000u || (p & 1u) == 0u) continue;
        switch (DJ) {
        case 1: td_lattice_prime<1>(p, a, b, c, K, J, mark); break;
        case 2: td_lattice_prime<2>(p, a, b, c, K, J, mark); break;
        case 3: td_lattice_prime<3>(p, a, b, c, K, J, mark); break;
        default: td_lattice_prime<4>(p, a, b, c, K, J, mark); break;
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
