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
