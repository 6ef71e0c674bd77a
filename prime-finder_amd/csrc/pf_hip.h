// pf_hip.h -- shared HIP helpers for the gfx950 kernels of this library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

// Evaluate a HIP call; on failure print "<where>: HIP error <name> (<call>)"
// and jump to the enclosing function's `fail:` label.
#define PF_TRY(where, call)                                                    \
    do {                                                                       \
        hipError_t pf_e_ = (call);                                             \
        if (pf_e_ != hipSuccess) {                                             \
            fprintf(stderr, "%s: HIP error %s (%s)\n", where,                  \
                    hipGetErrorName(pf_e_), #call);                            \
            goto fail;                                                         \
        }                                                                      \
    } while (0)

// ---------------------------------------------------------------------------
// 32-bit modular arithmetic on VALU.
// ---------------------------------------------------------------------------

// (x - y) mod p for x, y in [0, p), p <= 2^31: the wrapped difference is
// either already in range or exceeds 2^31, in which case adding p brings it
// back; min() picks the in-range one.  3 VALU ops, no branch.
__device__ __forceinline__ uint32_t pf_submod(uint32_t x, uint32_t y, uint32_t p)
{
    uint32_t d = x - y;
    return min(d, d + p);
}

// Montgomery arithmetic with R = 2^32 for an odd modulus p < 2^31.
struct PfMont {
    uint32_t p;
    uint32_t nprime;  // -p^{-1} mod 2^32
    uint32_t one;     // R mod p
    uint32_t r2;      // R^2 mod p
};

__device__ __forceinline__ PfMont pf_mont_init(uint32_t p)
{
    PfMont m;
    m.p = p;
    uint32_t inv = p;                 // correct to 3 bits for odd p
    inv *= 2u - p * inv;              // 6
    inv *= 2u - p * inv;              // 12
    inv *= 2u - p * inv;              // 24
    inv *= 2u - p * inv;              // 48 >= 32
    m.nprime = 0u - inv;
    m.one = (0u - p) % p;             // 2^32 mod p
    m.r2 = (uint32_t)(((uint64_t)m.one * m.one) % p);
    return m;
}

// a * b * R^{-1} mod p, for a, b < p.
__device__ __forceinline__ uint32_t pf_mont_mul(uint32_t a, uint32_t b, const PfMont &m)
{
    uint64_t t = (uint64_t)a * b;
    uint32_t lo = (uint32_t)t;
    uint32_t hi = (uint32_t)(t >> 32);
    uint32_t q = lo * m.nprime;
    // t + q p is divisible by 2^32; its low word is 0 with a carry iff lo != 0.
    uint32_t r = hi + __umulhi(q, m.p) + (lo != 0u ? 1u : 0u);
    return r >= m.p ? r - m.p : r;
}

__device__ __forceinline__ uint32_t pf_mont_in(uint32_t x, const PfMont &m)
{
    return pf_mont_mul(x, m.r2, m);
}

__device__ __forceinline__ uint32_t pf_mont_out(uint32_t x, const PfMont &m)
{
    return pf_mont_mul(x, 1u, m);
}

// base^e (both Montgomery form) for an exponent that is the same in every
// lane, so the bit loop is uniform control flow.
__device__ __forceinline__ uint32_t pf_mont_pow(uint32_t base, uint64_t e, const PfMont &m)
{
    uint32_t acc = m.one;
    if (e == 0) return acc;
    int top = 63 - __clzll((long long)e);
    for (int i = top; i >= 0; i--) {
        acc = pf_mont_mul(acc, acc, m);
        if ((e >> i) & 1u) acc = pf_mont_mul(acc, base, m);
    }
    return acc;
}
