// pf_mont.h -- 32-bit modular arithmetic shared by the gfx950 kernels and
// their host-side models (compiles as HIP device code and as plain C++).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define PF_HD __host__ __device__ __forceinline__
#else
#define PF_HD inline
#endif

// Work counters for host-side models (no-ops in the library): one call per
// Montgomery multiplication, fully or lazily reduced.
#ifndef PF_MONT_COUNT
#define PF_MONT_COUNT() ((void)0)
#endif
#ifndef PF_MONT_COUNT_LAZY
#define PF_MONT_COUNT_LAZY() PF_MONT_COUNT()
#endif

// High 32 bits of a 32 x 32-bit product (v_mul_hi_u32 on the device).
PF_HD uint32_t pf_mulhi(uint32_t a, uint32_t b)
{
    return (uint32_t)(((uint64_t)a * b) >> 32);
}

// (x - y) mod p for x, y in [0, p), p <= 2^31: the wrapped difference is
// either already in range or exceeds 2^31, in which case adding p brings it
// back; the minimum picks the in-range one.  3 VALU ops, no branch.
PF_HD uint32_t pf_submod(uint32_t x, uint32_t y, uint32_t p)
{
    const uint32_t d = x - y, e = d + p;
    return d < e ? d : e;
}

// (x + y) mod p for x, y in [0, p), p <= 2^31: the sum fits 32 bits, and the
// minimum of it and its difference with p (which wraps above 2^31 when the
// sum is below p) is the reduced value.  3 VALU ops, no branch.
PF_HD uint32_t pf_addmod(uint32_t x, uint32_t y, uint32_t p)
{
    const uint32_t s = x + y, d = s - p;
    return d < s ? d : s;
}

// Montgomery arithmetic with R = 2^32 for an odd modulus p < 2^31.
struct PfMont {
    uint32_t p;
    uint32_t nprime;  // -p^{-1} mod 2^32
    uint32_t one;     // R mod p
    uint32_t r2;      // R^2 mod p
};

PF_HD PfMont pf_mont_init(uint32_t p)
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
PF_HD uint32_t pf_mont_mul(uint32_t a, uint32_t b, const PfMont &m)
{
    PF_MONT_COUNT();
    const uint64_t t = (uint64_t)a * b;
    const uint32_t lo = (uint32_t)t;
    const uint32_t hi = (uint32_t)(t >> 32);
    const uint32_t q = lo * m.nprime;
    // t + q p is divisible by 2^32; its low word is 0 with a carry iff lo != 0.
    const uint32_t r = hi + pf_mulhi(q, m.p) + (lo != 0u ? 1u : 0u);
    return r >= m.p ? r - m.p : r;
}

// a * b * R^{-1} mod p without the final subtraction, for p < 2^30 and
// a, b < 2p: a b < 4 p^2 <= p R, so (a b + q p) / R < 2 p and the result
// stays a valid operand of the next lazy product.
PF_HD uint32_t pf_mont_mul_lazy(uint32_t a, uint32_t b, const PfMont &m)
{
    PF_MONT_COUNT_LAZY();
    const uint64_t t = (uint64_t)a * b;
    const uint32_t lo = (uint32_t)t;
    const uint32_t hi = (uint32_t)(t >> 32);
    const uint32_t q = lo * m.nprime;
    return hi + pf_mulhi(q, m.p) + (lo != 0u ? 1u : 0u);
}

PF_HD uint32_t pf_mont_in(uint32_t x, const PfMont &m) { return pf_mont_mul(x, m.r2, m); }

PF_HD uint32_t pf_mont_out(uint32_t x, const PfMont &m) { return pf_mont_mul(x, 1u, m); }

// base^e (both Montgomery form) for an exponent that is the same in every
// lane, so the bit loop is uniform control flow.
PF_HD uint32_t pf_mont_pow(uint32_t base, uint64_t e, const PfMont &m)
{
    uint32_t acc = m.one;
    if (e == 0) return acc;
    const int top = 63 - __builtin_clzll(e);
    for (int i = top; i >= 0; i--) {
        acc = pf_mont_mul(acc, acc, m);
        if ((e >> i) & 1u) acc = pf_mont_mul(acc, base, m);
    }
    return acc;
}

// base^e[k] (Montgomery form) for NE exponents that are the same in every
// lane, sharing the squarings of base: right to left over the bits of the
// largest exponent, one squaring per bit and one product per set bit after
// an exponent's lowest (which takes the power itself: R * sq / R = sq).
// LAZY (only for p < 2^30): the products skip their final subtraction
// (pf_mont_mul_lazy) and each power is reduced once at the end.
template <int NE, bool LAZY = false>
PF_HD void pf_mont_pow_multi(uint32_t base, const uint64_t (&e)[NE], uint32_t (&out)[NE],
                             const PfMont &m)
{
    uint64_t top = 0;
    for (int k = 0; k < NE; k++) {
        out[k] = m.one;
        top |= e[k];
    }
    uint32_t sq = base;
    for (int i = 0; (top >> i) != 0; i++) {
        for (int k = 0; k < NE; k++)
            if ((e[k] >> i) & 1u)
                out[k] = (e[k] & ((1ull << i) - 1u)) == 0u ? sq
                         : LAZY ? pf_mont_mul_lazy(out[k], sq, m) : pf_mont_mul(out[k], sq, m);
        if ((top >> (i + 1)) != 0) sq = LAZY ? pf_mont_mul_lazy(sq, sq, m) : pf_mont_mul(sq, sq, m);
    }
    if (LAZY)
        for (int k = 0; k < NE; k++) {
            const uint32_t d = out[k] - m.p;
            out[k] = d < out[k] ? d : out[k];
        }
}
