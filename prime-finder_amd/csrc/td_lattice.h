// td_lattice.h -- per-prime hit enumeration for large primes (the lattice
// kernel of td_kernels.hip), written as plain C++ over pf_mont.h so that the
// same code also compiles for the host (CPU model tests).
//
// Problem (reference trial_div_cpu.c:32-111 semantics, merged j digit
// groups, a - c >= D_j): for one prime p > 10^4 mark every (k, j) of the
// K x J grid with p | N = 10^a - k 10^b - j 10^c - 1 or p | the formula
// rev(N) = 10^a - rev(k) 10^(a-b-lk) - rev(j) 10^(a-c-lj) - 1.
//
// Each condition asks for the t in [0, W) with y(t) = (c0 + t sig) mod p < T:
//   N:          t = k - 1, y = j - 1, c0 = un(1), sig = -10^(b-c), T = J,
//               W = K;
//   rev, lk:    t = r (rev(k) written with lk digits), y = R - 1 (R = rev(j)
//               padded with zeros to D_j digits), c0 = cm, sig = -E_lk,
//               T = 10^D_j - 1, W = 10^lk; an r ending in 0, or whose k
//               exceeds K, is not a k of that class and is skipped.
// For p far above K and J a condition has about W T / p hits, so instead of
// stepping t through [0, W) the lane asks for the first hit from t0 on
// (td_first_hit, O(log p) divisions), marks it and asks again from the next
// t.  The class lk = 1 (9 values of r) is stepped directly.
#pragma once
#include "pf_mont.h"

// Work counters for host-side models (no-ops in the library): one call per
// continued-fraction level, per first-hit query and per direct step.
#ifndef TD_LAT_COUNT
#define TD_LAT_COUNT(what) ((void)0)
#endif

// 10^e for e in [0, 5]; a select chain keeps it in registers.
PF_HD uint32_t td_pow10(int e)
{
    uint32_t v = 1u;
    v = e >= 1 ? 10u : v;
    v = e >= 2 ? 100u : v;
    v = e >= 3 ? 1000u : v;
    v = e >= 4 ? 10000u : v;
    v = e >= 5 ? 100000u : v;
    return v;
}

// Reverse the lowest `len` decimal digits of x, leading zeros included.
PF_HD uint32_t td_reverse(uint32_t x, int len)
{
    uint32_t out = 0;
    for (int i = 0; i < len; i++) {
        const uint32_t q = x / 10u;
        out = out * 10u + (x - q * 10u);
        x = q;
    }
    return out;
}

// The j whose rev(j) padded with zeros to d digits is R (0 for R = 0): strip
// R's trailing zeros (lj = d minus their count) and reverse lj digits.
PF_HD uint32_t td_rdec_value(uint32_t R, int d)
{
    if (R == 0u) return 0u;
    int lj = d;
    while (R % 10u == 0u) {
        R /= 10u;
        lj--;
    }
    return td_reverse(R, lj);
}

// Smallest t >= 0 with (c + t sig) mod p < T, and y = that residue;
// 0xFFFFFFFF when there is none (only for a composite list entry).
// 1 <= T < p, c < p, 1 <= sig < p; siginvM = sig^-1 in Montgomery form.
//
// For c >= T this is the smallest x >= 1 with (sig x) mod p in [L, L + T - 1],
// L = p - c, solved by a nearest-integer continued-fraction descent: at each
// level (A, M, [L, R]) with R - L = T - 1, either a multiple of A lands in
// [L, R] without wrapping (done: w = A ceil(L / A)), or every solution wraps
// and A x - M y = w in [L, R] turns into ((-M) mod A) y mod A in
// [L mod A, R mod A] (the value w is L - (L mod A) plus the next level's);
// A > M/2 is first replaced by M - A with the interval mirrored (w -> M - w),
// so A at least halves per level.  The top-level value w is accumulated as
// +-w_level + tacc (wrapping uint32 arithmetic: the result is in [1, p)),
// and x = w sig^-1 mod p.
PF_HD uint32_t td_first_hit(uint32_t c, uint32_t sig, uint32_t T, uint32_t siginvM,
                            const PfMont &m, uint32_t &y)
{
    const uint32_t p = m.p;
    if (c < T) {
        y = c;
        return 0u;
    }
    uint32_t L = p - c, A = sig, M = p, tacc = 0u, w = 0u;
    bool neg = false;
    // A at least halves per level, so a prime p < 2^31 needs at most 32
    // levels; a composite list entry (gcd(sig, p) > 1) can make A divide M
    // or vanish, which ends the search with "no hit" instead of looping.
    TD_LAT_COUNT(query);
    for (int lvl = 0;; lvl++) {
        TD_LAT_COUNT(level);
        if (A > M - A) {                 // mirror: (M - A) x mod M in [M - R, M - L]
            const uint32_t R = L + T - 1u;
            A = M - A;
            L = M - R;
            tacc += neg ? 0u - M : M;
            neg = !neg;
        }
        if (lvl >= 40 || A == 0u) { y = 0u; return 0xFFFFFFFFu; }
        const uint32_t q = L / A;
        const uint32_t rL = L - q * A;
        if (rL == 0u) { w = L; break; }
        if (A - rL < T) { w = L + (A - rL); break; }
        tacc += neg ? rL - L : L - rL;
        const uint32_t mA = M % A;
        if (mA == 0u) { y = 0u; return 0xFFFFFFFFu; }
        M = A;
        A -= mA;
        L = rL;
    }
    const uint32_t wt = (neg ? 0u - w : w) + tacc;
    y = c + wt - p;
    return pf_mont_mul(wt, siginvM, m);
}

// Every hit of one prime p (odd, 10^4 < p < 2^31) on the K x J grid,
// a - c >= DJ (D_j = digits of J).  rdec[R] is the j whose rev(j) padded to
// DJ digits is R; mark(bit) receives bit (k - 1) J + (j - 1).
template <int DJ, class Mark>
PF_HD void td_lattice_prime(uint32_t p, long long a, long long b, long long c, uint32_t K,
                            uint32_t J, const uint16_t *rdec, Mark &mark)
{
    const PfMont m = pf_mont_init(p);
    const uint32_t ten = pf_mont_in(10u, m);
    // 10^{-1}: 10 x = 1 + t p with t p == 9 (mod 10), t fixed by p mod 10.
    const uint32_t pm = p % 10u;
    const uint32_t t = pm == 1u ? 9u : pm == 3u ? 3u : pm == 7u ? 7u : 1u;
    const uint32_t tenInv = pf_mont_in((uint32_t)((1ull + (uint64_t)t * p) / 10u), m);

    // Montgomery-form powers (suffix M) and plain residues; the powers of
    // 10 and of 10^-1 share their squarings.
    const long long eb1 = a - b - 1;
    uint32_t pw[4], nw[3];
    {
        const uint64_t pe[4] = {(uint64_t)a, (uint64_t)b, (uint64_t)(eb1 >= 0 ? eb1 : 0),
                                (uint64_t)(c >= b ? c - b : 0)};
        const uint64_t ne[3] = {(uint64_t)c, (uint64_t)(a - c - DJ), (uint64_t)(c < b ? b - c : 0)};
        pf_mont_pow_multi<4>(ten, pe, pw, m);
        pf_mont_pow_multi<3>(tenInv, ne, nw, m);
    }
    const uint32_t raM = pw[0];                                                  // 10^a
    const uint32_t rbM = pw[1];                                                  // 10^b
    const uint32_t icM = nw[0];                                                  // 10^-c
    const uint32_t ra1M = pf_submod(raM, m.one, p);                              // 10^a - 1
    const uint32_t bstepM = pf_mont_mul(rbM, icM, m);                            // 10^(b-c)
    const uint32_t bstep = pf_mont_out(bstepM, m);
    const uint32_t un1 = pf_submod(pf_submod(pf_mont_out(pf_mont_mul(ra1M, icM, m), m), 1u, p),
                                   bstep, p);                                    // j - 1 at k = 1
    const uint32_t binvM = c >= b ? pw[3] : nw[2];                               // 10^(c-b)

    // ---- N: (k, j) with j - 1 = (un1 - (k - 1) bstep) mod p < J ----
    {
        const uint32_t sig = p - bstep;
        const uint32_t siginvM = p - binvM;                 // (-10^(b-c))^-1
        uint32_t cc = un1, t0 = 0u;
        for (;;) {
            uint32_t y;
            const uint32_t x = td_first_hit(cc, sig, J, siginvM, m, y);
            if (x >= K - t0) break;                         // (no overflow: t0 < K)
            const uint32_t tk = t0 + x;
            mark(tk * J + y);                               // (k, j) = (tk + 1, y + 1)
            t0 = tk + 1u;
            if (t0 >= K) break;
            cc = y + sig;
            cc = cc >= p ? cc - p : cc;
        }
    }

    // ---- rev: per digit count lk of k, R - 1 = (cm - rev(k) E_lk) mod p ----
    if (eb1 < 0) return;                                    // no class has a rev term
    const uint32_t jinvM = nw[1];                           // 10^-(a-c-DJ)
    const uint32_t cm = pf_submod(pf_mont_out(pf_mont_mul(ra1M, jinvM, m), m), 1u, p);
    uint32_t kresM = pw[2];                                 // 10^(a-b-lk)
    const uint32_t T = td_pow10(DJ) - 1u;
    const int DK = (K >= 1000u) ? 4 : (K >= 100u) ? 3 : (K >= 10u) ? 2 : 1;
    // E_lk = 10^(a-b-lk) / 10^(a-c-DJ); E_lk^-1 = 10^(b-c) 10^(lk-DJ)
    uint32_t einvM = bstepM;
    for (int d = 1; d < DJ; d++) einvM = pf_mont_mul(einvM, tenInv, m);
    for (int lk = 1; lk <= DK; lk++) {
        if (a - b - lk < 0) break;
        const uint32_t E = pf_mont_out(pf_mont_mul(kresM, jinvM, m), m);
        const uint32_t W = td_pow10(lk);
        if (lk == 1) {
            uint32_t v = cm;
            for (uint32_t r = 1u; r < W; r++) {
                TD_LAT_COUNT(step);
                v = pf_submod(v, E, p);
                if (v < T && r % 10u != 0u) {
                    const uint32_t k = td_reverse(r, lk);
                    const uint32_t j = rdec[v + 1u];
                    if (k <= K && j <= J) mark((k - 1u) * J + (j - 1u));
                }
            }
        } else {
            const uint32_t sig = p - E;
            const uint32_t siginvM = p - einvM;
            uint32_t cc = cm, t0 = 0u;
            for (;;) {
                uint32_t y;
                const uint32_t x = td_first_hit(cc, sig, T, siginvM, m, y);
                if (x >= W - t0) break;                     // (no overflow: t0 < W)
                const uint32_t r = t0 + x;
                if (r % 10u != 0u) {
                    const uint32_t k = td_reverse(r, lk);
                    const uint32_t j = rdec[y + 1u];
                    if (k <= K && j <= J) mark((k - 1u) * J + (j - 1u));
                }
                t0 = r + 1u;
                if (t0 >= W) break;
                cc = y + sig;
                cc = cc >= p ? cc - p : cc;
            }
        }
        kresM = pf_mont_mul(kresM, tenInv, m);
        einvM = pf_mont_mul(einvM, ten, m);
    }
}
