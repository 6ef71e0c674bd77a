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

// td_rdec_value for R < 10^4 and d <= 4 without a loop or a table: the
// reversal of R written with d digits, whose leading zeros (R's trailing
// zeros) vanish in its value, so it is j itself (R = 1230, d = 4: 0321 =
// 321 = the reversal of 123).  The digits come from multiply-shift division
// by 10 (exact below 43 690).  Used on the hit paths, where a table load's
// latency would stall the whole wave.
PF_HD uint32_t td_rdec_arith(uint32_t R, int d)
{
    const uint32_t q1 = (R * 52429u) >> 19;
    const uint32_t d0 = R - 10u * q1;
    const uint32_t q2 = (q1 * 52429u) >> 19;
    const uint32_t d1 = q1 - 10u * q2;
    const uint32_t q3 = (q2 * 52429u) >> 19;
    const uint32_t d2 = q2 - 10u * q3;
    const uint32_t d3 = q3;
    const uint32_t v2 = d0 * 10u + d1, v3 = v2 * 10u + d2;
    return d <= 1 ? d0 : d == 2 ? v2 : d == 3 ? v3 : v3 * 10u + d3;
}

// ---- first-hit query, in steps ----------------------------------------
//
// The smallest t >= 0 with (c + t sig) mod p < T (1 <= T < p, c < p,
// 1 <= sig < p) and y = that residue.  For c >= T this is the smallest
// x >= 1 with (sig x) mod p in [L, L + T - 1], L = p - c, found by a
// nearest-integer continued-fraction descent: at each level (A, M, [L, R])
// with R - L = T - 1, either a multiple of A lands in [L, R] without
// wrapping (done: w = A ceil(L / A)), or every solution wraps and
// A x - M y = w in [L, R] turns into ((-M) mod A) y mod A in
// [L mod A, R mod A] (the value w is L - (L mod A) plus the next level's);
// A > M/2 is first replaced by M - A with the interval mirrored (w -> M - w),
// so A at least halves per level: a prime p < 2^31 needs at most 32 levels.
// The top-level value w is accumulated as +-w_level + tacc (wrapping uint32
// arithmetic: the result is in [1, p)), and x = w sig^-1 mod p.  A composite
// list entry (gcd(sig, p) > 1) can make A divide M or vanish; the search
// then ends with "no hit" instead of looping.
//
// A query that only needs a hit below W stops early: every solution of an
// undecided level wraps (its next-level unknown is at least 1), and the
// unknowns of successive levels satisfy x_i >= x_(i+1) M_i / A_i with
// M_(i+1) = A_i, so after undecided levels 0 .. k the product telescopes to
// x >= p / A_k.  Once A_k W <= p no hit lies in [0, W): the query ends with
// "no hit" there, without descending to its exact first hit.
struct TdFh {
    uint32_t c;                 // the query's start value
    uint32_t L, A, M, tacc, w;
    bool neg, live, fail;       // live: levels still to run; fail: no hit (none
                                // below the window, or a composite entry)
    bool imm;                   // c < T: t = 0 is the hit
};

// A query needs at most 32 levels for p < 2^31; a composite list entry that
// never decides is stopped after this many (wave-uniform loop bound).
constexpr int TD_FH_MAX_LEVELS = 40;

PF_HD void td_fh_start(TdFh &q, uint32_t c, uint32_t sig, uint32_t T, uint32_t p)
{
    q.c = c;
    q.L = p - c;
    q.A = sig;
    q.M = p;
    q.tacc = 0u;
    q.w = 0u;
    q.neg = false;
    q.fail = false;
    q.imm = c < T;
    q.live = !q.imm;
    if (q.imm) TD_LAT_COUNT(query);
}

// 1/A (A >= 1) scaled by 1 - 2^-20, so that trunc(x * td_rcp_lo(A)) never
// exceeds floor(x / A) for x < 2^31: the float rounding of x, of A and of
// the reciprocal (v_rcp_f32: 1 ulp) and of the product add up to less than
// 5 * 2^-24 relative, which the 2^-20 bias covers.
PF_HD float td_rcp_lo(uint32_t A)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const float r = __builtin_amdgcn_rcpf((float)A);
#else
    const float r = 1.0f / (float)A;
#endif
    return r * 0.99999904632568359375f;          // 1 - 2^-20, exact in binary32
}

// x mod A from the biased reciprocal: the estimate q of floor(x / A) is that
// quotient or one less while the quotient is below 2^19, which the caller
// guarantees by checking est < 2^18 (an estimate below 2^18 implies a
// quotient below 2^19); then x - q A is in [0, 2 A) and one conditional
// subtraction (as a minimum of unsigned values) finishes it.
PF_HD uint32_t td_mod_est(uint32_t x, uint32_t A, float rA, uint32_t &est)
{
    est = (uint32_t)((float)x * rA);
    const uint32_t r = x - est * A;
    const uint32_t d = r - A;
    return d < r ? d : r;
}

// floor(p / W) or less, for p < 2^31 and W >= 1: a query's early-exit
// threshold on A (the biased reciprocal's estimate never exceeds the
// quotient, and a lower threshold only exits later).
PF_HD uint32_t td_window_thr(uint32_t p, uint32_t W)
{
    return (uint32_t)((float)p * td_rcp_lo(W));
}

// The levels of a started query, until it is decided (or failed after
// TD_FH_MAX_LEVELS: the bound is the same in every lane, so the loop
// counter is wave-uniform).  Per level, with A > M/2 first mirrored:
// L mod A and M mod A from one biased reciprocal (exact division only for
// a partial quotient of 2^18 or more, which is rare), then either the next
// multiple of A lands in [L, L + T - 1] (decided) or the level moves to
// (A - M mod A, A, L mod A).  (M mod A == 0 makes the next level's A
// vanish, which fails there.)  athr = td_window_thr(p, W) ends the query
// with "no hit" once the first hit is known to be W or more (0: never).
PF_HD void td_fh_run(TdFh &q, uint32_t T, uint32_t athr)
{
    if (!q.live) return;
    // (the sign and the outcome are kept as 32-bit values, not booleans, so
    // that they stay in vector registers: lane-mask booleans carried across
    // the divergent exits cost several scalar mask operations per level)
    uint32_t L0 = q.L, A0 = q.A, M = q.M, tacc = q.tacc, w = 0u;
    uint32_t negm = q.neg ? ~0u : 0u;    // all ones: this level's value counts negated
    uint32_t st = 0u;                    // 1: decided, 2: no hit, 0: out of levels
    for (int it = 0; it < TD_FH_MAX_LEVELS; it++) {
        TD_LAT_COUNT(level);
        const uint32_t MA = M - A0;
        const uint32_t mirm = A0 > MA ? ~0u : 0u;   // mirror: (M - A) x mod M in [M - R, M - L]
        const uint32_t A = mirm ? MA : A0;
        const uint32_t L = mirm ? M - (L0 + (T - 1u)) : L0;
        tacc += ((M ^ negm) - negm) & mirm;          // +-M when mirrored
        negm ^= mirm;
        const uint32_t Ad = A ? A : 1u;
        const float rA = td_rcp_lo(Ad);
        uint32_t eL, eM;
        uint32_t rL = td_mod_est(L, Ad, rA, eL);
        uint32_t mA = td_mod_est(M, Ad, rA, eM);
        if ((eL | eM) >= (1u << 18)) {
            rL = L % Ad;
            mA = M % Ad;
        }
        const uint32_t up = rL ? Ad - rL : 0u;   // distance to the next multiple of A
        if (up < T || A == 0u) {                 // decided: a multiple lands in [L, R]
            st = A ? 1u : 2u;
            w = L + up;
            break;
        }
        tacc += ((L - rL) ^ negm) - negm;        // +-(L - L mod A)
        M = A;
        A0 = A - mA;
        L0 = rL;
        if (M <= athr) {                         // first hit >= p / A >= W
            st = 2u;
            break;
        }
    }
    TD_LAT_COUNT(query);
    q.tacc = tacc;
    q.neg = negm != 0u;
    q.w = w;
    q.fail = st != 1u;
    q.live = false;
}

// The decided query's t (0xFFFFFFFF: none) and residue y.
PF_HD uint32_t td_fh_result(const TdFh &q, uint32_t siginvM, const PfMont &m, uint32_t &y)
{
    if (q.imm) {
        y = q.c;
        return 0u;
    }
    if (q.fail) {
        y = 0u;
        return 0xFFFFFFFFu;
    }
    const uint32_t wt = (q.neg ? 0u - q.w : q.w) + q.tacc;
    y = q.c + wt - m.p;
    return pf_mont_mul(wt, siginvM, m);
}

// Smallest t >= 0 with (c + t sig) mod p < T, and y = that residue, when
// that t is below W (W >= 1); otherwise t >= W or 0xFFFFFFFF (no hit at
// all: only for a composite list entry).  siginvM = sig^-1 in Montgomery
// form.
PF_HD uint32_t td_first_hit(uint32_t c, uint32_t sig, uint32_t T, uint32_t W,
                            uint32_t siginvM, const PfMont &m, uint32_t &y)
{
    TdFh q;
    td_fh_start(q, c, sig, T, m.p);
    td_fh_run(q, T, td_window_thr(m.p, W));
    return td_fh_result(q, siginvM, m, y);
}

// ---- return times of the window ------------------------------------------
//
// Three-gap structure of the returns of a rotation to an interval: with
// e(n) = n sig mod p, let a = min{n >= 1 : e(n) < T} with alpha = e(a), and
// b = min{n >= 1 : e(n) > p - T} with beta = p - e(b).  Then alpha + beta >= T,
// and from any v in [0, T) the first n >= 1 with (v + e(n)) mod p < T is
//   a      (landing at v + alpha)          if v < T - alpha,
//   b      (landing at v - beta)           if v >= beta,
//   a + b  (landing at v + alpha - beta)   otherwise.
// So after a hit of a search every further hit is one compare-and-add step
// instead of a first-hit query.
struct TdGaps {
    uint32_t a, alpha, b, beta;
};

// floor(x / Y) for x < 2^31, Y >= 1, from td_rcp_lo's biased reciprocal:
// exact while the estimate is below 2^18 (then the quotient is the estimate
// or one more); `slow` is set otherwise, and the caller divides exactly.
PF_HD uint32_t td_div_est(uint32_t x, uint32_t Y, float rY, bool &slow)
{
    const uint32_t e = (uint32_t)((float)x * rY);
    slow = slow || e >= (1u << 18);
    return e + (x - e * Y >= Y ? 1u : 0u);
}

// The return times of [0, T) under sig (2 <= T < p), in one pass: the
// one-sided best approximations of sig / p are (n, e(n)) = (a, alpha) from
// above 0 and (b, p - beta) from below p, starting from (1, sig) and
// (1, p - sig); the larger distance is reduced by the smaller one as often
// as it stays positive (a run of the slow continued fraction, n adding up
// alongside; a beta + b alpha = p throughout), and the first value of a run
// below T, if any, is that side's first return.  About 7.7 division steps
// at T = 4 096 or 9 999 and p < 10^8, against 10.3 levels for two first-hit
// queries.  False when a run cannot proceed (a composite list entry).
PF_HD bool td_gaps(uint32_t sig, uint32_t T, uint32_t p, TdGaps &g)
{
    // (the flags are 32-bit values and the runs' updates selects; only the
    // rare first return and exact division take a branch)
    uint32_t a = 1u, al = sig, b = 1u, be = p - sig;
    uint32_t fa = al < T ? 1u : 0u, fb = be < T ? 1u : 0u, bad = 0u;
    uint32_t ga = 1u, galpha = al, gb = 1u, gbeta = be;
    for (int it = 0; it < 64; it++) {
        if ((fa & fb) | bad) break;
        TD_LAT_COUNT(gapcf);
        const bool up = al > be;                 // reduce alpha's side
        const uint32_t X = up ? al : be, Y0 = up ? be : al;
        const uint32_t cX = up ? a : b, cY = up ? b : a;
        bad = Y0 == 0u ? 1u : 0u;                // (a composite entry)
        const uint32_t Y = Y0 ? Y0 : 1u;
        const float rY = td_rcp_lo(Y);
        bool slow = false;
        uint32_t k = td_div_est(X - 1u, Y, rY, slow);          // length of the run
        if (slow) k = (X - 1u) / Y;
        const uint32_t nK = cX + k * cY, dK = X - k * Y;        // the run's last value
        // this side's first return lies inside the run iff the run ends
        // below T (X >= T while the side is open); its position j is only
        // computed then (rare: a branch the wave mostly skips), exactly
        // when either estimate reached 2^18.
        const bool need = (up ? fa : fb) == 0u;
        if (need && dK < T && !bad) {
            bool s2 = slow;
            uint32_t j = td_div_est(X - T, Y, rY, s2) + 1u;
            if (s2) j = (X - T) / Y + 1u;
            const uint32_t nJ = cX + j * cY, dJ = X - j * Y;
            if (up) {
                fa = 1u;
                ga = nJ;
                galpha = dJ;
            } else {
                fb = 1u;
                gb = nJ;
                gbeta = dJ;
            }
        }
        al = up ? dK : al;
        a = up ? nK : a;
        be = up ? be : dK;
        b = up ? b : nK;
    }
    g.a = ga;
    g.alpha = galpha;
    g.b = gb;
    g.beta = gbeta;
    return (fa & fb) != 0u && !bad;
}

// ---- per-prime setup and conditions -----------------------------------
//
// Conditions of one prime: cond 0 is N, cond 1..3 the rev classes
// l_k = 2..4 (the class l_k = 1, 9 values of r, is stepped in the setup).
constexpr int TD_LAT_CONDS = 4;

// Per-prime constants of the lattice path: the start value, step and step
// inverse of every condition to be searched.
struct TdLatRec {
    uint32_t p, nprime;
    uint32_t un1, sigN, sinvN;       // N: start value, step, step^-1 (Montgomery)
    uint32_t cm;                     // rev: start value of every class
    uint32_t sig[TD_LAT_CONDS - 1];  // rev classes l_k = 2..4: step, step^-1
    uint32_t sinv[TD_LAT_CONDS - 1];
    uint32_t flags;                  // bit i: condition i is to be searched
};

// The search of one condition: hits t in [0, W) with (c0 + t sig) mod p < T.
struct TdLatItem {
    uint32_t c0, sig, sinv, T, W;
};

// The search window of the top rev class (l_k = D_k, the digit count of
// K): 1 + the largest t = rev(k) over its k <= K, chosen digit by digit
// from k's last digit (t's leading digit) on, each the largest digit that
// some k of [10^(D_k-1), K] with the digits chosen so far can take: below
// the top, the smallest such k is 10^(D_k-1) + the chosen digits; at the
// top, the digits themselves.  A full class (K = 10^D_k - 1) keeps 10^D_k;
// K = 100 leaves t = 1 (k = 100) and a window of 2 instead of 1 000.
// K in [1, 9999].
PF_HD uint32_t td_top_window(uint32_t K)
{
    const int DK = (K >= 1000u) ? 4 : (K >= 100u) ? 3 : (K >= 10u) ? 2 : 1;
    const uint32_t lo = td_pow10(DK - 1);
    uint32_t s = 0u, t = 0u;                 // chosen low digits of k; t so far
    for (int m = 0; m < DK; m++) {
        const uint32_t pm = td_pow10(m);
        const uint32_t room = m + 1 < DK ? K - lo - s : K - s;
        const uint32_t d = room / pm > 9u ? 9u : room / pm;
        s += d * pm;
        t = t * 10u + d;
    }
    return t + 1u;
}

// Condition cond's search; wtop = td_top_window(K) bounds the top class.
PF_HD bool td_lattice_item(const TdLatRec &r, int cond, uint32_t K, uint32_t J, uint32_t T_rev,
                           uint32_t wtop, TdLatItem &it)
{
    if (!((r.flags >> cond) & 1u)) return false;
    if (cond == 0) {
        it = TdLatItem{r.un1, r.sigN, r.sinvN, J, K};
    } else {
        const bool top = td_pow10(cond + 1) > K;            // l_k = D_k
        it = TdLatItem{r.cm, r.sig[cond - 1], r.sinv[cond - 1], T_rev,
                       top ? wtop : td_pow10(cond + 1)};
    }
    return true;
}

// Mark hit t (residue y) of condition cond when it names a pair of the grid:
// N: (k, j) = (t + 1, y + 1); rev class l_k: k = t reversed with l_k digits
// (t ending in 0 names no k of the class), j = the j whose rev(j) padded
// to DJ digits is y + 1.
template <int DJ, class Mark>
PF_HD void td_lattice_mark(int cond, uint32_t t, uint32_t y, uint32_t K, uint32_t J,
                           Mark &mark)
{
    TD_LAT_COUNT(hit);
    if (cond == 0) {
        mark(t * J + y);
        return;
    }
    if (t % 10u == 0u) return;
    const uint32_t k = td_reverse(t, cond + 1);
    const uint32_t j = td_rdec_arith(y + 1u, DJ);
    if (k <= K && j <= J) mark((k - 1u) * J + (j - 1u));
}

// Setup of one prime p (odd, 10^4 < p < 2^31) on the K x J grid, a - c >= DJ
// (D_j = digits of J): the record r, and the marks of the rev class l_k = 1.
// mark(bit) receives bit (k - 1) J + (j - 1).
template <int DJ, class Mark>
PF_HD void td_lattice_setup(uint32_t p, long long a, long long b, long long c, uint32_t K,
                            uint32_t J, Mark &mark, TdLatRec &r)
{
    const PfMont m = pf_mont_init(p);          // (its R^2 mod p is not needed here)
    // 10 and 10^-1 in Montgomery form without R^2 mod p: 10 R by doubling
    // R mod p, and R / 10 = (R + k p) / 10 with k < 10 the multiple of p
    // that makes R + k p divisible by 10 (k = -R p^-1 mod 10; p is prime to
    // 10), divided exactly: halve, then multiply by 5^-1 mod 2^32 (the
    // quotient is below p < 2^31).
    const uint32_t two = pf_addmod(m.one, m.one, p), four = pf_addmod(two, two, p);
    const uint32_t ten = pf_addmod(pf_addmod(four, four, p), two, p);
    const uint32_t pm = p % 10u;
    const uint32_t pinv10 = pm == 3u ? 7u : pm == 7u ? 3u : pm;          // p^-1 mod 10
    const uint32_t r10 = m.one - 10u * (uint32_t)(((uint64_t)m.one * 0xCCCCCCCDull) >> 35);
    const uint32_t kx = (r10 ? 10u - r10 : 0u) * pinv10;                  // < 90
    const uint32_t k10 = kx - 10u * ((kx * 52429u) >> 19);
    const uint64_t X = (uint64_t)m.one + (uint64_t)k10 * p;               // 10 | X, X < 10 p
    const uint32_t tenInv = (uint32_t)(X >> 1) * 0xCCCCCCCDu;

    // Montgomery-form powers (suffix M) and plain residues; the powers of
    // 10 and of 10^-1 share their squarings, and 10^(b-c) and 10^(c-b) are
    // one exponent |b - c| in each chain.
    const long long eb1 = a - b - 1;
    const uint64_t bc = (uint64_t)(c <= b ? b - c : c - b);
    // (lazily reduced products where every prime of the wave is below 2^30)
    uint32_t pw[2], nw[3];
    {
        const uint64_t pe[2] = {(uint64_t)a, bc};
        const uint64_t ne[3] = {(uint64_t)c, (uint64_t)(a - c - DJ), bc};
#if defined(__HIP_DEVICE_COMPILE__)
        const bool lazy = __ballot(p >= (1u << 30)) == 0ull;
#else
        const bool lazy = p < (1u << 30);
#endif
        if (lazy) {
            pf_mont_pow_multi<2, true>(ten, pe, pw, m);
            pf_mont_pow_multi<3, true>(tenInv, ne, nw, m);
        } else {
            pf_mont_pow_multi<2>(ten, pe, pw, m);
            pf_mont_pow_multi<3>(tenInv, ne, nw, m);
        }
    }
    const uint32_t raM = pw[0];                                                  // 10^a
    const uint32_t icM = nw[0];                                                  // 10^-c
    const uint32_t ra1M = pf_submod(raM, m.one, p);                              // 10^a - 1
    const uint32_t bstepM = c <= b ? pw[1] : nw[2];                              // 10^(b-c)
    const uint32_t binvM = c <= b ? nw[2] : pw[1];                               // 10^(c-b)
    const uint32_t bstep = pf_mont_out(bstepM, m);

    // N: (k, j) with j - 1 = (un1 - (k - 1) bstep) mod p < J
    r.p = p;
    r.nprime = m.nprime;
    r.un1 = pf_submod(pf_submod(pf_mont_out(pf_mont_mul(ra1M, icM, m), m), 1u, p), bstep, p);
    r.sigN = p - bstep;
    r.sinvN = p - binvM;                                    // (-10^(b-c))^-1
    r.flags = 1u;
    r.cm = 0u;
    for (int i = 0; i < TD_LAT_CONDS - 1; i++) r.sig[i] = r.sinv[i] = 0u;

    // rev: per digit count lk of k, R - 1 = (cm - rev(k) E_lk) mod p with
    // E_lk = 10^(a-b-lk) / 10^(a-c-DJ) = 10^(c-b+DJ-lk), for the classes
    // lk = 1 .. LK that exist on the grid and keep a rev term (a - b >= lk)
    if (eb1 < 0) return;                                    // no class has a rev term
    const uint32_t jinvM = nw[1];                           // 10^-(a-c-DJ)
    const uint32_t cm = pf_submod(pf_mont_out(pf_mont_mul(ra1M, jinvM, m), m), 1u, p);
    r.cm = cm;
    const uint32_t T = td_pow10(DJ) - 1u;
    const int DK = (K >= 1000u) ? 4 : (K >= 100u) ? 3 : (K >= 10u) ? 2 : 1;
    const int LK = a - b < DK ? (int)(a - b) : DK;
    // E_LK and E_LK^-1 from 10^(c-b) and 10^(b-c), then one factor of 10
    // per class downwards
    uint32_t eM = binvM, einvM = bstepM;
    for (int d = LK; d < DJ; d++) {
        eM = pf_mont_mul(eM, ten, m);
        einvM = pf_mont_mul(einvM, tenInv, m);
    }
    for (int d = DJ; d < LK; d++) {
        eM = pf_mont_mul(eM, tenInv, m);
        einvM = pf_mont_mul(einvM, ten, m);
    }
    // (unrolled: the record's arrays are indexed by constants, so they stay
    // in registers)
#if defined(__HIPCC__)
#pragma unroll
#endif
    for (int lk = TD_LAT_CONDS; lk >= 1; lk--) {
        if (lk > LK) continue;
        const uint32_t E = pf_mont_out(eM, m);
        if (lk == 1) {
            uint32_t v = cm;
            for (uint32_t rr = 1u; rr < 10u; rr++) {
                TD_LAT_COUNT(step);
                v = pf_submod(v, E, p);
                if (v < T) {
                    const uint32_t j = td_rdec_arith(v + 1u, DJ);
                    if (rr <= K && j <= J) mark((rr - 1u) * J + (j - 1u));
                }
            }
        } else {
            r.sig[lk - 2] = p - E;
            r.sinv[lk - 2] = p - einvM;
            r.flags |= 1u << (lk - 1);
            eM = pf_mont_mul(eM, ten, m);
            einvM = pf_mont_mul(einvM, tenInv, m);
        }
    }
}

// Every hit of one condition, in order of t: the first hit by a first-hit
// query, then the window's return times (td_gaps) step from hit
// to hit.  A search expecting fewer than half a hit (2 W T < p) looks for
// its next hits by queries too, and takes the return times only from a
// fifth (a lane that needs the return times makes its whole wave wait for
// two queries, which pays only where most lanes have further hits; replayed
// over waves of 64 primes of the 4096 x 4096 workload this issues 7 % fewer
// instructions than return times from the second hit on with W T >= 2 p).
template <int DJ, class Mark>
PF_HD void td_lattice_run(const TdLatRec &r, int cond, const TdLatItem &it, uint32_t K, uint32_t J,
                          Mark &mark)
{
    PfMont m;
    m.p = r.p;
    m.nprime = r.nprime;
    m.one = m.r2 = 0u;                  // (not used by the queries)
    const bool many = 2ull * it.W * it.T >= r.p;
    uint32_t cc = it.c0, t0 = 0u, t = 0u, y = 0u;
    for (int n = 0;; n++) {
        const uint32_t x = td_first_hit(cc, it.sig, it.T, it.W - t0, it.sinv, m, y);
        if (x >= it.W - t0) return;                         // (no overflow: t0 < W)
        t = t0 + x;
        td_lattice_mark<DJ>(cond, t, y, K, J, mark);
        t0 = t + 1u;
        if (t0 >= it.W) return;
        if (it.T >= 2u && (many || n == 3)) break;
        cc = y + it.sig;
        cc = cc >= r.p ? cc - r.p : cc;
    }
    TdGaps g;
    if (!td_gaps(it.sig, it.T, r.p, g)) return;            // (composite entry: stop)
    // steps capped at W (any step reaching W ends the search): t stays below 2 W
    const uint32_t da = g.a < it.W ? g.a : it.W;
    const uint32_t db = g.b < it.W ? g.b : it.W;
    const uint32_t dab = (uint64_t)g.a + g.b < it.W ? g.a + g.b : it.W;
    const uint32_t ya = it.T - g.alpha;
    for (;;) {
        TD_LAT_COUNT(gapstep);
        // (selects, not branches: y + alpha, y - beta or y + alpha - beta)
        const bool s1 = y < ya, s2 = y >= g.beta;
        t += s1 ? da : s2 ? db : dab;
        y += s1 ? g.alpha : s2 ? 0u - g.beta : g.alpha - g.beta;
        if (t >= it.W) return;
        td_lattice_mark<DJ>(cond, t, y, K, J, mark);
    }
}

// Every hit of one prime, condition by condition (wtop =
// td_top_window(K)).  (Chaining the four
// searches into one loop, so that a lane whose search is exhausted starts
// its next condition without waiting for the rest of its wave, issued 5 %
// fewer VALU instructions but ran 3 % longer on the 4096 x 4096 grid.)
template <int DJ, class Mark>
PF_HD void td_lattice_prime(uint32_t p, long long a, long long b, long long c, uint32_t K,
                            uint32_t J, uint32_t wtop, Mark &mark)
{
    TdLatRec r;
    td_lattice_setup<DJ>(p, a, b, c, K, J, mark, r);
    // (unrolled, so that the record's arrays are indexed by constants and
    // stay in registers)
#if defined(__HIPCC__)
#pragma unroll
#endif
    for (int cond = 0; cond < TD_LAT_CONDS; cond++) {
        TdLatItem it;
        if (td_lattice_item(r, cond, K, J, td_pow10(DJ) - 1u, wtop, it))
            td_lattice_run<DJ>(r, cond, it, K, J, mark);
    }
}
