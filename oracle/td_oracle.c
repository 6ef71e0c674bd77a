/*
 * td_oracle.c -- CPU oracle for Phase-1 trial division.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
 * this file; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg load it, and only as the checker.
 *
 * Two independent routines are provided:
 *
 *  td_oracle_sieve  a restatement of the reference's CPU sieve
 *                   (trial_div_cpu.c:32-111), generalised from the fixed
 *                   100x100 grid to K x J with 1..4 decimal digit groups.
 *                   Per prime it solves for j algebraically (one modular
 *                   inverse per digit group), exactly as the reference does,
 *                   including its edge semantics:
 *                     - normal sieve only when 10^c mod p != 0
 *                       (trial_div_cpu.c:41);
 *                     - j0 == 0 maps to j0 = p (trial_div_cpu.c:46,76,85);
 *                     - a rev exponent a-b-l or a-c-l < 0 gives residue 0 and
 *                       disables that k (trial_div_cpu.c:57-60,68) or that
 *                       j digit group (trial_div_cpu.c:62-63,73,82,97);
 *                     - rev(x) is the digit reversal of x read as an integer
 *                       (rev(100) = 1), l_x its digit count
 *                       (trial_div_cpu.c:9-20).
 *                   The reference's l_j = 2 mapping (trial_div_cpu.c:87-93)
 *                   and its l_j = 3 special case (j = 100 only,
 *                   trial_div_cpu.c:97-100) are the K=J=100 instances of one
 *                   rule used here for every group: the l-digit j with
 *                   rev(j) = r exists iff r < 10^l and r mod 10 != 0, and is
 *                   the reversal of r written with l digits.
 *
 *  td_oracle_brute  direct per-(k, j, p) evaluation of N mod p and of the
 *                   formula rev(N) mod p (the pattern of the reference's
 *                   tests/test_gpu.c:31-73), with the same residue-0 gating.
 *                   It shares no algebra with td_oracle_sieve (no inverses,
 *                   no enumeration, no reversal mapping), so agreement of the
 *                   two pins the solve.
 *
 * Bitmap contract (sieve_table.h:15-17, sieve.c:24-31, worker.c:45-54):
 * bit (k-1)*J + (j-1), LSB-first in uint32 words, 1 = eliminated; bits are
 * OR-ed into the caller's array.
 */
#include <stdint.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Square-and-multiply with the reference's conventions (trial_div_cpu.c:22-30):
 * the result for exp == 0 is 1 even before reduction. */
static uint64_t or_powmod(uint64_t base, uint64_t exp, uint64_t mod)
{
    uint64_t acc = 1;
    uint64_t sq = base % mod;
    for (; exp; exp >>= 1) {
        if (exp & 1u) acc = acc * sq % mod;
        sq = sq * sq % mod;
    }
    return acc;
}

static int or_ndigits(uint32_t x)
{
    int n = 1;
    while (x >= 10u) { x /= 10u; n++; }
    return n;
}

/* Reverse the lowest `len` decimal digits of x (leading zeros included). */
static uint32_t or_reverse(uint32_t x, int len)
{
    uint32_t out = 0;
    for (int i = 0; i < len; i++) { out = out * 10u + x % 10u; x /= 10u; }
    return out;
}

static uint32_t or_pow10(int e)
{
    uint32_t v = 1;
    while (e-- > 0) v *= 10u;
    return v;
}

static void or_mark(uint32_t *bits, uint32_t J, uint32_t k, uint32_t j)
{
    uint64_t bit = (uint64_t)(k - 1u) * J + (j - 1u);
    bits[bit >> 5] |= 1u << (bit & 31u);
}

/* 10^e mod p for a possibly negative exponent: 0 when e < 0. */
static uint64_t or_rev_residue(long long e, uint32_t p)
{
    return (e >= 0) ? or_powmod(10u, (uint64_t)e, p) : 0u;
}

static void or_sieve_prime(uint32_t p, long long a, long long b, long long c,
                           uint32_t K, uint32_t J, uint32_t *bits)
{
    const uint64_t P = p;
    const uint64_t ra = or_powmod(10u, (uint64_t)a, P);
    const uint64_t rb = or_powmod(10u, (uint64_t)b, P);
    const uint64_t rc = or_powmod(10u, (uint64_t)c, P);

    /* N = 10^a - k 10^b - j 10^c - 1 == 0 (mod p)  <=>
     * j == (ra - k rb - 1) / rc (mod p).            (trial_div_cpu.c:41-52) */
    if (rc != 0) {
        const uint64_t rc_inv = or_powmod(rc, P - 2u, P);
        for (uint32_t k = 1; k <= K; k++) {
            uint64_t rhs = (ra + 2u * P - (uint64_t)k * rb % P - 1u) % P;
            uint64_t j = rhs * rc_inv % P;
            if (j == 0) j = P;
            for (; j <= J; j += P) or_mark(bits, J, k, (uint32_t)j);
        }
    }

    /* rev(N) = 10^a - rev(k) 10^(a-b-lk) - rev(j) 10^(a-c-lj) - 1
     * (the formula of miller_rabin.c:54-74).         (trial_div_cpu.c:55-101) */
    const int dk = or_ndigits(K), dj = or_ndigits(J);
    uint64_t kres[5] = {0}, jres[5] = {0}, jinv[5] = {0};
    for (int l = 1; l <= dk; l++) kres[l] = or_rev_residue(a - b - l, p);
    for (int l = 1; l <= dj; l++) {
        jres[l] = or_rev_residue(a - c - l, p);
        jinv[l] = jres[l] ? or_powmod(jres[l], P - 2u, P) : 0u;
    }

    for (uint32_t k = 1; k <= K; k++) {
        const int lk = or_ndigits(k);
        const uint64_t rk = or_reverse(k, lk);
        if (kres[lk] == 0) continue;
        const uint64_t rhs = (ra + 2u * P - rk * kres[lk] % P - 1u) % P;
        for (int lj = 1; lj <= dj; lj++) {
            if (jinv[lj] == 0) continue;
            const uint32_t lim = or_pow10(lj);      /* rev(j) < 10^lj */
            uint64_t r = rhs * jinv[lj] % P;
            if (r == 0) r = P;
            for (; r < lim; r += P) {
                if (r % 10u == 0) continue;        /* no lj-digit j has it */
                uint32_t j = or_reverse((uint32_t)r, lj);
                if (j <= J) or_mark(bits, J, k, j);
            }
        }
    }
}

int td_oracle_sieve(const uint32_t *primes, uint32_t n_primes,
                    long long a, long long b, long long c,
                    uint32_t K, uint32_t J, uint32_t *bits)
{
    if (!primes && n_primes) return -1;
    if (K < 1 || J < 1 || K > 9999 || J > 9999) return -1;
    for (uint32_t i = 0; i < n_primes; i++)
        if (primes[i] >= 2u) or_sieve_prime(primes[i], a, b, c, K, J, bits);
    return 0;
}

/* Direct evaluation, one (k, j, p) at a time. */
int td_oracle_brute(const uint32_t *primes, uint32_t n_primes,
                    long long a, long long b, long long c,
                    uint32_t K, uint32_t J, uint32_t *bits)
{
    if (!primes && n_primes) return -1;
    if (K < 1 || J < 1 || K > 9999 || J > 9999) return -1;
    for (uint32_t i = 0; i < n_primes; i++) {
        const uint64_t P = primes[i];
        if (P < 2u) continue;
        const uint64_t ra = or_powmod(10u, (uint64_t)a, P);
        const uint64_t rb = or_powmod(10u, (uint64_t)b, P);
        const uint64_t rc = or_powmod(10u, (uint64_t)c, P);
        for (uint32_t k = 1; k <= K; k++) {
            const int lk = or_ndigits(k);
            const uint64_t rk = or_reverse(k, lk);
            const uint64_t kr = or_rev_residue(a - b - lk, (uint32_t)P);
            for (uint32_t j = 1; j <= J; j++) {
                int hit = 0;
                if (rc != 0) {
                    uint64_t v = (ra + 3u * P - (uint64_t)k * rb % P
                                  - (uint64_t)j * rc % P - 1u) % P;
                    hit = (v == 0);
                }
                if (!hit && kr != 0) {
                    const int lj = or_ndigits(j);
                    const uint64_t jr = or_rev_residue(a - c - lj, (uint32_t)P);
                    if (jr != 0) {
                        const uint64_t rj = or_reverse(j, lj);
                        uint64_t v = (ra + 3u * P - rk * kr % P - rj * jr % P - 1u) % P;
                        hit = (v == 0);
                    }
                }
                if (hit) or_mark(bits, J, k, j);
            }
        }
    }
    return 0;
}

/* Full-list verdicts for a sample of (k, j) pairs: flags[i] = 1 iff some
 * prime of the list eliminates (ks[i], js[i]) under the bitmap contract above
 * (direct residue evaluation, as td_oracle_brute).  Threads split the prime
 * list; with nthreads <= 0 OpenMP picks the count. */
int td_oracle_check_pairs(const uint32_t *primes, uint32_t n_primes,
                          long long a, long long b, long long c,
                          const uint32_t *ks, const uint32_t *js, uint32_t n_pairs,
                          uint8_t *flags, int nthreads)
{
    if ((!primes && n_primes) || (!ks && n_pairs) || (!js && n_pairs)) return -1;
    for (uint32_t i = 0; i < n_pairs; i++) {
        flags[i] = 0;
        if (ks[i] < 1 || js[i] < 1 || ks[i] > 9999 || js[i] > 9999) return -1;
    }
    uint8_t lk[n_pairs ? n_pairs : 1], lj[n_pairs ? n_pairs : 1];
    uint32_t rk[n_pairs ? n_pairs : 1], rj[n_pairs ? n_pairs : 1];
    for (uint32_t i = 0; i < n_pairs; i++) {
        lk[i] = (uint8_t)or_ndigits(ks[i]); lj[i] = (uint8_t)or_ndigits(js[i]);
        rk[i] = or_reverse(ks[i], lk[i]);   rj[i] = or_reverse(js[i], lj[i]);
    }
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel
    {
        uint8_t local[n_pairs ? n_pairs : 1];
        memset(local, 0, sizeof local);
#pragma omp for schedule(dynamic, 4096)
        for (uint32_t pi = 0; pi < n_primes; pi++) {
            const uint64_t P = primes[pi];
            if (P < 2u) continue;
            const uint64_t ra = or_powmod(10u, (uint64_t)a, P);
            const uint64_t rb = or_powmod(10u, (uint64_t)b, P);
            const uint64_t rc = or_powmod(10u, (uint64_t)c, P);
            uint64_t kr[5], jr[5];
            for (int l = 1; l <= 4; l++) {
                kr[l] = or_rev_residue(a - b - l, (uint32_t)P);
                jr[l] = or_rev_residue(a - c - l, (uint32_t)P);
            }
            for (uint32_t i = 0; i < n_pairs; i++) {
                if (local[i]) continue;
                if (rc != 0) {
                    uint64_t v = (ra + 3u * P - (uint64_t)ks[i] * rb % P
                                  - (uint64_t)js[i] * rc % P - 1u) % P;
                    if (v == 0) { local[i] = 1; continue; }
                }
                const uint64_t x = kr[lk[i]], y = jr[lj[i]];
                if (x != 0 && y != 0) {
                    uint64_t v = (ra + 3u * P - rk[i] * x % P - rj[i] * y % P - 1u) % P;
                    if (v == 0) local[i] = 1;
                }
            }
        }
#pragma omp critical
        for (uint32_t i = 0; i < n_pairs; i++) flags[i] |= local[i];
    }
    return 0;
}

/* td_oracle_sieve with the prime list split over threads, each into a
 * private bitmap, OR-ed at the end.  Same result as td_oracle_sieve. */
int td_oracle_sieve_mt(const uint32_t *primes, uint32_t n_primes,
                       long long a, long long b, long long c,
                       uint32_t K, uint32_t J, uint32_t *bits, int nthreads)
{
    if (!primes && n_primes) return -1;
    if (K < 1 || J < 1 || K > 9999 || J > 9999) return -1;
    const size_t nw = ((size_t)K * J + 31u) / 32u;
    int fail = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel
    {
        uint32_t *mine = (uint32_t *)calloc(nw, sizeof(uint32_t));
        if (!mine) {
#pragma omp atomic write
            fail = 1;
        }
#pragma omp for schedule(dynamic, 256)
        for (uint32_t i = 0; i < n_primes; i++)
            if (mine && primes[i] >= 2u) or_sieve_prime(primes[i], a, b, c, K, J, mine);
#pragma omp critical
        if (mine) for (size_t w = 0; w < nw; w++) bits[w] |= mine[w];
        free(mine);
    }
    return fail ? -1 : 0;
}
