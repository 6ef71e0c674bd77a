/*
 * mr_oracle.c -- GMP oracle for Phase-2 (Miller-Rabin on N and rev(N)).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg load it, and only as the checker.
 *
 * Big integers cross this interface as little-endian uint32 limb arrays.
 *
 *  - N = 10^a - k 10^b - j 10^c - 1 exactly (miller_rabin.c:32-46,
 *    candidate.cuh:131-137), with GMP mpz_ui_pow_ui / mul_ui / sub.
 *  - rev(N): mode 1 is the decimal-string reversal of N, read back with
 *    mpz_set_str; mode 2 is the reference's closed formula
 *    10^a - rev(k) 10^(a-b-lk) - rev(j) 10^(a-c-lj) - 1
 *    (miller_rabin.c:54-74), generalised to any digit count.
 *  - d, s with N - 1 = 2^s d, d odd (candidate.cuh:95-104).
 *  - Strong-probable-prime round to one base: r = base^d mod N by
 *    mpz_powm; pass iff r == 1, r == N-1, or r^(2^i) == N-1 for some
 *    1 <= i < s (miller_rabin_runner.cu:281-473 semantics).
 *  - The reference's production verdict mpz_probab_prime_p(N, reps)
 *    (miller_rabin.c:91,104) is exposed separately; on GMP >= 6.2 it is
 *    Baillie-PSW, not per-base Miller-Rabin.
 */
#include <gmp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static void mo_from_limbs(mpz_t x, const uint32_t *v, size_t n)
{
    mpz_import(x, n, -1, sizeof(uint32_t), 0, 0, v);
}

/* Writes |x| into out[0..cap), zero-padded; returns the number of
 * significant limbs, or -1 when cap is too small. */
static long mo_to_limbs(const mpz_t x, uint32_t *out, size_t cap)
{
    size_t cnt = (mpz_sizeinbase(x, 2) + 31) / 32;
    if (mpz_sgn(x) == 0) cnt = 0;
    if (cnt > cap) return -1;
    memset(out, 0, cap * sizeof(uint32_t));
    size_t w = 0;
    mpz_export(out, &w, -1, sizeof(uint32_t), 0, 0, x);
    return (long)w;
}

static int mo_ndigits(unsigned long x)
{
    int n = 1;
    while (x >= 10ul) { x /= 10ul; n++; }
    return n;
}

static unsigned long mo_reverse(unsigned long x, int len)
{
    unsigned long out = 0;
    for (int i = 0; i < len; i++) { out = out * 10ul + x % 10ul; x /= 10ul; }
    return out;
}

static void mo_build_n(mpz_t out, long long a, long long b, long long c,
                       unsigned long k, unsigned long j)
{
    mpz_t t;
    mpz_init(t);
    mpz_ui_pow_ui(out, 10u, (unsigned long)a);
    mpz_ui_pow_ui(t, 10u, (unsigned long)b);
    mpz_mul_ui(t, t, k);
    mpz_sub(out, out, t);
    mpz_ui_pow_ui(t, 10u, (unsigned long)c);
    mpz_mul_ui(t, t, j);
    mpz_sub(out, out, t);
    mpz_sub_ui(out, out, 1u);
    mpz_clear(t);
}

/* mode 0: N; mode 1: decimal reversal of N; mode 2: reference formula.
 * Returns significant limb count, -1 on overflow / bad arguments. */
long mr_oracle_build(long long a, long long b, long long c,
                     unsigned long k, unsigned long j, int mode,
                     uint32_t *out, size_t cap)
{
    if (a < 1 || b < 0 || c < 0 || k < 1 || j < 1) return -1;
    mpz_t x;
    mpz_init(x);
    if (mode == 0 || mode == 1) {
        mo_build_n(x, a, b, c, k, j);
        if (mpz_sgn(x) <= 0) { mpz_clear(x); return -1; }
        if (mode == 1) {
            char *s = mpz_get_str(NULL, 10, x);
            size_t L = strlen(s);
            for (size_t i = 0; i < L / 2; i++) {
                char tmp = s[i]; s[i] = s[L - 1 - i]; s[L - 1 - i] = tmp;
            }
            mpz_set_str(x, s, 10);
            void (*freefn)(void *, size_t);
            mp_get_memory_functions(NULL, NULL, &freefn);
            freefn(s, L + 1);
        }
    } else if (mode == 2) {
        int lk = mo_ndigits(k), lj = mo_ndigits(j);
        long long eb = a - b - lk, ec = a - c - lj;
        if (eb < 0 || ec < 0) { mpz_clear(x); return -1; }
        mpz_t t;
        mpz_init(t);
        mpz_ui_pow_ui(x, 10u, (unsigned long)a);
        mpz_ui_pow_ui(t, 10u, (unsigned long)eb);
        mpz_mul_ui(t, t, mo_reverse(k, lk));
        mpz_sub(x, x, t);
        mpz_ui_pow_ui(t, 10u, (unsigned long)ec);
        mpz_mul_ui(t, t, mo_reverse(j, lj));
        mpz_sub(x, x, t);
        mpz_sub_ui(x, x, 1u);
        mpz_clear(t);
        if (mpz_sgn(x) <= 0) { mpz_clear(x); return -1; }
    } else {
        mpz_clear(x);
        return -1;
    }
    long r = mo_to_limbs(x, out, cap);
    mpz_clear(x);
    return r;
}

/* out = base^E mod N (all limb arrays of length n, E of length ne). */
int mr_oracle_powm(const uint32_t *N, size_t n, const uint32_t *base,
                   const uint32_t *E, size_t ne, uint32_t *out)
{
    mpz_t m, b, e, r;
    mpz_inits(m, b, e, r, NULL);
    mo_from_limbs(m, N, n);
    mo_from_limbs(b, base, n);
    mo_from_limbs(e, E, ne);
    int rc = 0;
    if (mpz_sgn(m) <= 0) rc = -1;
    else {
        mpz_powm(r, b, e, m);
        if (mo_to_limbs(r, out, n) < 0) rc = -1;
    }
    mpz_clears(m, b, e, r, NULL);
    return rc;
}

/* Strong probable-prime round to `base`.  Writes base^d mod N into resid
 * (n limbs, may be NULL) and s into *s_out (may be NULL).  A base that is a
 * multiple of N counts as passed (it says nothing about N).
 * Returns 1 = pass, 0 = composite witness, -1 = bad input (N < 5 or even). */
int mr_oracle_sprp(const uint32_t *N, size_t n, unsigned long base,
                   uint32_t *resid, int *s_out)
{
    mpz_t m, m1, d, r, bb;
    mpz_inits(m, m1, d, r, bb, NULL);
    mo_from_limbs(m, N, n);
    int rc;
    if (mpz_cmp_ui(m, 5u) < 0 || mpz_even_p(m)) { rc = -1; goto out; }
    mpz_sub_ui(m1, m, 1u);
    unsigned long s = mpz_scan1(m1, 0);
    mpz_tdiv_q_2exp(d, m1, s);
    if (s_out) *s_out = (int)s;
    mpz_set_ui(bb, base);
    mpz_mod(bb, bb, m);
    mpz_powm(r, bb, d, m);
    if (resid && mo_to_limbs(r, resid, n) < 0) { rc = -1; goto out; }
    /* a base that is a multiple of N tests nothing: the round passes */
    if (mpz_sgn(bb) == 0) { rc = 1; goto out; }
    if (mpz_cmp_ui(r, 1u) == 0 || mpz_cmp(r, m1) == 0) { rc = 1; goto out; }
    rc = 0;
    for (unsigned long i = 1; i < s; i++) {
        mpz_mul(r, r, r);
        mpz_mod(r, r, m);
        if (mpz_cmp(r, m1) == 0) { rc = 1; break; }
        if (mpz_cmp_ui(r, 1u) == 0) break;
    }
out:
    mpz_clears(m, m1, d, r, bb, NULL);
    return rc;
}

/* The reference's production verdict (miller_rabin.c:91): GMP's
 * mpz_probab_prime_p, 0 = composite, 1 = probable prime, 2 = prime. */
int mr_oracle_probab_prime(const uint32_t *N, size_t n, int reps)
{
    mpz_t m;
    mpz_init(m);
    mo_from_limbs(m, N, n);
    int r = mpz_probab_prime_p(m, reps);
    mpz_clear(m);
    return r;
}

const char *mr_oracle_gmp_version(void) { return gmp_version; }
