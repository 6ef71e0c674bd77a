/*
 * miller_rabin.h -- the reference's Phase-2 entry point (miller_rabin.h:1-25),
 * served by the gfx950 Miller-Rabin path of libpfamd.so instead of GMP.
 *
 * worker.c (worker.c:83-84) includes this header and calls
 * miller_rabin_test unchanged.
 */
#pragma once

typedef struct {
    double build_n;    /* building N (device), seconds                   */
    double mr_n;       /* rounds on N, seconds                           */
    double build_revn; /* building rev(N), seconds (0 if N composite)    */
    double mr_revn;    /* rounds on rev(N), seconds                      */
} mr_timing_t;

#ifdef __cplusplus
extern "C" {
#endif

/*
 * N = 10^a - k 10^b - j 10^c - 1 and, when N passes, the reference's rev
 * formula 10^a - rev(k) 10^(a-b-lk) - rev(j) 10^(a-c-lj) - 1
 * (miller_rabin.c:32-74), tested on the current HIP device
 * (cuda_set_device).  Replaces miller_rabin.c:76-118.
 *
 * Verdict: strong probable prime to bases 2, 3, 5 and 7 (exact below
 * 3 215 031 751), where the reference runs GMP's mpz_probab_prime_p(., 5)
 * (Baillie-PSW); the two agree on every input of known answer.
 *
 * Returns 0 = N composite (rev not tested), 1 = N probable prime and rev(N)
 * composite, 2 = both probable primes; -1 when the candidate cannot be
 * formed (k, j < 1, N <= 0, a negative rev exponent, a >= 2 000 000; the
 * reference has no defined result for these) or on a HIP error, with a
 * message on stderr.  timing (nullable) receives the four durations.
 */
int miller_rabin_test(long long a, long long b, long long c, int k, int j,
                      mr_timing_t *timing);

#ifdef __cplusplus
}
#endif
