/*
 * prime_list.h -- the prime table handed to Phase-1 trial division.
 *
 * Same type and loader contract as the reference's prime_list.h:4-15
 * (struct layout is part of the trial_div_gpu ABI), plus one generator.
 */
#pragma once
#include <stdint.h>

typedef struct {
    uint32_t *primes;
    uint32_t  count;
} prime_list_t;

#ifdef __cplusplus
extern "C" {
#endif

/* Load the reference's gap-encoded file: [u32 count][u32 p0][u8 gap ...],
 * a zero gap byte being followed by a little-endian u16 gap
 * (reference prime_list.c:5-55).  Returns 0 on success, -1 on error. */
int  prime_list_load(prime_list_t *pl, const char *path);

/* Release what prime_list_load / prime_list_sieve allocated. */
void prime_list_free(prime_list_t *pl);

/* Extension (no counterpart in the reference, which ships gen_primes.c as a
 * separate tool): fill pl with every prime p with lo <= p < limit, ascending.
 * lo = 3 reproduces gen_primes' file contents (2 excluded,
 * gen_primes.c:119-122); lo = 2 gives the full list.  Returns 0 / -1. */
int  prime_list_sieve(prime_list_t *pl, uint32_t lo, uint32_t limit);

#ifdef __cplusplus
}
#endif
