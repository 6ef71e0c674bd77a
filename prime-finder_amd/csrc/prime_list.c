/*
 * prime_list.c -- host-side prime tables for Phase-1 trial division.
 *
 * prime_list_load reads the reference's gap-encoded file format
 * (reference prime_list.c:5-55, written by gen_primes.c:91-127):
 *   [u32 count][u32 first prime][u8 gap]...   a zero gap byte is followed by
 *   the gap as a little-endian u16.
 * prime_list_sieve builds the same list in memory with an odd-only,
 * segment-free sieve of Eratosthenes (limit <= 2^32 - 1).
 */
#include "../../include/prime_list.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int prime_list_load(prime_list_t *pl, const char *path)
{
    if (!pl || !path) return -1;
    pl->primes = NULL;
    pl->count = 0;
    FILE *f = fopen(path, "rb");
    if (!f) { perror(path); return -1; }
    uint32_t count = 0;
    if (fread(&count, sizeof count, 1, f) != 1) { fclose(f); return -1; }
    uint32_t *v = (uint32_t *)malloc((size_t)(count ? count : 1) * sizeof(uint32_t));
    if (!v) { fclose(f); return -1; }
    if (count > 0 && fread(&v[0], sizeof(uint32_t), 1, f) != 1) goto bad;
    for (uint32_t i = 1; i < count; i++) {
        int g = fgetc(f);
        if (g == EOF) goto bad;
        uint32_t gap = (uint32_t)g;
        if (gap == 0) {
            int lo = fgetc(f), hi = fgetc(f);
            if (lo == EOF || hi == EOF) goto bad;
            gap = (uint32_t)lo | ((uint32_t)hi << 8);
        }
        v[i] = v[i - 1] + gap;
    }
    fclose(f);
    pl->primes = v;
    pl->count = count;
    return 0;
bad:
    free(v);
    fclose(f);
    return -1;
}

void prime_list_free(prime_list_t *pl)
{
    if (!pl) return;
    free(pl->primes);
    pl->primes = NULL;
    pl->count = 0;
}

int prime_list_sieve(prime_list_t *pl, uint32_t lo, uint32_t limit)
{
    if (!pl) return -1;
    pl->primes = NULL;
    pl->count = 0;
    if (limit <= 2u || lo >= limit) {
        pl->primes = (uint32_t *)malloc(sizeof(uint32_t));
        return pl->primes ? 0 : -1;
    }
    /* composite[i] describes the odd number 2 i + 1 */
    const uint64_t n_odd = ((uint64_t)limit + 1u) / 2u;
    uint8_t *composite = (uint8_t *)calloc((size_t)n_odd, 1);
    if (!composite) return -1;
    composite[0] = 1; /* 1 is not prime */
    for (uint64_t i = 1; (2 * i + 1) * (2 * i + 1) < limit; i++) {
        if (composite[i]) continue;
        const uint64_t p = 2 * i + 1;
        for (uint64_t m = p * p; m < limit; m += 2 * p) composite[m / 2] = 1;
    }
    uint64_t cnt = (lo <= 2u) ? 1u : 0u;
    for (uint64_t i = 1; i < n_odd; i++)
        if (!composite[i] && 2 * i + 1 >= lo && 2 * i + 1 < limit) cnt++;
    uint32_t *v = (uint32_t *)malloc((size_t)(cnt ? cnt : 1) * sizeof(uint32_t));
    if (!v) { free(composite); return -1; }
    uint64_t w = 0;
    if (lo <= 2u) v[w++] = 2u;
    for (uint64_t i = 1; i < n_odd; i++) {
        const uint64_t x = 2 * i + 1;
        if (!composite[i] && x >= lo && x < limit) v[w++] = (uint32_t)x;
    }
    free(composite);
    pl->primes = v;
    pl->count = (uint32_t)cnt;
    return 0;
}
