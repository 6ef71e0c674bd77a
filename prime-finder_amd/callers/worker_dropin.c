/*
 * worker_dropin -- a C caller of the library's drop-in entry points making
 * the calls of the reference's worker (worker.c:37-42 trial_div_gpu, the
 * survivor walk of worker.c:45-54, miller_rabin_test of worker.c:84) and of
 * main.c:92-94 (cuda_get_device_count / cuda_set_device), compiled as C and
 * linked against libpfamd.so exactly as prime_hunter would be.  Used by the
 * GPU tests.
 *
 *   worker_dropin <primes.bin> a b c [n_mr]
 *
 * Prints "bits" and the KJ_WORDS words of the bitmap (hex), "survivors n",
 * "timing upload kernel download" (seconds), then "mr k j code" for the
 * first n_mr survivors (default 0).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "miller_rabin.h"
#include "prime_list.h"
#include "trial_div.h"

int main(int argc, char **argv)
{
    if (argc < 5) {
        fprintf(stderr, "usage: %s <primes.bin> a b c [n_mr]\n", argv[0]);
        return 2;
    }
    const long long a = atoll(argv[2]), b = atoll(argv[3]), c = atoll(argv[4]);
    const int n_mr = argc > 5 ? atoi(argv[5]) : 0;
    if (cuda_get_device_count() < 1 || cuda_set_device(0) != 0) {
        fprintf(stderr, "worker_dropin: no GPU\n");
        return 1;
    }
    prime_list_t pl;
    if (prime_list_load(&pl, argv[1]) != 0) {
        fprintf(stderr, "worker_dropin: cannot load %s\n", argv[1]);
        return 1;
    }
    uint32_t bits[KJ_WORDS];
    memset(bits, 0, sizeof(bits));
    gpu_timing_t gtiming = {0};
    if (trial_div_gpu(&pl, a, b, c, bits, &gtiming) != 0) {
        fprintf(stderr, "worker_dropin: GPU trial division failed\n");
        prime_list_free(&pl);
        return 1;
    }
    printf("bits");
    for (int i = 0; i < KJ_WORDS; i++) printf(" %08x", bits[i]);
    printf("\n");
    int ks[KJ_BITS], js[KJ_BITS], n = 0;
    for (int k = 1; k <= 100; k++) {
        for (int j = 1; j <= 100; j++) {
            const int bit = (k - 1) * 100 + (j - 1);
            if (!(bits[bit >> 5] & (1u << (bit & 31)))) {
                ks[n] = k;
                js[n] = j;
                n++;
            }
        }
    }
    printf("survivors %d\n", n);
    printf("timing %.6f %.6f %.6f\n", gtiming.upload, gtiming.kernel, gtiming.download);
    for (int i = 0; i < n && i < n_mr; i++) {
        mr_timing_t mrt = {0};
        const int code = miller_rabin_test(a, b, c, ks[i], js[i], &mrt);
        printf("mr %d %d %d\n", ks[i], js[i], code);
    }
    prime_list_free(&pl);
    return 0;
}
