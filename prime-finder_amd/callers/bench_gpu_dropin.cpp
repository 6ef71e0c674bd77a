// bench_gpu_dropin -- a C++ caller of trial_div_launch with the reference's
// C++ linkage, making the calls of tests/bench_gpu.cu:25-27,46,56: the prime
// list and a KJ_WORDS bitmap in device memory, the bitmap cleared, one
// launch per run, the device synchronised; here the device memory comes
// from the library's pf_runtime helpers instead of cudaMalloc.  Used by the
// GPU tests.
//
//   bench_gpu_dropin <primes.bin> a b c [runs]
//
// Prints "bits" and the KJ_WORDS words of the last run's bitmap (hex), then
// "ms_per_run t".
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>

extern "C" {
#include "prime_list.h"
}
#include "pf_runtime.h"
#include "trial_div.h"

// trial_div_launch is a C++ function (reference trial_div.cu:127-133).
void trial_div_launch(const uint32_t *d_primes, uint32_t n_primes, uint64_t a, uint64_t b,
                      uint64_t c, uint32_t *d_bits, int block_size);

int main(int argc, char **argv)
{
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s <primes.bin> a b c [runs]\n", argv[0]);
        return 2;
    }
    const long long a = std::atoll(argv[2]), b = std::atoll(argv[3]), c = std::atoll(argv[4]);
    const int runs = argc > 5 ? std::atoi(argv[5]) : 3;
    prime_list_t pl;
    if (prime_list_load(&pl, argv[1]) != 0) {
        std::fprintf(stderr, "bench_gpu_dropin: cannot load %s\n", argv[1]);
        return 1;
    }
    uint32_t host_bits[KJ_WORDS];
    std::memset(host_bits, 0, sizeof(host_bits));
    const size_t pb = (size_t)pl.count * sizeof(uint32_t), bb = KJ_WORDS * sizeof(uint32_t);
    uint32_t *d_primes = (uint32_t *)pf_malloc(pb ? pb : 4);
    uint32_t *d_bits = (uint32_t *)pf_malloc(bb);
    if (!d_primes || !d_bits || pf_memcpy_htod(d_primes, pl.primes, pb) != 0) {
        std::fprintf(stderr, "bench_gpu_dropin: device memory\n");
        return 1;
    }
    const auto t0 = std::chrono::steady_clock::now();
    for (int r = 0; r < runs; r++) {
        pf_memcpy_htod(d_bits, host_bits, bb);
        trial_div_launch(d_primes, pl.count, (uint64_t)a, (uint64_t)b, (uint64_t)c, d_bits, 256);
    }
    if (pf_sync() != 0 || trial_div_check_error() != 0) {
        std::fprintf(stderr, "bench_gpu_dropin: launch failed\n");
        return 1;
    }
    const double ms = std::chrono::duration<double, std::milli>(
                          std::chrono::steady_clock::now() - t0).count() / runs;
    uint32_t out[KJ_WORDS];
    pf_memcpy_dtoh(out, d_bits, bb);
    std::printf("bits");
    for (int i = 0; i < KJ_WORDS; i++) std::printf(" %08x", out[i]);
    std::printf("\nms_per_run %.4f\n", ms);
    pf_free(d_primes);
    pf_free(d_bits);
    prime_list_free(&pl);
    return 0;
}
