/*
 * mr_gpu.h -- batched Miller-Rabin on an MI355X (gfx950), C ABI.
 *
 * Replaces the reference's C++-only GPU runner surface
 * (miller_rabin_gpu/miller_rabin_runner.cuh:28-49 gpu_miller_rabin_s1 /
 * gpu_miller_rabin, miller_rabin_gpu/batch_mod_ctx.cuh:52-179 BatchModCtx)
 * with entry points a C caller such as worker.c can bind.
 *
 * A context holds n_batch odd moduli N (all of one size class) resident in
 * HBM together with their Montgomery constants (R = 2^(24 L), N' = -N^-1 mod
 * R, R mod N, (N-1)R mod N) and the transforms of N and N'.  mr_run then
 * performs strong-probable-prime rounds: for each base a, r = a^d mod N with
 * N - 1 = 2^s d, d odd; the round passes iff r == 1 or r^(2^i) == N - 1 for
 * some 0 <= i < s (miller_rabin_runner.cu:281-473 semantics).  A candidate
 * is "alive" while every base so far passed; a failed candidate is retired
 * and not tested with later bases.  This is per-base Miller-Rabin; the
 * reference CPU path (miller_rabin.c:91, GMP mpz_probab_prime_p) is
 * Baillie-PSW instead.
 *
 * Big integers cross the ABI as little-endian uint32 word arrays.
 * Return codes: 0 success, -1 invalid argument, -2 HIP error (message on
 * stderr).  No C++ exception crosses this interface.  A context is used by
 * one host thread at a time; it owns one HIP stream.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mr_ctx mr_ctx_t;

/* Build a context from n_batch moduli, each n_words little-endian uint32
 * words, laid out [n_batch][n_words].  Bit length <= 196 000 in this build
 * (transforms up to 2^14).  Even moduli and 0, 1 are reported composite and
 * 2, 3 prime without running rounds.  device = HIP device index.  NULL on
 * error.
 * Replaces BatchModCtx(N_all, n_limbs, n_batch, device)
 * (batch_mod_ctx.cuh:132-133). */
mr_ctx_t *mr_ctx_create(const uint32_t *N_words, int n_words, int n_batch, int device);

/* Same, from the reference's limb format: 16-bit limbs stored one per
 * uint64 (Data64), [n_batch][n_limbs16] (batch_mod_ctx.cuh:128-133,
 * candidate.cuh:62-70). */
mr_ctx_t *mr_ctx_create_limbs16(const uint64_t *N_all, int n_limbs16, int n_batch, int device);

/* Candidates of the search form, built on the device from their decimal
 * structure: N = 10^a - k 10^b - j 10^c - 1 (which = 0), or rev(N) as the
 * decimal-string reversal of N (which = 1), or the reference's closed rev
 * formula 10^a - rev(k) 10^(a-b-lk) - rev(j) 10^(a-c-lj) - 1 (which = 2,
 * miller_rabin.c:54-74; equal to which = 1 whenever the k and j digit
 * blocks of N do not overlap and N keeps a digits). */
typedef struct {
    int64_t a, b, c;
    uint32_t k, j;
    int32_t which;
} mr_candidate_t;

mr_ctx_t *mr_ctx_create_sparse(const mr_candidate_t *cands, int n_batch, int device);

/* Strong-probable-prime rounds to each base in order (every base in
 * [2, 2^24)).  alive_out[n_batch]: 1 = passed every base, 0 = composite.
 * residues_out (NULL or [n_batch][n_bases][n_words] words) receives
 * a^d mod N for every base that was run; rows of bases skipped after an
 * earlier failure are zero-filled.  Synchronous. */
int mr_run(mr_ctx_t *ctx, const uint32_t *bases, int n_bases,
           uint8_t *alive_out, uint32_t *residues_out);

/* Per-context geometry: words per number, limbs L (24-bit), log2 of the
 * transform length, and s (N - 1 = 2^s d) per candidate (may be NULL). */
int mr_ctx_info(const mr_ctx_t *ctx, int *n_words, int *limbs, int *log_len, int *s_out);

/* Read back the moduli as uint32 words [n_batch][n_words] (for checking
 * the device-built candidates). */
int mr_ctx_moduli(const mr_ctx_t *ctx, uint32_t *out);

void mr_ctx_destroy(mr_ctx_t *ctx);

/* Benchmark helper (no counterpart in the reference, whose runner prints
 * per-phase timings itself, miller_rabin_runner.cu:205-275): starts a round
 * with `base` on every candidate and times its first n_bits squaring steps
 * on the context's stream; *ms = elapsed milliseconds. */
int mr_time_squarings(mr_ctx_t *ctx, uint32_t base, int n_bits, float *ms);

/* Arithmetic self-check entry points (used by the parity tests):
 * out = a * b (full product), a * b * R^-1 mod N, or a * s mod N for a
 * small s, computed by the same workgroup routines as mr_run on every
 * candidate of the context.  a, b: [n_batch][n_words], a, b < N for the
 * modular forms; out: [n_batch][2 * n_words] for the product, [n_batch]
 * [n_words] otherwise. */
int mr_debug_product(mr_ctx_t *ctx, const uint32_t *a, const uint32_t *b, uint32_t *out);
int mr_debug_montmul(mr_ctx_t *ctx, const uint32_t *a, const uint32_t *b, uint32_t *out);
int mr_debug_mulsmall(mr_ctx_t *ctx, const uint32_t *a, uint32_t s, uint32_t *out);

#ifdef __cplusplus
}
#endif
