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
 * words, laid out [n_batch][n_words].  Bit length <= 384 000 (24-bit limbs,
 * transforms of 2^11 .. 2^15 points; 50 000 digits use 2^14, 100 000
 * digits 2^15).  Even moduli and 0, 1 are reported composite and
 * 2, 3 prime without running rounds.  device = HIP device index, or -1
 * for the calling thread's current device.  NULL on error.
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

/* mr_run with options.  flags = 0 behaves as mr_run; MR_RUN_ALL_BASES runs
 * every base on every candidate (no retirement), as the reference runner
 * computes every witness (miller_rabin_runner.cu:305-354), so residues_out
 * holds a^d mod N for every (candidate, base); alive_out is then the AND of
 * the rounds.  pass_out (NULL or [n_batch][n_bases]): 1 where that base's
 * round passed (without MR_RUN_ALL_BASES: where the candidate was still
 * alive after it).  Returns 0, -1 (invalid argument) or -2 (HIP error). */
#define MR_RUN_ALL_BASES 1
int mr_run_ex(mr_ctx_t *ctx, const uint32_t *bases, int n_bases, int flags,
              uint8_t *alive_out, uint8_t *pass_out, uint32_t *residues_out);

/* One strong-probable-prime round in which candidate i uses its own base
 * bases[i] (n_batch entries, each in [2, 2^24)): every (number, base) test
 * of a batch in one launch, the way the reference runner computes all
 * witnesses of all candidates (miller_rabin_runner.cu:305-354), with the
 * numbers repeated in the context once per base.  pass_out[n_batch]: 1 =
 * the round passed.  residues_out (NULL or [n_batch][n_words]): a^d mod N. */
int mr_run_per_slot(mr_ctx_t *ctx, const uint32_t *bases, uint8_t *pass_out,
                    uint32_t *residues_out);

/* ---- Arithmetic on caller-owned device vectors ----------------------
 * The device-pointer methods of the reference's BatchModCtx
 * (batch_mod_ctx.cuh:141-161: modmul_batch, modsq_batch, check_passed,
 * to_residue_batch, from_residue_batch) in this library's working form: the
 * residue of x for candidate c is x R mod N_c with R = 2^(24 L_c) (L_c =
 * that candidate's 24-bit limb count, at most the L of mr_ctx_info), stored
 * as L_c little-endian 24-bit limbs, one per uint64_t, at V + c * ld
 * (ld >= L; entries from L_c up to ld are written as zero).  The reference
 * keeps 16-bit limbs one per uint64_t (Data64) with R = 2^(16 n_limbs), so
 * a buffer of n_batch * n_limbs16 elements holds either form.  The _dev
 * calls are asynchronous on `stream` (a hipStream_t, NULL = null stream)
 * and use the context's scratch: one call at a time per context.
 * Candidates decided without rounds (even N, N < 4) give zero vectors and
 * passed = 0.  Return 0, -1 (invalid argument, or a value not below its
 * modulus) or -2 (HIP error). */

/* d_out = a b R^-1 mod N per candidate (Montgomery product). */
int mr_ctx_modmul_dev(mr_ctx_t *ctx, const uint64_t *d_a, const uint64_t *d_b, uint64_t *d_out,
                      int ld, void *stream);
/* d_out = a^2 R^-1 mod N per candidate. */
int mr_ctx_modsq_dev(mr_ctx_t *ctx, const uint64_t *d_a, uint64_t *d_out, int ld, void *stream);
/* d_passed[c] = 1 if r == R mod N or r == (N-1) R mod N (the round's pass
 * test in working form), else 0. */
int mr_ctx_check_passed_dev(mr_ctx_t *ctx, const uint64_t *d_r, int ld, uint8_t *d_passed,
                            void *stream);
/* Host to host: x (16-bit limbs, one per uint64_t, [n_batch][n_limbs16],
 * each x < N) -> its residue x R mod N in working form, out [n_batch][ld].
 * Synchronous. */
int mr_ctx_to_residue(mr_ctx_t *ctx, const uint64_t *x16, int n_limbs16, uint64_t *out, int ld);
/* Device to host: residues d_x (working form) -> x R^-1 mod N as 16-bit
 * limbs, out16 [n_batch][n_limbs16].  Synchronous. */
int mr_ctx_from_residue_dev(mr_ctx_t *ctx, const uint64_t *d_x, int ld, uint64_t *out16,
                            int n_limbs16);

/* Per-context geometry: words per number, limbs L (24-bit), log2 of the
 * transform length, and s (N - 1 = 2^s d) per candidate (may be NULL). */
int mr_ctx_info(const mr_ctx_t *ctx, int *n_words, int *limbs, int *log_len, int *s_out);

/* Read back the moduli as uint32 words [n_batch][n_words] (for checking
 * the device-built candidates). */
int mr_ctx_moduli(const mr_ctx_t *ctx, uint32_t *out);

void mr_ctx_destroy(mr_ctx_t *ctx);

/* ---- Continuous batching of candidates of mixed sizes ----------------
 *
 * A scheduler tests (a, b, c, k, j, which) candidates against a fixed list
 * of bases, retiring each at its first failing base.  Candidates are sorted
 * into size buckets by transform length (2^11 .. 2^15 points); each bucket
 * owns a pool of slots in HBM (one workgroup each).  Squarings run in chunks
 * of `chunk` steps over every running slot on one stream (one HIP graph
 * launch per chunk), while a second
 * stream ends rounds, starts the next base and builds newly admitted
 * candidates into freed slots, so the GPU never waits for admission.  At
 * most `max_active` slots (0: the GPU's compute-unit count) are in use at
 * once over all buckets; the slot memory of all pools together stays within
 * `hbm_budget` bytes (0: 64 GiB).  Admission is first come, first served.
 *
 * No counterpart in the reference: its runner tests one fixed batch of equal
 * size per call (miller_rabin_runner.cu:281-473) and keeps squaring
 * candidates that already failed (:346-353).
 *
 * Not thread-safe: one host thread drives a scheduler. */
typedef struct mr_sched mr_sched_t;

typedef struct {
    int64_t tag;          /* the caller's id from mr_sched_submit          */
    int32_t verdict;      /* 1 = strong probable prime to every base, 0 = composite */
    int32_t rounds;       /* bases actually run (early retirement)         */
    double t_submit;      /* CLOCK_MONOTONIC seconds                       */
    double t_admit;       /* built into a slot                             */
    double t_done;        /* verdict seen by the host                      */
} mr_result_t;

typedef struct {
    int64_t submitted, completed, pending, active;
    int64_t chunks;       /* squaring-chunk launches                       */
    int64_t admin_batches;
    int64_t squarings;    /* slot-steps issued (upper bound of work done)  */
    int64_t hbm_bytes;    /* slot memory allocated (the pools' slots)      */
    int32_t max_active, n_pools;
    int64_t job_bytes;    /* footprint of the admitted jobs (<= hbm_budget) */
    int64_t peak_job_bytes;
    int64_t reclaimed_pools; /* idle pools released to admit another size  */
} mr_sched_stats_t;

mr_sched_t *mr_sched_create(const uint32_t *bases, int n_bases, int device,
                            int max_active, int chunk, int64_t hbm_budget);
/* Device bytes one job of a candidate with `bits` bits occupies while
 * admitted (its slot: N, NTT(N), NTT(N'), residue, transform scratch), the
 * footprint admission counts against hbm_budget; -1 beyond this build's
 * sizes. */
int64_t mr_sched_job_bytes(int64_t bits);
/* Queue n candidates (validated first; -1 and nothing queued if any cannot
 * be formed or is too large). */
int mr_sched_submit(mr_sched_t *s, const mr_candidate_t *cands, const int64_t *tags, int n);
/* Queue the survivors of one trial-division call in place: d_list is a
 * device list as trial_div_survivors writes it ((k << 16) | j per entry),
 * n how many of its entries to test (e.g. the count trial_div_survivors
 * returned), each as the candidate (a, b, c, k, j, which) with tag
 * tag_base + i.  The (k, j) values stay on the device: a job's number is
 * built from its list entry when a slot takes it.  The list must stay
 * valid and unchanged until every job of the call is decided.  When
 * max(b, c) > a - 7 the decimal length may vary between survivors; the list
 * is then read to the host and queued as mr_sched_submit would.  Returns
 * 0, -1 for invalid arguments or candidates (nothing queued), -2 on a HIP
 * error.  No counterpart in the reference, whose worker sends survivors to
 * its master over MPI (worker.c:45-54, master.c:518-531). */
int mr_sched_submit_survivors(mr_sched_t *s, long long a, long long b, long long c,
                              const uint32_t *d_list, uint32_t n, int which, int64_t tag_base);
/* Advance the scheduler.  block = 0: launch what can be launched and
 * return; block = 1: also wait for the next chunk or admission batch to
 * finish.  Returns the number of candidates not yet decided (>= 0), or -2
 * on a HIP error. */
int mr_sched_pump(mr_sched_t *s, int block);
/* Move up to cap finished results to out; returns how many. */
int mr_sched_results(mr_sched_t *s, mr_result_t *out, int cap);
int mr_sched_stats(const mr_sched_t *s, mr_sched_stats_t *out);
void mr_sched_destroy(mr_sched_t *s);

/* Per-bucket choices, replacing the reference's configure-time parameters
 * (params.cmake:6-64, miller_rabin_gpu/config.h.in:11-95) with values a
 * tuner measured on the device.  mr_sched_set_bucket: candidates of
 * transform length 2^log_len (11..15) run in chunks of `chunk` squaring
 * steps in a pool of `slots` slots (0: the scheduler-wide chunk and
 * max_active); when max_active was 0 at creation, the active cap rises to
 * the largest bucket's slots.  Must precede the bucket's first admission
 * (its pool's graph holds the chunk): returns 0, or -1 (invalid values, or
 * the pool exists already).
 *
 * A tuning table is a JSON object whose "buckets" array holds objects of
 * integer fields "log_len", "chunk" and "slots_per_cu" (slots per compute
 * unit; other fields are ignored, nested values are not allowed inside a
 * bucket), as prime-finder_amd/tune_mr.py writes it.  mr_tuning_read fills
 * up to cap entries and returns the number of buckets in the file, or -1
 * (unreadable or malformed); no GPU call.  mr_sched_load_tuning applies a
 * table to s (slots = slots_per_cu x the device's compute units) and returns
 * the number of buckets set, or -1 (nothing set).
 * No counterpart in the reference. */
int mr_sched_set_bucket(mr_sched_t *s, int log_len, int chunk, int slots);
int mr_tuning_read(const char *path, int *log_len, int *chunk, int *slots_per_cu, int cap);
int mr_sched_load_tuning(mr_sched_t *s, const char *path);

/* Checkpoint and resume of a serving scheduler.  The reference's master
 * dumps its pending Phase-2 queue as "a b c k j" lines (master.c:279-369,
 * checkpoint_save_phase2 / checkpoint_load_phase2) and restarts every test
 * from scratch; here the candidates already admitted keep their progress.
 *
 * mr_sched_checkpoint waits for the work in flight (launching nothing new),
 * then writes every undecided candidate to the text file at path: its tag,
 * candidate, completed rounds, and, once admitted, the index of the base in
 * progress, the next exponent bit and the residue (Montgomery form, 24-bit
 * limbs).  Results already decided stay in mr_sched_results.  The scheduler
 * can go on pumping afterwards.  Returns the number of candidates written,
 * or -1 (file error), -2 (HIP error).
 *
 * mr_sched_restore queues the candidates of such a file on s, which must
 * have been created with the same list of bases; admitted ones resume at
 * their saved exponent bit.  Returns the number restored, or -1 (unreadable
 * or malformed file, other bases, invalid candidate; nothing queued), -2
 * (HIP error or out of memory while queueing). */
int mr_sched_checkpoint(mr_sched_t *s, const char *path);
int mr_sched_restore(mr_sched_t *s, const char *path);

/* One call over a whole list: verdict[i] (1 / 0) and rounds[i] (nullable)
 * for cands[i].  Returns 0, -1 (invalid candidate) or -2 (HIP error). */
int mr_test_batch(const mr_candidate_t *cands, int n, const uint32_t *bases, int n_bases,
                  int device, uint8_t *verdict, int32_t *rounds);

/* Smallest transform length (log2, 11..15) for contexts created afterwards;
 * a context uses the smallest length >= this that holds its products.  The
 * size-bucket tuner and the tests use it; process-wide, not thread-safe.
 * No counterpart in the reference (its length is fixed at configure time,
 * params.cmake). */
int mr_set_min_log_len(int log_len);

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
