/*
 * trial_div.h -- Phase-1 trial division on an MI355X (gfx950), C ABI.
 *
 * Drop-in for the reference's trial_div.h: the first four declarations keep
 * the reference's names, argument meaning, return codes and the
 * OR-into-caller-bits contract, so worker.c (worker.c:37-42), main.c
 * (main.c:92-94) and tests/test_gpu.c call them unchanged.  The cuda_* names
 * are kept for ABI compatibility; they are implemented over HIP.
 *
 * Bitmap contract: bit (k-1)*J + (j-1), LSB-first in uint32 words,
 * 1 = (k,j) eliminated because some prime p of the list divides
 *     N = 10^a - k 10^b - j 10^c - 1, or divides
 *     rev(N) = 10^a - rev(k) 10^(a-b-lk) - rev(j) 10^(a-c-lj) - 1
 *     (rev(x) = decimal reversal of x, lx = its digit count; a rev term with
 *     a negative exponent eliminates nothing).  Identical to the reference's
 *     CPU sieve trial_div_cpu.c:32-111 on every input.
 */
#pragma once
#include "prime_list.h"
#include <stddef.h>
#include <stdint.h>

/* The reference's fixed grid (sieve_table.h:15-17): K = J = 100. The
 * definitions are token-identical to the reference's so that including both
 * headers in one translation unit is a benign redefinition. */
#define BITS_PER_WORD   32
#define KJ_BITS         10000
#define KJ_WORDS        ((KJ_BITS + BITS_PER_WORD - 1) / BITS_PER_WORD)

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double upload;    /* host->device copies (bits; primes when not cached), s */
    double kernel;    /* all trial-division kernels, s                        */
    double download;  /* device->host copy of the bits, s                     */
} gpu_timing_t;       /* reference trial_div.h:10-14 */

/* Replaces reference trial_div.h:26-29 / trial_div.cu:155-224.
 * K = J = 100; sieve_bits holds KJ_WORDS words, in/out (OR-accumulate).
 * Synchronous.  Returns 0, or -1 after printing
 * "trial_div_gpu: HIP error ..." to stderr (bits then unchanged).
 * The device copy of pl->primes is cached across calls and refreshed when
 * the pointer, the count or the contents change (compared word for word
 * with a host copy of the last upload).
 * Every prime must satisfy 2 <= p < 2^31. */
int trial_div_gpu(const prime_list_t *pl,
                  long long a, long long b, long long c,
                  uint32_t *sieve_bits,
                  gpu_timing_t *timing);

/* Replaces reference trial_div.h:31 (trial_div.cu:137-140): number of HIP
 * devices, 0 on error. */
int cuda_get_device_count(void);

/* Replaces reference trial_div.h:32 (trial_div.cu:142-145): 0 / -1. */
int cuda_set_device(int index);

/* Extension of trial_div_gpu to a K x J grid (1 <= K, J <= 9999), for the
 * 4096 x 4096 grids the reference's fixed 100 x 100 cannot express
 * (sieve_table.h:15-17).  bits holds n_words >= ceil(K*J/32) words, in/out.
 * Same semantics, caching and error convention as trial_div_gpu. */
int trial_div_gpu_kj(const prime_list_t *pl,
                     long long a, long long b, long long c,
                     uint32_t K, uint32_t J,
                     uint32_t *bits, size_t n_words,
                     gpu_timing_t *timing);

/* Device-resident form, asynchronous on `stream` (a hipStream_t, NULL =
 * the null stream): d_primes (n_primes words) and d_bits (ceil(K*J/32)
 * words, OR-accumulated) are device pointers.  This is the generalised
 * trial_div_launch (reference trial_div.cu:127-133) that the benchmark and
 * batch callers use; it returns -1 only for invalid arguments or launch
 * failure.  Input-data errors (a prime outside [2, 2^31)) are latched on the
 * device and reported by trial_div_check_error().  Calls on one device share
 * scratch memory: issue them from one host thread, on one stream. */
int trial_div_launch_kj(const uint32_t *d_primes, uint32_t n_primes,
                        long long a, long long b, long long c,
                        uint32_t K, uint32_t J,
                        uint32_t *d_bits, void *stream);

/* trial_div_launch_kj that also records two hipEvent_t on `stream` around
 * its dominant kernel (the lattice search of the primes from
 * trial_div_lattice_min() up, or the k-loop kernel of every prime from 2^16
 * when the j digit groups do not merge), so a benchmark can time it on the
 * stream it runs on; the kernels of the smaller primes run beside it on a
 * second stream.  Either event may
 * be NULL.  No counterpart in the reference (its bench times the whole
 * launch, tests/bench_gpu.cu:53-62). */
int trial_div_launch_kj_ev(const uint32_t *d_primes, uint32_t n_primes,
                           long long a, long long b, long long c,
                           uint32_t K, uint32_t J,
                           uint32_t *d_bits, void *stream,
                           void *ev_lane_start, void *ev_lane_stop);

/* Synchronises the current device, then returns and clears the error latched
 * by earlier launches: 0 = none, else a bit mask: 1 = a prime outside
 * [2, 2^31), 2 = more than 65536 list entries below 2^16, 4 = more than
 * 1 100 000 entries from 2^16 to the lattice range (2^17) (both: a
 * duplicated list). */
int trial_div_check_error(void);

/* trial_div_launch_kj that allocates nothing: no hipMalloc, hipFree, stream
 * or event creation, on this call or any other.  No counterpart in the
 * reference; it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  Its
 * scratch (small-prime list and setup table, control words, rev(j) decode
 * table, row-kernel work items, mid-range prime list, class masks of the
 * small primes' word fills: 11.75 MB) is one
 * static device array that
 * each GPU the library runs on holds until the process ends; every kernel
 * runs in order on `stream`.  Limits: K, J in 1..9999 (-1 and a message
 * otherwise), at most 65536 list entries below 2^16 and 1 100 000 from 2^16
 * to 2^17 (latched, reported by trial_div_check_error() as 2 and 4).  All
 * calls on one GPU share the array:
 * calls that overlap in time on one GPU (other streams, other host threads)
 * race and must be serialised by the caller.  Same arguments, results and
 * return codes as trial_div_launch_kj. */
int trial_div_launch_kj_noalloc(const uint32_t *d_primes, uint32_t n_primes,
                                long long a, long long b, long long c,
                                uint32_t K, uint32_t J,
                                uint32_t *d_bits, void *stream);

/* The lower bound of the primes whose hits the lattice kernel enumerates
 * directly (when the j digit groups merge, a - c >= D_j); primes from 2^16
 * below it take the split k loop.  No counterpart in the reference. */
uint32_t trial_div_lattice_min(void);

/* Algorithmic work of one trial_div_launch_kj call, counted on the host
 * (no GPU call) with `threads` threads, for rating the kernels: out[] is
 * indexed by the TD_WORK_* fields below.  The lattice counts come from the
 * lattice kernel's own per-prime routine run with counting hooks.  Returns 0,
 * or -1 for invalid arguments.  No counterpart in the reference. */
enum {
    TD_WORK_PRIME = 0,   /* primes on the lattice path                      */
    TD_WORK_MODMUL,      /* their 32-bit Montgomery multiplications         */
    TD_WORK_QUERY,       /* first-hit queries                               */
    TD_WORK_LEVEL,       /* continued-fraction levels of those queries      */
    TD_WORK_HIT,         /* hits found (before the k <= K, j <= J filter)   */
    TD_WORK_STEP,        /* direct steps of the one-digit rev class         */
    TD_WORK_KSTEP,       /* prime x k steps of the k-loop kernels (>= 2^16) */
    TD_WORK_GAPSTEP,     /* return-time steps from hit to hit               */
    TD_WORK_GAPCF,       /* division steps computing those return times     */
    TD_WORK_MODMUL_LAZY, /* Montgomery multiplications without the final    */
                         /* subtraction (primes below 2^30)                 */
    TD_WORK_FIELDS
};
int trial_div_work(const uint32_t *primes, uint32_t n_primes,
                   long long a, long long b, long long c,
                   uint32_t K, uint32_t J, uint32_t threads,
                   uint64_t out[TD_WORK_FIELDS]);

/* Survivor extraction of worker.c:45-54 on the device: the pairs (k, j) of
 * a K x J bitmap (device pointer, layout as trial_div_gpu_kj) whose bit is
 * clear, in lexicographic order, written to d_out (device, cap entries) as
 * (k << 16) | j.  *count (host) receives the number of survivors, which may
 * exceed cap (then only cap are written).  Synchronous on `stream`.
 * Returns 0, or -1 (message on stderr).  No counterpart in the reference,
 * which walks the bitmap on the host. */
int trial_div_survivors(const uint32_t *d_bits, uint32_t K, uint32_t J,
                        uint32_t *d_out, uint32_t cap, uint32_t *count, void *stream);

#ifdef __cplusplus
} /* extern "C" */

/* The reference's configurable launcher keeps its C++ linkage
 * (trial_div.cu:127-133, declared by its caller tests/bench_gpu.cu:25-27):
 * K = J = 100 on device buffers, asynchronous on the null stream.
 * block_size is accepted for signature compatibility; the launch shapes are
 * fixed by this implementation. */
void trial_div_launch(const uint32_t *d_primes, uint32_t n_primes,
                      uint64_t a, uint64_t b, uint64_t c,
                      uint32_t *d_bits, int block_size);
#endif
