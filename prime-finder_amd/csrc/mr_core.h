// mr_core.h -- workgroup-level big-integer arithmetic for gfx950.
//
// One workgroup of MR_THREADS (1024) threads owns one big integer operation
// at a time.  Numbers are little-endian arrays of 24-bit limbs.  A product
// is computed with two number-theoretic transforms (one per 31-bit prime,
// P0 = 63*2^25 + 1 and P1 = 127*2^24 + 1) of length M = 2^LOGM held in LDS,
// recombined exactly by CRT (P0*P1 > 2^61.9 > L * (2^24 - 1)^2 for every
// L <= 16002 limbs) and normalised to limbs with a two-level carry:
// a per-thread ripple over a contiguous run of C = M / 1024 coefficients,
// then a workgroup-wide 1-bit carry-lookahead scan built from wave ballots.
//
// Register layout of a limb vector: thread t holds positions
// [t*C, t*C + C) in v[0..C).  Forward transforms are decimation-in-frequency
// (natural order in, bit-reversed out), inverse transforms
// decimation-in-time (bit-reversed in, natural out), so no permutation pass
// is ever needed and pointwise products work on bit-reversed data.
//
// Modular multiplications: twiddles use Shoup's precomputed quotient
// (3 multiplies); data-by-data products use Montgomery with 2^32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mr {

constexpr int      THREADS   = 1024;
constexpr int      WAVES     = THREADS / 64;
constexpr uint32_t LIMB_BITS = 24;
constexpr uint32_t LIMB_MASK = (1u << LIMB_BITS) - 1u;

constexpr uint32_t P0 = 2113929217u;   // 63 * 2^25 + 1, primitive root 5
constexpr uint32_t P1 = 2130706433u;   // 127 * 2^24 + 1, primitive root 3
constexpr uint32_t G0 = 5u, G1 = 3u;
constexpr uint32_t P0_INV_MOD_P1 = 127u;
constexpr int      MAX_LIMBS = 16002;  // CRT exactness bound for 24-bit limbs

__host__ __device__ constexpr uint32_t prime(int q) { return q == 0 ? P0 : P1; }

// -p^{-1} mod 2^32
__host__ __device__ constexpr uint32_t neg_inv32(uint32_t p)
{
    uint32_t x = p;
    for (int i = 0; i < 5; i++) x *= 2u - p * x;
    return 0u - x;
}

// ---------------------------------------------------------------- mod p ops
__device__ __forceinline__ uint32_t add_mod(uint32_t a, uint32_t b, uint32_t p)
{
    uint32_t s = a + b;          // < 2^32 since a, b < p < 2^31
    return min(s, s - p);
}

__device__ __forceinline__ uint32_t sub_mod(uint32_t a, uint32_t b, uint32_t p)
{
    uint32_t d = a - b;
    return min(d, d + p);
}

// x * w mod p for any x < 2^32, wq = floor(w * 2^32 / p), w < p.
__device__ __forceinline__ uint32_t shoup(uint32_t x, uint32_t w, uint32_t wq, uint32_t p)
{
    uint32_t q = __umulhi(x, wq);
    uint32_t r = x * w - q * p;
    return min(r, r - p);
}

// a * b * 2^-32 mod p (Montgomery), a, b < p.
__device__ __forceinline__ uint32_t mont(uint32_t a, uint32_t b, uint32_t p, uint32_t np)
{
    uint64_t t = (uint64_t)a * b;
    uint32_t lo = (uint32_t)t, hi = (uint32_t)(t >> 32);
    uint32_t m = lo * np;
    uint32_t r = hi + __umulhi(m, p) + (lo != 0u ? 1u : 0u);
    return min(r, r - p);
}

// ------------------------------------------------------------- tables
// Per (prime, M): tw[i] = w^i and itw[i] = w^-i for i < M/2 with their Shoup
// quotients, w a primitive M-th root of unity.  Layout of one table block:
// [tw | twq | itw | itwq], each M/2 words; prime 1 follows prime 0.
struct Twiddles {
    const uint32_t *base;   // 2 primes * 4 * (M/2) words
    int half;               // M/2
    __device__ __forceinline__ const uint32_t *tw(int q) const { return base + (size_t)q * 4 * half; }
    __device__ __forceinline__ const uint32_t *twq(int q) const { return tw(q) + half; }
    __device__ __forceinline__ const uint32_t *itw(int q) const { return tw(q) + 2 * half; }
    __device__ __forceinline__ const uint32_t *itwq(int q) const { return tw(q) + 3 * half; }
};

// ------------------------------------------------------------- transforms
// Forward DIF, natural -> bit-reversed, on NP buffers X + q*M (prime q0+q)
// in the same passes (one barrier per radix-4 pass for all of them).
template <int LOGM, int NP>
__device__ void ntt_fwd(uint32_t *X, const Twiddles &T, int q0)
{
    constexpr int M = 1 << LOGM;
    constexpr int GROUPS = M / 4;
    const int tid = threadIdx.x;
    int s = 0;
    for (; s + 1 < LOGM; s += 2) {
        const int h2 = M >> (s + 2);
#pragma unroll
        for (int q = 0; q < NP; q++) {
            uint32_t *Y = X + q * M;
            const uint32_t p = prime(q0 + q);
            const uint32_t *tw = T.tw(q0 + q), *twq = T.twq(q0 + q);
            for (int g = tid; g < GROUPS; g += THREADS) {
                const int pos = g & (h2 - 1);
                const int base = ((g - pos) << 2) + pos;
                uint32_t x0 = Y[base], x1 = Y[base + h2], x2 = Y[base + 2 * h2], x3 = Y[base + 3 * h2];
                const int ia = pos << s, ib = (pos + h2) << s, ic = pos << (s + 1);
                // stage s: (x0, x2), (x1, x3)
                uint32_t y0 = add_mod(x0, x2, p);
                uint32_t y2 = shoup(x0 - x2 + p, tw[ia], twq[ia], p);
                uint32_t y1 = add_mod(x1, x3, p);
                uint32_t y3 = shoup(x1 - x3 + p, tw[ib], twq[ib], p);
                // stage s+1: (y0, y1), (y2, y3)
                const uint32_t wc = tw[ic], wcq = twq[ic];
                Y[base] = add_mod(y0, y1, p);
                Y[base + h2] = shoup(y0 - y1 + p, wc, wcq, p);
                Y[base + 2 * h2] = add_mod(y2, y3, p);
                Y[base + 3 * h2] = shoup(y2 - y3 + p, wc, wcq, p);
            }
        }
        __syncthreads();
    }
    if (s < LOGM) {   // odd LOGM: final radix-2 stage, half = 1, twiddle 1
#pragma unroll
        for (int q = 0; q < NP; q++) {
            uint32_t *Y = X + q * M;
            const uint32_t p = prime(q0 + q);
            for (int g = tid; g < M / 2; g += THREADS) {
                uint32_t x0 = Y[2 * g], x1 = Y[2 * g + 1];
                Y[2 * g] = add_mod(x0, x1, p);
                Y[2 * g + 1] = sub_mod(x0, x1, p);
            }
        }
        __syncthreads();
    }
}

// Inverse DIT, bit-reversed -> natural, result scaled by M.
template <int LOGM, int NP>
__device__ void ntt_inv(uint32_t *X, const Twiddles &T, int q0)
{
    constexpr int M = 1 << LOGM;
    constexpr int GROUPS = M / 4;
    const int tid = threadIdx.x;
    int t = 0;
    if (LOGM & 1) {   // first radix-2 stage, half = 1, twiddle 1
#pragma unroll
        for (int q = 0; q < NP; q++) {
            uint32_t *Y = X + q * M;
            const uint32_t p = prime(q0 + q);
            for (int g = tid; g < M / 2; g += THREADS) {
                uint32_t x0 = Y[2 * g], x1 = Y[2 * g + 1];
                Y[2 * g] = add_mod(x0, x1, p);
                Y[2 * g + 1] = sub_mod(x0, x1, p);
            }
        }
        __syncthreads();
        t = 1;
    }
    for (; t + 1 < LOGM; t += 2) {
        const int h = 1 << t;
#pragma unroll
        for (int q = 0; q < NP; q++) {
            uint32_t *Y = X + q * M;
            const uint32_t p = prime(q0 + q);
            const uint32_t *itw = T.itw(q0 + q), *itwq = T.itwq(q0 + q);
            for (int g = tid; g < GROUPS; g += THREADS) {
                const int pos = g & (h - 1);
                const int base = ((g - pos) << 2) + pos;
                uint32_t x0 = Y[base], x1 = Y[base + h], x2 = Y[base + 2 * h], x3 = Y[base + 3 * h];
                // stage t (half h): (x0, x1), (x2, x3), twiddle w^-(pos * M/(2h))
                const int ia = pos << (LOGM - 1 - t);
                const uint32_t wa = itw[ia], waq = itwq[ia];
                uint32_t b1 = shoup(x1, wa, waq, p);
                uint32_t b3 = shoup(x3, wa, waq, p);
                uint32_t y0 = add_mod(x0, b1, p), y1 = sub_mod(x0, b1, p);
                uint32_t y2 = add_mod(x2, b3, p), y3 = sub_mod(x2, b3, p);
                // stage t+1 (half 2h): (y0, y2) with w^-(pos*M/(4h)),
                // (y1, y3) with w^-((pos+h)*M/(4h))
                const int ib = pos << (LOGM - 2 - t), ic = (pos + h) << (LOGM - 2 - t);
                uint32_t c2 = shoup(y2, itw[ib], itwq[ib], p);
                uint32_t c3 = shoup(y3, itw[ic], itwq[ic], p);
                Y[base] = add_mod(y0, c2, p);
                Y[base + 2 * h] = sub_mod(y0, c2, p);
                Y[base + h] = add_mod(y1, c3, p);
                Y[base + 3 * h] = sub_mod(y1, c3, p);
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------- scans
struct Scratch {            // LDS scratch beside the transform buffers
    uint64_t carry[THREADS];
    uint32_t wave[2 * WAVES + 8];
};

// Workgroup-wide carry-lookahead on one bit per thread: returns the carry
// into this thread given generate g and propagate p (never both), with a
// carry of 0 into thread 0.  Wave level: the carries of (G) + (G|P) are the
// ripple carries of the ballots; across waves a 16-step fold in LDS.
__device__ __forceinline__ uint32_t block_carry_in(bool g, bool pr, Scratch &sc)
{
    const int lane = __lane_id();
    const int w = threadIdx.x >> 6;
    const uint64_t G = __ballot(g), P = __ballot(pr);
    const uint64_t Y = G | P;
    const uint64_t s0 = G + Y;
    if (lane == 0) {
        sc.wave[2 * w] = (s0 < G) ? 1u : 0u;          // carry out of the wave, carry-in 0
        sc.wave[2 * w + 1] = (P == ~0ull) ? 1u : 0u;   // a carry-in passes straight through
    }
    __syncthreads();
    uint32_t cin = 0;
    for (int i = 0; i < w; i++) cin = sc.wave[2 * i] | (sc.wave[2 * i + 1] & cin);
    __syncthreads();
    const uint64_t s = G + Y + cin;
    return (uint32_t)(((s ^ G ^ Y) >> lane) & 1u);
}

__device__ __forceinline__ bool block_any(bool v, Scratch &sc)
{
    const int w = threadIdx.x >> 6;
    const uint64_t b = __ballot(v);
    if (__lane_id() == 0) sc.wave[2 * WAVES + 0 + 0] = 0u;
    __syncthreads();
    if (__lane_id() == 0 && b) atomicOr(&sc.wave[2 * WAVES], 1u);
    (void)w;
    __syncthreads();
    const bool r = sc.wave[2 * WAVES] != 0u;
    __syncthreads();
    return r;
}

// ---------------------------------------------------------- CRT and carry
// CRT constants: result x = a0 + P0 * t with a0 = y0 * s0 mod P0 and
// t = (y1 * s1 - a0) / P0 mod P1, where s0, s1 undo the transform's scale
// (M^-1, times 2^32 when the pointwise step was a Montgomery square).
struct CrtConst {
    uint32_t s0, s0q;       // scale for prime 0 (Shoup pair, mod P0)
    uint32_t c1, c1q;       // s1 * P0^-1 mod P1
    uint32_t c2, c2q;       // P0^-1 mod P1
};

// Limbs from the two residue vectors of a product: thread t reads positions
// [t*C, t*C+C) of Y0 and Y1 (LDS or registers via the functor) and writes
// normalised limbs to v[]; positions >= keep are zeroed (keep = M for all).
template <int C, class Get0, class Get1>
__device__ __forceinline__ void crt_carry(uint32_t (&v)[C], Get0 get0, Get1 get1,
                                          const CrtConst &k, int keep, Scratch &sc)
{
    const int t = threadIdx.x;
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < C; i++) {
        const uint32_t y0 = get0(i), y1 = get1(i);
        const uint32_t a0 = shoup(y0, k.s0, k.s0q, P0);
        const uint32_t u = shoup(y1, k.c1, k.c1q, P1);
        const uint32_t w = shoup(a0, k.c2, k.c2q, P1);
        const uint32_t tt = sub_mod(u, w, P1);
        const uint64_t x = (uint64_t)tt * P0 + a0;
        acc += x;
        v[i] = (uint32_t)acc & LIMB_MASK;
        acc >>= LIMB_BITS;
    }
    sc.carry[t] = acc;               // < 2^40
    __syncthreads();
    uint64_t cin = t ? sc.carry[t - 1] : 0ull;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < C; i++) {
        cin += v[i];
        v[i] = (uint32_t)cin & LIMB_MASK;
        cin >>= LIMB_BITS;
    }
    bool all1 = true;
#pragma unroll
    for (int i = 0; i < C; i++) all1 &= (v[i] == LIMB_MASK);
    const uint32_t c = block_carry_in(cin != 0ull, all1, sc);
    uint32_t cc = c;
#pragma unroll
    for (int i = 0; i < C; i++) {
        const uint32_t s = v[i] + cc;
        v[i] = s & LIMB_MASK;
        cc = s >> LIMB_BITS;
    }
    const int pos0 = t * C;
#pragma unroll
    for (int i = 0; i < C; i++)
        if (pos0 + i >= keep) v[i] = 0u;
}

// v += w (+ cin0 at position 0 of the whole vector), normalised, in place.
// Returns nothing; the carry out of the top thread is dropped.
template <int C>
__device__ __forceinline__ void add_vec(uint32_t (&v)[C], const uint32_t (&w)[C], uint32_t cin0,
                                        Scratch &sc)
{
    uint32_t cc = threadIdx.x == 0 ? cin0 : 0u;
#pragma unroll
    for (int i = 0; i < C; i++) {
        const uint32_t s = v[i] + w[i] + cc;
        v[i] = s & LIMB_MASK;
        cc = s >> LIMB_BITS;
    }
    bool all1 = true;
#pragma unroll
    for (int i = 0; i < C; i++) all1 &= (v[i] == LIMB_MASK);
    const uint32_t c = block_carry_in(cc != 0u, all1 && cc == 0u, sc);
    cc = c;
#pragma unroll
    for (int i = 0; i < C; i++) {
        const uint32_t s = v[i] + cc;
        v[i] = s & LIMB_MASK;
        cc = s >> LIMB_BITS;
    }
}

// d = v - w over the whole vector; returns true (in every thread) when the
// subtraction borrowed out of the top, i.e. v < w.
template <int C>
__device__ __forceinline__ bool sub_vec(uint32_t (&d)[C], const uint32_t (&v)[C],
                                        const uint32_t (&w)[C], Scratch &sc)
{
    uint32_t bb = 0;
#pragma unroll
    for (int i = 0; i < C; i++) {
        const uint32_t s = v[i] - w[i] - bb;        // in (-2^25, 2^24)
        d[i] = s & LIMB_MASK;
        bb = (s >> 31) & 1u;
    }
    bool all0 = true;
#pragma unroll
    for (int i = 0; i < C; i++) all0 &= (d[i] == 0u);
    // a borrow into this thread passes through iff every limb is 0
    const uint32_t b = block_carry_in(bb != 0u, all0 && bb == 0u, sc);
    bb = b;
#pragma unroll
    for (int i = 0; i < C; i++) {
        const uint32_t s = d[i] - bb;
        d[i] = s & LIMB_MASK;
        bb = (s >> 31) & 1u;
    }
    // borrow out of the top thread
    if (threadIdx.x == THREADS - 1) sc.wave[2 * WAVES + 1] = bb;
    __syncthreads();
    const bool r = sc.wave[2 * WAVES + 1] != 0u;
    __syncthreads();
    return r;
}

} // namespace mr
