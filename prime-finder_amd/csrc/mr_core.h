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

// The thread index, read through an empty asm statement.  Every address a
// thread derives from its index is loop-invariant, and the optimiser would
// otherwise hoist all of them (LDS and global addresses, twiddle loads) out
// of the squaring loop at once and spill registers; an opaque index makes
// each use recompute its few address operations where it is needed.
__device__ __forceinline__ int tid_()
{
    int t = (int)threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
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

// ------------------------------------------------------------- LDS layout
// A transform buffer of M words occupies MP = M + M/64 words of LDS: one
// pad word after every 64.  When each thread walks its own contiguous run
// of C = 8, 16 or 32 coefficients, the 64 lanes of a wave (stride C words)
// would share 64/C banks; with the pad they reach 64 different banks.  64
// lanes reading 64 consecutive coefficients still reach 64 consecutive
// words.
__device__ __forceinline__ int px(int n) { return n + (n >> 6); }
template <int LOGM>
struct Lds {
    static constexpr int M = 1 << LOGM;
    static constexpr int MP = M + (M >> 6);
};

// Padded address of element base + j*H of a radix-4 group with butterfly
// half-width H, from one per-group register: the padded base when H is a
// multiple of 64 (the j offsets are then constants), else the plain base.
template <int H>
__device__ __forceinline__ int grp_base(int base) { return H % 64 == 0 ? px(base) : base; }
template <int H>
__device__ __forceinline__ int grp_addr(int b, int j)
{
    if constexpr (H % 64 == 0) return b + j * (H + H / 64);
    else return px(b + j * H);
}

// ------------------------------------------------------------- tables
// Per (prime, M): tw[i] = w^i and itw[i] = w^-i for i < M/2 with their Shoup
// quotients, w a primitive M-th root of unity.  Layout of one table block:
// [tw | twq | itw | itwq], each M/2 words; prime 1 follows prime 0.
//
// Per-pass tables (`pass`): the three twiddles (and quotients) a radix-4
// pass needs at butterfly position pos, stored at consecutive pos, six
// arrays of H words per pass ([a | a' | b | b' | c | c']), so that the
// lanes of a wave, whose positions are consecutive, read consecutive words.
// Indexing the full tables instead (w^(pos << S)) puts the lanes 2^S words
// apart: one cache line per lane and per load.  Block of 2M words per
// (prime, direction): forward passes at their fwd_off, inverse passes at
// their inv_off (both below 2M).
struct Twiddles {
    const uint32_t *base;   // 2 primes * 4 * (M/2) words
    int half;               // M/2
    const uint32_t *pass;   // 2 primes * 2 directions * 2M words
    // Products of 2^15 points run as two transforms of 2^14 (base, half and
    // pass are then the 2^14 tables); big holds the 2^15 table w^i, w^-i
    // (i < 2^14, with quotients), big_half = 2^14, for the half twist.
    const uint32_t *big;
    int big_half;
    // Tables of the transform of M/2 points (base, pass, half) for the
    // wrapped products of unsplit lengths.
    const uint32_t *hbase, *hpass;
    int hhalf;
    __device__ __forceinline__ Twiddles halved() const
    {
        Twiddles h = *this;
        h.base = hbase; h.pass = hpass; h.half = hhalf;
        return h;
    }
    __device__ __forceinline__ const uint32_t *tw(int q) const { return base + (size_t)q * 4 * half; }
    __device__ __forceinline__ const uint32_t *twq(int q) const { return tw(q) + half; }
    __device__ __forceinline__ const uint32_t *itw(int q) const { return tw(q) + 2 * half; }
    __device__ __forceinline__ const uint32_t *itwq(int q) const { return tw(q) + 3 * half; }
    __device__ __forceinline__ const uint32_t *fwd(int q) const { return pass + (size_t)(2 * q) * 4 * half; }
    __device__ __forceinline__ const uint32_t *inv(int q) const { return pass + (size_t)(2 * q + 1) * 4 * half; }
    __device__ __forceinline__ const uint32_t *btw(int q) const { return big + (size_t)q * 4 * big_half; }
    __device__ __forceinline__ const uint32_t *btwq(int q) const { return btw(q) + big_half; }
    __device__ __forceinline__ const uint32_t *bitw(int q) const { return btw(q) + 2 * big_half; }
    __device__ __forceinline__ const uint32_t *bitwq(int q) const { return btw(q) + 3 * big_half; }
};

// Word offsets of the per-pass blocks (same formulas on the host).
__host__ __device__ constexpr int fwd_off(int logm, int S)
{
    int o = 0;
    for (int s2 = 0; s2 < S; s2 += 2) o += 6 * ((1 << logm) >> (s2 + 2));
    return o;
}
__host__ __device__ constexpr int inv_first(int logm) { return (logm & 1) ? 1 : 2; }
__host__ __device__ constexpr int inv_off(int logm, int T0)
{
    int o = 0;
    for (int t = inv_first(logm); t < T0; t += 2) o += 6 * (1 << t);
    return o;
}

// ------------------------------------------------------------- transforms
// ------------------------------------------------- fused transform passes
// The same transforms as ntt_fwd / ntt_inv, organised for throughput:
//  * per-thread trip counts are compile-time constants;
//  * in every pass whose butterfly half-width is <= THREADS a thread's
//    position inside its block, hence its twiddles, is the same for all its
//    groups, so the twiddles are loaded once per pass;
//  * the last forward pass and the first inverse pass act on the same four
//    consecutive elements (two for odd LOGM) with the constant twiddles 1 and
//    w^(+-M/4), so they are fused with the pointwise step in registers.

// Workgroup barrier for LDS hand-offs only: waits for this wave's LDS
// operations, not for its outstanding global loads (__syncthreads waits for
// both), so loads issued for the next pass stay in flight.  The memory
// clobber keeps the compiler from moving memory accesses across it.  Only
// for sections that write no global memory.
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Per-thread twiddles of one radix-4 pass (wa, wa', wb, wb', wc, wc' with
// their Shoup quotients) for NP primes.
template <int NP>
struct TwSet {
    uint32_t v[NP][6];
};

// A pass whose butterfly half-width is <= THREADS gives each thread one
// position inside its block, hence one twiddle set for all its groups.
template <int LOGM, int S>
__device__ constexpr bool dif_hoisted() { return ((1 << LOGM) >> (S + 2)) <= THREADS; }
template <int LOGM, int T0>
__device__ constexpr bool dit_hoisted() { return (1 << T0) <= THREADS; }

template <int LOGM, int S, int NP>
__device__ __forceinline__ TwSet<NP> dif_tw(const Twiddles &T, int q0)
{
    constexpr int H2 = (1 << LOGM) >> (S + 2), OFF = fwd_off(LOGM, S);
    const int pos = tid_() & (H2 - 1);
    TwSet<NP> t;
#pragma unroll
    for (int q = 0; q < NP; q++) {
        const uint32_t *b = T.fwd(q0 + q) + OFF + pos;
#pragma unroll
        for (int k = 0; k < 6; k++) t.v[q][k] = b[k * H2];
    }
    return t;
}

template <int LOGM, int T0, int NP>
__device__ __forceinline__ TwSet<NP> dit_tw(const Twiddles &T, int q0)
{
    constexpr int H = 1 << T0, OFF = inv_off(LOGM, T0);
    const int pos = tid_() & (H - 1);
    TwSet<NP> t;
#pragma unroll
    for (int q = 0; q < NP; q++) {
        const uint32_t *b = T.inv(q0 + q) + OFF + pos;
#pragma unroll
        for (int k = 0; k < 6; k++) t.v[q][k] = b[k * H];
    }
    return t;
}

// Group handled by a lane in a pass of butterfly half-width H < 64.  With
// the natural order (group = thread index + k * THREADS) a wave's 64 lanes
// cover 64/H consecutive blocks of 4H words, whose padded addresses fall on
// 16 to 32 of the 64 banks (2- to 4-way conflicts).  Rotating lane
// bits 4..5 (bit 5 when H = 32) up above the block bits that fall on
// distinct banks makes the wave's blocks lie 16 blocks apart, where the pad
// word per 64 shifts each block to its own banks: every access of these
// passes is then conflict-free.  A bit permutation of the group index, so
// still every group exactly once; the position inside a block (low bits)
// and hence the twiddles are unchanged.
template <int H, int GROUPS>
__device__ __forceinline__ int pass_group(int g)
{
    constexpr int h = __builtin_ctz(H);
    if constexpr (H >= 64) {
        return g;
    } else {
        constexpr int lo = h <= 4 ? 4 : 5, w = h <= 4 ? h + 2 : 5, k = h <= 4 ? 2 : 1;
        if constexpr ((1 << (lo + w)) > GROUPS) {
            return g;
        } else {
            const int F = (g >> lo) & ((1 << w) - 1);
            const int F2 = (F >> k) | ((F & ((1 << k) - 1)) << (w - k));
            return (g & ~(((1 << w) - 1) << lo)) | (F2 << lo);
        }
    }
}

// Regular forward radix-4 pass starting at stage S.  A thread's groups go
// in batches of GB per prime (at most 8 groups, 32 words, in registers):
// every word of a batch, for all NP primes, is read before the first
// butterfly result is written.  The compiler cannot tell that one group's
// stores miss the next group's loads, so loads placed after stores would
// wait for them and each wave would alternate between LDS latency and
// arithmetic.
template <int LOGM, int S, int NP>
__device__ __forceinline__ void dif_pass(uint32_t *X, const Twiddles &T, int q0, const TwSet<NP> &cur)
{
    constexpr int M = 1 << LOGM, H2 = M >> (S + 2), GROUPS = M / 4, MP = Lds<LOGM>::MP;
    constexpr int IT = GROUPS >= THREADS ? GROUPS / THREADS : 1;
    constexpr int MAXG = LOGM >= 15 ? 2 : 8;     // 2^15: registers hold 32 limbs already
    constexpr int GB = NP * IT > MAXG ? MAXG / NP : IT;
    const int tid = tid_();
    if (GROUPS >= THREADS || tid < GROUPS) {
#pragma unroll
        for (int i0 = 0; i0 < IT; i0 += GB) {
            int ad[GB][4];
            uint32_t x[NP][GB][4];
#pragma unroll
            for (int ib = 0; ib < GB; ib++) {
                const int g = pass_group<H2, GROUPS>(tid + (i0 + ib) * THREADS);
                const int pos = g & (H2 - 1);
                const int gb = grp_base<H2>(((g - pos) << 2) + pos);
#pragma unroll
                for (int j = 0; j < 4; j++) ad[ib][j] = grp_addr<H2>(gb, j);
            }
#pragma unroll
            for (int q = 0; q < NP; q++)
#pragma unroll
                for (int ib = 0; ib < GB; ib++)
#pragma unroll
                    for (int j = 0; j < 4; j++) x[q][ib][j] = X[q * MP + ad[ib][j]];
#pragma unroll
            for (int q = 0; q < NP; q++) {
                uint32_t *Y = X + q * MP;
                const uint32_t p = prime(q0 + q);
                const uint32_t *tp = T.fwd(q0 + q) + fwd_off(LOGM, S);
                uint32_t wa = 0, waq = 0, wb = 0, wbq = 0, wc = 0, wcq = 0;
                if constexpr (H2 <= THREADS) {
                    wa = cur.v[q][0]; waq = cur.v[q][1]; wb = cur.v[q][2];
                    wbq = cur.v[q][3]; wc = cur.v[q][4]; wcq = cur.v[q][5];
                }
#pragma unroll
                for (int ib = 0; ib < GB; ib++) {
                    if constexpr (H2 > THREADS) {
                        const int pos = pass_group<H2, GROUPS>(tid + (i0 + ib) * THREADS) & (H2 - 1);
                        const uint32_t *tb = tp + pos;
                        wa = tb[0]; waq = tb[H2]; wb = tb[2 * H2]; wbq = tb[3 * H2]; wc = tb[4 * H2]; wcq = tb[5 * H2];
                    }
                    const uint32_t x0 = x[q][ib][0], x1 = x[q][ib][1], x2 = x[q][ib][2], x3 = x[q][ib][3];
                    const uint32_t y0 = add_mod(x0, x2, p);
                    const uint32_t y2 = shoup(x0 - x2 + p, wa, waq, p);
                    const uint32_t y1 = add_mod(x1, x3, p);
                    const uint32_t y3 = shoup(x1 - x3 + p, wb, wbq, p);
                    Y[ad[ib][0]] = add_mod(y0, y1, p);
                    Y[ad[ib][1]] = shoup(y0 - y1 + p, wc, wcq, p);
                    Y[ad[ib][2]] = add_mod(y2, y3, p);
                    Y[ad[ib][3]] = shoup(y2 - y3 + p, wc, wcq, p);
                }
            }
        }
    }
    lds_barrier();
}

// Regular inverse radix-4 pass starting at stage T0 (half 2^T0); batched
// like dif_pass.
template <int LOGM, int T0, int NP>
__device__ __forceinline__ void dit_pass(uint32_t *X, const Twiddles &T, int q0, const TwSet<NP> &cur)
{
    constexpr int M = 1 << LOGM, H = 1 << T0, GROUPS = M / 4, MP = Lds<LOGM>::MP;
    constexpr int IT = GROUPS >= THREADS ? GROUPS / THREADS : 1;
    constexpr int MAXG = LOGM >= 15 ? 2 : 8;     // 2^15: registers hold 32 limbs already
    constexpr int GB = NP * IT > MAXG ? MAXG / NP : IT;
    const int tid = tid_();
    if (GROUPS >= THREADS || tid < GROUPS) {
#pragma unroll
        for (int i0 = 0; i0 < IT; i0 += GB) {
            int ad[GB][4];
            uint32_t x[NP][GB][4];
#pragma unroll
            for (int ib = 0; ib < GB; ib++) {
                const int g = pass_group<H, GROUPS>(tid + (i0 + ib) * THREADS);
                const int pos = g & (H - 1);
                const int gb = grp_base<H>(((g - pos) << 2) + pos);
#pragma unroll
                for (int j = 0; j < 4; j++) ad[ib][j] = grp_addr<H>(gb, j);
            }
#pragma unroll
            for (int q = 0; q < NP; q++)
#pragma unroll
                for (int ib = 0; ib < GB; ib++)
#pragma unroll
                    for (int j = 0; j < 4; j++) x[q][ib][j] = X[q * MP + ad[ib][j]];
#pragma unroll
            for (int q = 0; q < NP; q++) {
                uint32_t *Y = X + q * MP;
                const uint32_t p = prime(q0 + q);
                const uint32_t *tp = T.inv(q0 + q) + inv_off(LOGM, T0);
                uint32_t wa = 0, waq = 0, wb = 0, wbq = 0, wc = 0, wcq = 0;
                if constexpr (H <= THREADS) {
                    wa = cur.v[q][0]; waq = cur.v[q][1]; wb = cur.v[q][2];
                    wbq = cur.v[q][3]; wc = cur.v[q][4]; wcq = cur.v[q][5];
                }
#pragma unroll
                for (int ib = 0; ib < GB; ib++) {
                    if constexpr (H > THREADS) {
                        const int pos = pass_group<H, GROUPS>(tid + (i0 + ib) * THREADS) & (H - 1);
                        const uint32_t *tb = tp + pos;
                        wa = tb[0]; waq = tb[H]; wb = tb[2 * H]; wbq = tb[3 * H]; wc = tb[4 * H]; wcq = tb[5 * H];
                    }
                    const uint32_t x0 = x[q][ib][0], x1 = x[q][ib][1], x2 = x[q][ib][2], x3 = x[q][ib][3];
                    const uint32_t b1 = shoup(x1, wa, waq, p);
                    const uint32_t b3 = shoup(x3, wa, waq, p);
                    const uint32_t y0 = add_mod(x0, b1, p), y1 = sub_mod(x0, b1, p);
                    const uint32_t y2 = add_mod(x2, b3, p), y3 = sub_mod(x2, b3, p);
                    const uint32_t c2 = shoup(y2, wb, wbq, p);
                    const uint32_t c3 = shoup(y3, wc, wcq, p);
                    Y[ad[ib][0]] = add_mod(y0, c2, p);
                    Y[ad[ib][2]] = sub_mod(y0, c2, p);
                    Y[ad[ib][1]] = add_mod(y1, c3, p);
                    Y[ad[ib][3]] = sub_mod(y1, c3, p);
                }
            }
        }
    }
    lds_barrier();
}

// Forward passes S, S+2, ... while the pass is not the last one.  The
// next pass's twiddles are loaded before this pass computes, so their
// latency hides behind it (the pass barriers wait for LDS only).
template <int LOGM, int S, int NP>
__device__ __forceinline__ void dif_regular(uint32_t *X, const Twiddles &T, int q0, const TwSet<NP> &cur)
{
    constexpr int LAST = (LOGM & 1) ? LOGM - 1 : LOGM - 2;   // first stage of the final pass
    if constexpr (S < LAST) {
        constexpr int S2 = S + 2;
        TwSet<NP> nxt{};
        if constexpr (S2 < LAST && dif_hoisted<LOGM, S2>()) nxt = dif_tw<LOGM, S2, NP>(T, q0);
        dif_pass<LOGM, S, NP>(X, T, q0, cur);
        dif_regular<LOGM, S2, NP>(X, T, q0, nxt);
    }
}

// Inverse passes after the first one.
template <int LOGM, int T0, int NP>
__device__ __forceinline__ void dit_regular(uint32_t *X, const Twiddles &T, int q0, const TwSet<NP> &cur)
{
    if constexpr (T0 + 1 < LOGM) {
        constexpr int T2 = T0 + 2;
        TwSet<NP> nxt{};
        if constexpr (T2 + 1 < LOGM && dit_hoisted<LOGM, T2>()) nxt = dit_tw<LOGM, T2, NP>(T, q0);
        dit_pass<LOGM, T0, NP>(X, T, q0, cur);
        dit_regular<LOGM, T2, NP>(X, T, q0, nxt);
    }
}

// Final forward pass (stages LOGM-2, LOGM-1, or the lone radix-2 stage),
// optionally followed by the pointwise step and the first inverse pass.
// MODE 0: forward only; MODE 1: pointwise square when B is null, else
// times B (kept times 2^32) -- one code path, a uniform branch.
template <int LOGM, int NP, int MODE>
__device__ __forceinline__ void middle(uint32_t *X, const Twiddles &T, int q0, const uint32_t *B,
                                       const uint32_t *np_tab, int bstride)
{
    constexpr int M = 1 << LOGM, MP = Lds<LOGM>::MP;
    const int tid = tid_();
#pragma unroll
    for (int q = 0; q < NP; q++) {
        uint32_t *Y = X + q * MP;
        const uint32_t p = prime(q0 + q), np = np_tab[q0 + q];
        const uint32_t *Bq = B ? B + (size_t)(q0 + q) * bstride : nullptr;
        if constexpr ((LOGM & 1) == 0) {
            constexpr int GROUPS = M / 4;
            constexpr int IT = GROUPS >= THREADS ? GROUPS / THREADS : 1;
            const uint32_t *tw = T.tw(q0 + q), *itw = T.itw(q0 + q);
            const uint32_t w4 = tw[M / 4], w4q = tw[M / 2 + M / 4];     // tw, twq blocks
            const uint32_t v4 = itw[M / 4], v4q = itw[M / 2 + M / 4];
            if (GROUPS >= THREADS || tid < GROUPS) {
#pragma unroll
                for (int it = 0; it < IT; it++) {
                    const int g = tid + it * THREADS;
                    uint32_t *Yg = Y + px(4 * g);      // 4 words inside one 64-word block
                    uint4 v = make_uint4(Yg[0], Yg[1], Yg[2], Yg[3]);
                    // DIF stages LOGM-2 (twiddles 1, w^(M/4)) and LOGM-1 (twiddle 1)
                    const uint32_t y0 = add_mod(v.x, v.z, p), y2 = sub_mod(v.x, v.z, p);
                    const uint32_t y1 = add_mod(v.y, v.w, p);
                    const uint32_t y3 = shoup(v.y - v.w + p, w4, w4q, p);
                    uint32_t z0 = add_mod(y0, y1, p), z1 = sub_mod(y0, y1, p);
                    uint32_t z2 = add_mod(y2, y3, p), z3 = sub_mod(y2, y3, p);
                    if constexpr (MODE != 0) {
                        uint4 bb = make_uint4(z0, z1, z2, z3);
                        if (Bq) bb = *reinterpret_cast<const uint4 *>(Bq + 4 * g);
                        z0 = mont(z0, bb.x, p, np); z1 = mont(z1, bb.y, p, np);
                        z2 = mont(z2, bb.z, p, np); z3 = mont(z3, bb.w, p, np);
                        // DIT stages 0 (twiddle 1) and 1 (twiddles 1, w^-(M/4))
                        const uint32_t a0 = add_mod(z0, z1, p), a1 = sub_mod(z0, z1, p);
                        const uint32_t a2 = add_mod(z2, z3, p), a3 = sub_mod(z2, z3, p);
                        const uint32_t c3 = shoup(a3, v4, v4q, p);
                        v.x = add_mod(a0, a2, p);
                        v.z = sub_mod(a0, a2, p);
                        v.y = add_mod(a1, c3, p);
                        v.w = sub_mod(a1, c3, p);
                    } else {
                        v.x = z0; v.y = z1; v.z = z2; v.w = z3;
                    }
                    Yg[0] = v.x; Yg[1] = v.y; Yg[2] = v.z; Yg[3] = v.w;
                }
            }
        } else {
            constexpr int PAIRS = M / 2;
            constexpr int IT = PAIRS >= THREADS ? PAIRS / THREADS : 1;
            if (PAIRS >= THREADS || tid < PAIRS) {
#pragma unroll
                for (int it = 0; it < IT; it++) {
                    const int g = tid + it * THREADS;
                    uint32_t *Yg = Y + px(2 * g);
                    uint2 v = make_uint2(Yg[0], Yg[1]);
                    uint32_t z0 = add_mod(v.x, v.y, p), z1 = sub_mod(v.x, v.y, p);
                    if constexpr (MODE != 0) {
                        uint2 bb = make_uint2(z0, z1);
                        if (Bq) bb = *reinterpret_cast<const uint2 *>(Bq + 2 * g);
                        z0 = mont(z0, bb.x, p, np); z1 = mont(z1, bb.y, p, np);
                        v.x = add_mod(z0, z1, p);
                        v.y = sub_mod(z0, z1, p);
                    } else {
                        v.x = z0; v.y = z1;
                    }
                    Yg[0] = v.x; Yg[1] = v.y;
                }
            }
        }
    }
    lds_barrier();
}

// Cyclic convolution of the NP buffers with themselves (B null) or with the
// stored transform B (prime q's at B + q * bstride); output scaled by M (and
// by 2^-32 for the square).
template <int LOGM, int NP>
__device__ __forceinline__ void convolve(uint32_t *X, const Twiddles &T, int q0, const uint32_t *B,
                                         const uint32_t *np_tab, int bstride = 1 << LOGM)
{
    constexpr int T0 = (LOGM & 1) ? 1 : 2;     // first inverse pass after middle
    TwSet<NP> f{}, d{};
    if constexpr (dif_hoisted<LOGM, 0>()) f = dif_tw<LOGM, 0, NP>(T, q0);
    dif_regular<LOGM, 0, NP>(X, T, q0, f);
    if constexpr (T0 + 1 < LOGM && dit_hoisted<LOGM, T0>()) d = dit_tw<LOGM, T0, NP>(T, q0);
    middle<LOGM, NP, 1>(X, T, q0, B, np_tab, bstride);
    dit_regular<LOGM, T0, NP>(X, T, q0, d);
}

// Forward transform only (bit-reversed output).
template <int LOGM, int NP>
__device__ __forceinline__ void forward(uint32_t *X, const Twiddles &T, int q0, const uint32_t *np_tab)
{
    TwSet<NP> f{};
    if constexpr (dif_hoisted<LOGM, 0>()) f = dif_tw<LOGM, 0, NP>(T, q0);
    dif_regular<LOGM, 0, NP>(X, T, q0, f);
    middle<LOGM, NP, 0>(X, T, q0, nullptr, np_tab, 0);
}

// ------------------------------------------------------------- scans
// (their barriers hand off LDS only: lds_barrier, which does not wait for
// the wave's outstanding global stores)
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
    const int w = tid_() >> 6;
    const uint64_t G = __ballot(g), P = __ballot(pr);
    const uint64_t Y = G | P;
    const uint64_t s0 = G + Y;
    if (lane == 0) {
        sc.wave[2 * w] = (s0 < G) ? 1u : 0u;          // carry out of the wave, carry-in 0
        sc.wave[2 * w + 1] = (P == ~0ull) ? 1u : 0u;   // a carry-in passes straight through
    }
    lds_barrier();
    uint32_t cin = 0;
    for (int i = 0; i < w; i++) cin = sc.wave[2 * i] | (sc.wave[2 * i + 1] & cin);
    lds_barrier();
    const uint64_t s = G + Y + cin;
    return (uint32_t)(((s ^ G ^ Y) >> lane) & 1u);
}

__device__ __forceinline__ bool block_any(bool v, Scratch &sc)
{
    const int w = tid_() >> 6;
    const uint64_t b = __ballot(v);
    if (__lane_id() == 0) sc.wave[2 * WAVES + 0 + 0] = 0u;
    lds_barrier();
    if (__lane_id() == 0 && b) atomicOr(&sc.wave[2 * WAVES], 1u);
    (void)w;
    lds_barrier();
    const bool r = sc.wave[2 * WAVES] != 0u;
    lds_barrier();
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
// normalised limbs to v[]; positions >= keep are zeroed (keep = M for all);
// inputs at positions >= live are known to be zero.
template <int C, class Get0, class Get1>
__device__ __forceinline__ void crt_carry(uint32_t (&v)[C], Get0 get0, Get1 get1,
                                          const CrtConst &k, int keep, Scratch &sc,
                                          int live = 1 << 30)
{
    const int t = tid_();
    uint64_t acc = 0;
    // inputs in batches of CB, all of a batch's reads issued before use
    // (some callers read them from global memory).  A thread whose run lies
    // wholly at or above keep (its limbs are zeroed below, and a carry out
    // of it lands above keep as well) or at or above live (all its inputs
    // are zero) reads and converts nothing; carries still pass through it.
    constexpr int CB = C < 8 ? C : 8;
    if (t * C >= min(keep, live)) {
#pragma unroll
        for (int i = 0; i < C; i++) v[i] = 0u;
    } else {
#pragma unroll
        for (int i0 = 0; i0 < C; i0 += CB) {
            uint32_t y0[CB], y1[CB];
#pragma unroll
            for (int i = 0; i < CB; i++) { y0[i] = get0(i0 + i); y1[i] = get1(i0 + i); }
#pragma unroll
            for (int i = 0; i < CB; i++) {
                const uint32_t a0 = shoup(y0[i], k.s0, k.s0q, P0);
                const uint32_t u = shoup(y1[i], k.c1, k.c1q, P1);
                const uint32_t w = shoup(a0, k.c2, k.c2q, P1);
                const uint32_t tt = sub_mod(u, w, P1);
                const uint64_t x = (uint64_t)tt * P0 + a0;
                acc += x;
                v[i0 + i] = (uint32_t)acc & LIMB_MASK;
                acc >>= LIMB_BITS;
            }
        }
    }
    sc.carry[t] = acc;               // < 2^40
    lds_barrier();
    uint64_t cin = t ? sc.carry[t - 1] : 0ull;
    lds_barrier();
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
    uint32_t cc = tid_() == 0 ? cin0 : 0u;
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
    const uint32_t own = bb;                        // borrow this thread generates
    // a borrow into this thread passes through iff every limb is 0
    const uint32_t b = block_carry_in(bb != 0u, all0 && bb == 0u, sc);
    bb = b;
#pragma unroll
    for (int i = 0; i < C; i++) {
        const uint32_t s = d[i] - bb;
        d[i] = s & LIMB_MASK;
        bb = (s >> 31) & 1u;
    }
    // borrow out of the top thread: generated there, or passed through it
    // (never both: a thread that generates one has a nonzero limb left)
    if (tid_() == THREADS - 1) sc.wave[2 * WAVES + 1] = own | bb;
    lds_barrier();
    const bool r = sc.wave[2 * WAVES + 1] != 0u;
    lds_barrier();
    return r;
}

} // namespace mr
