// mr_kernels.hip -- batched Miller-Rabin on gfx950 (kernels, host context,
// C ABI of include/mr_gpu.h).
//
// Reference behaviour replaced: miller_rabin_gpu/miller_rabin_runner.cu
// (window_exp_loop :89-203, gpu_miller_rabin_s1 :281-359, gpu_miller_rabin
// :365-473) and batch_mod_ctx.cu (modsq_batch / modmul_batch :257-316,
// check_passed_kernel :18-49).  Design differs: one workgroup owns one
// candidate for a whole chunk of squarings, keeping the transform buffers in
// LDS and every intermediate in registers, instead of ~39 separate batched
// launches per squaring.
//
// Arithmetic per squaring (Montgomery, R = 2^(24 L)):
//   T = r^2;  m = (T mod R) N' mod R;  t = (T + m N) / R;  t -= N if t >= N
// as three transform products (r*r, T_lo*N', m*N) with NTT(N') and NTT(N)
// precomputed.  Since T + m N == 0 (mod R), the low half of the sum is not
// formed: t = T_hi + (mN)_hi + [T_lo != 0].
// Exponentiation is left-to-right binary over the bits of d; a set bit
// multiplies by the base, which is below 2^24, by double-and-add of modular
// additions instead of a transform product.
#include "mr_core.h"
#include "../../include/mr_gpu.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <algorithm>
#include <string>
#include <new>
#include <type_traits>
#include <vector>

namespace mr {

// ------------------------------------------------------------------ device

struct Dev {                    // per-batch device state (kernel argument)
    int L;                      // largest limb count of the context, 2L < M
    int stride;                 // words per linear limb vector (>= L + 1)
    int n;                      // candidate slots
    const int32_t *list;        // launch index -> slot (nullptr: identity)
    const uint8_t *run;         // [n] scheduler graph: 1 = the slot squares in this chunk
                                //   (nullptr: every launched slot)
    int32_t *Ls;                // [n] limbs of the slot's number (R = 2^(24 Ls))
    uint32_t *base;             // [n] base of the current round
    int32_t *bidx;              // [n] index of that base in the scheduler's list
    uint8_t *done;              // [n] scheduler: 1 = verdict final (in alive)
    uint32_t *N;                // [n][stride] modulus limbs
    uint32_t *Nh;               // [n][M] N in the high layout (positions [L, 2L)),
                                //   lane-interleaved: element i of thread t at i * THREADS + t
    uint32_t *Nf;               // [n][2M] NTT(N) * 2^32 per prime
    uint32_t *Npf;              // [n][2M] NTT(N') * 2^32 per prime
    uint32_t *Nf2;              // [n][M] NTT of M/2 points of N * 2^32 per prime (HALF lengths)
    uint32_t *one;              // [n][stride] R mod N
    uint32_t *nm1;              // [n][stride] (N - 1) R mod N
    uint32_t *res;              // [n][stride] current residue (Montgomery form)
    uint32_t *scr;              // [n][2M] transform scratch
    uint32_t *ysv;              // [n][M + M/2] for split products: first half's inverse, staged operand
    int32_t *nbits;             // [n]
    int32_t *sval;              // [n] s with N - 1 = 2^s d
    int32_t *bitpos;            // [n] next exponent bit (bit index of N), < s when done
    uint8_t *alive;             // [n]
    uint8_t *pre;               // [n] 0 = test, 1 = prime (2, 3), 2 = composite (even, < 2)
    uint32_t *err;              // [1] error latch
};

__device__ __forceinline__ int slot_id(const Dev &D)
{
    return D.list ? D.list[blockIdx.x] : (int)blockIdx.x;
}

struct Consts {
    CrtConst sq, mul;           // CRT scales after Montgomery square / exact product
    CrtConst mulh;              // exact product through a transform of M/2 points
    uint32_t r2[2];             // 2^64 mod p (to Montgomery form)
    uint32_t np[2];             // -p^-1 mod 2^32
};

// Geometry of a product of M = 2^LOGM points.  Up to 2^14 one transform of
// M points per prime; at 2^15 (SPLIT) two transforms of 2^14 per prime: the
// operands' upper halves are zero, so the 2^15 transform's first stage
// splits it into the transforms of a and of a * w^n (w a 2^15-th root of
// unity), which give the even and the odd frequencies (the two halves of the
// bit-reversed output); the inverse's last stage recombines
// c[n] = u[n] + w^-n v[n], c[n + 2^14] = u[n] - w^-n v[n].  Both primes'
// buffers of TL = 14 fit the LDS at once either way.
template <int LOGM>
struct G {
    static constexpr int M = 1 << LOGM;
    static constexpr int C = M / THREADS;
    static constexpr int NP = 2;
    static constexpr bool SPLIT = LOGM > 14;
    static constexpr int TL = SPLIT ? 14 : LOGM;      // log2 of each transform
    static constexpr int H = 1 << TL;
    static constexpr int YSV = SPLIT ? M + H : 0;      // per-slot words of Dev::ysv
    // Products known modulo B^(M/2) - 1 only (m N in the Montgomery step)
    // run through transforms of M/2 points: at 2^15 the first split half,
    // below that a transform of LOGM - 1 (none at the smallest length).
    static constexpr bool WRAP = LOGM >= 12;
    static constexpr bool HALF = WRAP && !SPLIT;       // own M/2 transform
};

template <int LOGM>
struct Smem {
    uint32_t X[2 * Lds<G<LOGM>::TL>::MP];           // padded transform buffers, one per prime
    Scratch sc;
    uint32_t bcast[8];
};

// v[i] <- src[pos - off] for pos = t*C + i in [off, off + len), else 0.
template <int C>
__device__ __forceinline__ void load_vec(uint32_t (&v)[C], const uint32_t *src, int len, int off)
{
    const int p0 = tid_() * C;
#pragma unroll
    for (int i = 0; i < C; i++) {
        const int idx = p0 + i - off;
        v[i] = (idx >= 0 && idx < len) ? src[idx] : 0u;
    }
}

template <int C>
__device__ __forceinline__ void store_vec(const uint32_t (&v)[C], uint32_t *dst, int len, int off)
{
    const int p0 = tid_() * C;
#pragma unroll
    for (int i = 0; i < C; i++) {
        const int idx = p0 + i - off;
        if (idx >= 0 && idx < len) dst[idx] = v[i];
    }
}

// A thread's own elements (positions t*C + i in [lo, hi)) parked in global
// memory lane-interleaved, element i of thread t at buf[i * THREADS + t], so
// that each store or load instruction of a wave touches consecutive words;
// only the same thread reads them back.  load_own zero-fills outside
// [lo, hi).
template <int C>
__device__ __forceinline__ void save_own(const uint32_t (&v)[C], uint32_t *buf, int lo, int hi)
{
    const int t = tid_(), p0 = t * C;
#pragma unroll
    for (int i = 0; i < C; i++)
        if (p0 + i >= lo && p0 + i < hi) buf[i * THREADS + t] = v[i];
}

template <int C>
__device__ __forceinline__ void load_own(uint32_t (&v)[C], const uint32_t *buf, int lo, int hi)
{
    const int t = tid_(), p0 = t * C;
#pragma unroll
    for (int i = 0; i < C; i++)
        v[i] = (p0 + i >= lo && p0 + i < hi) ? buf[i * THREADS + t] : 0u;
}

template <int C>
__device__ __forceinline__ void zero_below(uint32_t (&v)[C], int lim)
{
    const int p0 = tid_() * C;
#pragma unroll
    for (int i = 0; i < C; i++)
        if (p0 + i < lim) v[i] = 0u;
}

// Put positions [off, off+len) of v at X_q[0..len), zero elsewhere, for the
// NPW buffers starting at X.
// Only positions below H = 2^TLS exist in the buffers of 2^TLS points (the
// operand has at most L < H limbs).
template <int LOGM, int NPW, int TLS = G<LOGM>::TL>
__device__ __forceinline__ void stage_in(uint32_t *X, const uint32_t (&v)[G<LOGM>::C], int off, int len)
{
    constexpr int C = G<LOGM>::C, MP = Lds<TLS>::MP, H = 1 << TLS;
    const int p0 = tid_() * C;
    if (off != 0) {
        if (p0 < H) {
#pragma unroll
            for (int q = 0; q < NPW; q++)
#pragma unroll
                for (int i = 0; i < C; i++) X[q * MP + px(p0 + i)] = 0u;
        }
        lds_barrier();
#pragma unroll
        for (int i = 0; i < C; i++) {
            const int idx = p0 + i - off;
            if (idx >= 0 && idx < len)
#pragma unroll
                for (int q = 0; q < NPW; q++) X[q * MP + px(idx)] = v[i];
        }
    } else if (p0 < H) {
#pragma unroll
        for (int i = 0; i < C; i++) {
            const uint32_t x = (p0 + i < len) ? v[i] : 0u;
#pragma unroll
            for (int q = 0; q < NPW; q++) X[q * MP + px(p0 + i)] = x;
        }
    }
    lds_barrier();
}

// Split products: the staged operand (natural order, prime 0's buffer) is
// copied to save (H words, coalesced) ...
template <int LOGM>
__device__ __forceinline__ void split_save(const uint32_t *X, uint32_t *save)
{
    constexpr int H = G<LOGM>::H, KN = H / THREADS;
    const int t = tid_();
#pragma unroll
    for (int k = 0; k < KN; k++) save[t + k * THREADS] = X[px(t + k * THREADS)];
    __syncthreads();
}

// ... and restaged for the second half as a[n] * w^n in both buffers.  The
// loops have compile-time trip counts and issue a batch's global loads
// before using any, so that their latency is paid once per batch.
template <int LOGM>
__device__ __forceinline__ void split_twist(uint32_t *X, const Twiddles &T, const uint32_t *save)
{
    constexpr int H = G<LOGM>::H, MP = Lds<G<LOGM>::TL>::MP, KN = H / THREADS, KB = 8;
    const int t = tid_();
    const uint32_t *w0 = T.btw(0), *q0 = T.btwq(0), *w1 = T.btw(1), *q1 = T.btwq(1);
#pragma unroll
    for (int k0 = 0; k0 < KN; k0 += KB) {
        uint32_t x[KB], a0[KB], b0[KB], a1[KB], b1[KB];
#pragma unroll
        for (int k = 0; k < KB; k++) {
            const int n = t + (k0 + k) * THREADS;
            x[k] = save[n]; a0[k] = w0[n]; b0[k] = q0[n]; a1[k] = w1[n]; b1[k] = q1[n];
        }
#pragma unroll
        for (int k = 0; k < KB; k++) {
            const int n = t + (k0 + k) * THREADS;
            X[px(n)] = shoup(x[k], a0[k], b0[k], P0);
            X[MP + px(n)] = shoup(x[k], a1[k], b1[k], P1);
        }
    }
    lds_barrier();
}

// Inverse of the second half, times w^-n (the recombination's weight).
template <int LOGM>
__device__ __forceinline__ void split_unweight(uint32_t *X, const Twiddles &T)
{
    constexpr int H = G<LOGM>::H, MP = Lds<G<LOGM>::TL>::MP, KN = H / THREADS, KB = 8;
    const int t = tid_();
    const uint32_t *w0 = T.bitw(0), *q0 = T.bitwq(0), *w1 = T.bitw(1), *q1 = T.bitwq(1);
#pragma unroll
    for (int k0 = 0; k0 < KN; k0 += KB) {
        uint32_t a0[KB], b0[KB], a1[KB], b1[KB];
#pragma unroll
        for (int k = 0; k < KB; k++) {
            const int n = t + (k0 + k) * THREADS;
            a0[k] = w0[n]; b0[k] = q0[n]; a1[k] = w1[n]; b1[k] = q1[n];
        }
#pragma unroll
        for (int k = 0; k < KB; k++) {
            const int n = t + (k0 + k) * THREADS;
            X[px(n)] = shoup(X[px(n)], a0[k], b0[k], P0);
            X[MP + px(n)] = shoup(X[MP + px(n)], a1[k], b1[k], P1);
        }
    }
    lds_barrier();
}

// out = A * B as normalised limbs (positions >= keep zeroed).  A is
// positions [off, off+len) of a; B is a (square) when Bf == nullptr, else the
// precomputed transform Bf.  a and out may alias.  Split products (M = 2^15)
// park the first half's inverse u in ysave (M words, lane-interleaved by
// the reading thread) and the staged operand at ysave + M (H words).
// wrap (split products only): out = the cyclic product of length H, i.e.
// A * B mod (x^H - 1) evaluated at x = 2^24, normalised over [0, M) -- the
// first half alone, whose inverse transform is that cyclic product.
template <int LOGM>
__device__ __forceinline__ void tproduct(Smem<LOGM> &sm, const Twiddles &T, const Consts &K,
                         const uint32_t (&a)[G<LOGM>::C], int off, int len,
                         const uint32_t *Bf, uint32_t (&out)[G<LOGM>::C], int keep,
                         uint32_t *ysave, bool wrap = false)
{
    constexpr int C = G<LOGM>::C, M = G<LOGM>::M, TL = G<LOGM>::TL, MP = Lds<TL>::MP;
    const CrtConst k = Bf ? K.mul : K.sq;
    const int p0 = tid_() * C;
    const uint32_t *X0 = sm.X, *X1 = sm.X + MP;
    if constexpr (!G<LOGM>::SPLIT) {
        // unsplit: M points, or M/2 points (buffers at the smaller pitch) for
        // a wrapped product
        bool half = false;
        if constexpr (G<LOGM>::HALF) {
            if (wrap) {
                half = true;
                stage_in<LOGM, 2, LOGM - 1>(sm.X, a, off, len);
                convolve<LOGM - 1, 2>(sm.X, T.halved(), 0, Bf, K.np);
            }
        }
        if (!half) {
            stage_in<LOGM, 2>(sm.X, a, off, len);
            convolve<LOGM, 2>(sm.X, T, 0, Bf, K.np);
        }
        // a thread's run lies wholly below or above M/2; loads stay in the
        // buffer either way (unconditional)
        const bool in = !half || p0 < M / 2;
        const int b0 = in ? p0 : 0;
        const uint32_t *Y1 = half ? sm.X + Lds<G<LOGM>::HALF ? LOGM - 1 : LOGM>::MP : X1;
        crt_carry<C>(out,
                     [&](int i) {
                         const uint32_t y = X0[px(b0 + i)];
                         return in ? y : 0u;
                     },
                     [&](int i) {
                         const uint32_t y = Y1[px(b0 + i)];
                         return in ? y : 0u;
                     },
                     half ? K.mulh : k, keep, sm.sc, half ? M / 2 : M);
    } else {
        constexpr int H = G<LOGM>::H, HALF_T = THREADS / 2;   // threads covering [0, H)
        static_assert(C * HALF_T == H, "split layout");
        stage_in<LOGM, 2>(sm.X, a, off, len);
        uint32_t *asave = ysave + M;
        if (!wrap) split_save<LOGM>(sm.X, asave);
        convolve<TL, 2>(sm.X, T, 0, Bf, K.np, M);                      // even frequencies
        if (wrap) {
            const bool lo = p0 < H;
            crt_carry<C>(out, [&](int i) { return lo ? X0[px(p0 + i)] : 0u; },
                         [&](int i) { return lo ? X1[px(p0 + i)] : 0u; }, K.mulh, keep, sm.sc, H);
            return;
        }
        {   // u -> ysave: element i of position block t' (n = t' C + i) at
            // q H + i HALF_T + t'; thread t writes half of t mod HALF_T's run
            const int t = tid_(), tp = t & (HALF_T - 1), i0 = (t / HALF_T) * (C / 2);
#pragma unroll
            for (int i = 0; i < C / 2; i++) {
                ysave[(i0 + i) * HALF_T + tp] = X0[px(tp * C + i0 + i)];
                ysave[H + (i0 + i) * HALF_T + tp] = X1[px(tp * C + i0 + i)];
            }
        }
        __syncthreads();
        split_twist<LOGM>(sm.X, T, asave);
        convolve<TL, 2>(sm.X, T, 0, Bf ? Bf + H : nullptr, K.np, M);   // odd frequencies
        split_unweight<LOGM>(sm.X, T);
        const int tp = tid_() & (HALF_T - 1);
        const bool hi = p0 >= H;
        crt_carry<C>(out,
                     [&](int i) {
                         const uint32_t u = ysave[i * HALF_T + tp], v = X0[px(tp * C + i)];
                         return hi ? sub_mod(u, v, P0) : add_mod(u, v, P0);
                     },
                     [&](int i) {
                         const uint32_t u = ysave[H + i * HALF_T + tp], v = X1[px(tp * C + i)];
                         return hi ? sub_mod(u, v, P1) : add_mod(u, v, P1);
                     },
                     k, keep, sm.sc);
    }
}

// dst[0..2M) = transform of positions [0, len) of v, times 2^32, per prime
// (split products: the two halves' transforms side by side).  save: H
// scratch words for split products.
template <int LOGM>
__device__ __forceinline__ void to_fft(Smem<LOGM> &sm, const Twiddles &T, const Consts &K,
                       const uint32_t (&v)[G<LOGM>::C], int len, uint32_t *dst, uint32_t *save)
{
    constexpr int M = G<LOGM>::M, TL = G<LOGM>::TL, H = G<LOGM>::H, MP = Lds<TL>::MP;
    stage_in<LOGM, 2>(sm.X, v, 0, len);
    if constexpr (G<LOGM>::SPLIT) split_save<LOGM>(sm.X, save);
    for (int h = 0; h < (G<LOGM>::SPLIT ? 2 : 1); h++) {
        if (h == 1) split_twist<LOGM>(sm.X, T, save);
        forward<TL, 2>(sm.X, T, 0, K.np);
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const uint32_t p = prime(q), np = K.np[q], r2 = K.r2[q];
            for (int i = tid_(); i < H; i += THREADS)
                dst[(size_t)q * M + h * H + i] = mont(sm.X[q * MP + px(i)], r2, p, np);
        }
        __syncthreads();
    }
}

// dst[0..M) = transform of M/2 points of positions [0, len) of v (len <
// M/2), times 2^32, per prime (prime q's at dst + q M/2).
template <int LOGM>
__device__ __forceinline__ void to_fft_half(Smem<LOGM> &sm, const Twiddles &T, const Consts &K,
                            const uint32_t (&v)[G<LOGM>::C], int len, uint32_t *dst)
{
    constexpr int WL = LOGM - 1, HW = 1 << WL, WP = Lds<WL>::MP;
    stage_in<LOGM, 2, WL>(sm.X, v, 0, len);
    forward<WL, 2>(sm.X, T.halved(), 0, K.np);
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const uint32_t p = prime(q), np = K.np[q], r2 = K.r2[q];
        for (int i = tid_(); i < HW; i += THREADS)
            dst[(size_t)q * HW + i] = mont(sm.X[q * WP + px(i)], r2, p, np);
    }
    __syncthreads();
}

// r <- r - N when r >= N; N read from Nl at layout offset off, or (Nil
// non-null) from its lane-interleaved copy (coalesced).  r < 2N.
template <int C>
__device__ __forceinline__ void reduce_once(uint32_t (&r)[C], const uint32_t *Nl, int L, int off,
                                            Scratch &sc, const uint32_t *Nil = nullptr)
{
    uint32_t t[C], d[C];
    if (Nil) load_own<C>(t, Nil, 0, C * THREADS);
    else load_vec<C>(t, Nl, L, off);
    const bool lt = sub_vec<C>(d, r, t, sc);
    if (!lt) {
#pragma unroll
        for (int i = 0; i < C; i++) r[i] = d[i];
    }
}

// r <- r * s mod N for r < N, 1 <= s < 2^24 (double-and-add of modular
// additions).  The multiplicand waits in LDS (stash, padded, MP words) so
// that at most three limb vectors are live.
template <int C>
__device__ __forceinline__ void mulsmall_modN(uint32_t (&r)[C], uint32_t s, const uint32_t *Nl, int L, int off,
                              uint32_t *stash, Scratch &sc, const uint32_t *Nil = nullptr)
{
    const int p0 = tid_() * C;
#pragma unroll
    for (int i = 0; i < C; i++) stash[px(p0 + i)] = r[i];   // own slots only: no barrier
    const int top = 31 - __clz((int)s);
    for (int b = top - 1; b >= 0; b--) {
        add_vec<C>(r, r, 0u, sc);                            // 2r < 2N
        reduce_once<C>(r, Nl, L, off, sc, Nil);
        if ((s >> b) & 1u) {
            uint32_t t[C];
#pragma unroll
            for (int i = 0; i < C; i++) t[i] = stash[px(p0 + i)];
            add_vec<C>(r, t, 0u, sc);
            reduce_once<C>(r, Nl, L, off, sc, Nil);
        }
    }
}

// Montgomery tail after a wrapped m N (W = m N mod (B^H - 1), B = 2^24,
// H = M/2 > L, normalised over [0, M) in r; T parked in park over [0, 2L)):
// T + m N = t B^L with t < 2N, so V = T + W == t B^L (mod B^H - 1).  The
// residue of t B^L has t's limbs below position H - L at [L, H) and the
// rest wrapped to [0, 2L + 1 - H), which ends at or below L.  So: fold V
// (limbs [H, M) added at 0, one end-around carry), then unwrap to t at
// [L, 2L + 1) -- the high layout.  V == 0 (mod B^H - 1) only when t = 0,
// and then V = 0 exactly (T = 0 forces m = 0), so the fold never yields the
// all-ones form of zero.
template <int LOGM>
__device__ __forceinline__ void wrap_tail(Smem<LOGM> &sm, uint32_t (&r)[G<LOGM>::C], const uint32_t *park, int L)
{
    constexpr int C = G<LOGM>::C, M = G<LOGM>::M, H = M / 2;
    static_assert(sizeof(sm.X) / 4 >= M + M / 64, "LDS staging of M limbs");
    uint32_t *Y = sm.X;                  // the transform buffers hold M padded words
    const int p0 = tid_() * C;
    uint32_t t[C];
    load_own<C>(t, park, 0, 2 * L);
    add_vec<C>(r, t, 0u, sm.sc);         // V = T + W, within [0, M)
#pragma unroll
    for (int i = 0; i < C; i++) Y[px(p0 + i)] = r[i];
    lds_barrier();
#pragma unroll
    for (int i = 0; i < C; i++) {          // unconditional loads (indexes wrap mod M)
        const bool lo = p0 + i < H;
        const uint32_t y = Y[px((p0 + i + H) & (M - 1))];
        t[i] = lo ? y : 0u;
        if (!lo) r[i] = 0u;
    }
    lds_barrier();
    add_vec<C>(r, t, 0u, sm.sc);         // V_lo + V_hi < 2 B^H: a carry lands at H
    bool top = false;
#pragma unroll
    for (int i = 0; i < C; i++)
        if (p0 + i == H && r[i] != 0u) { top = true; r[i] = 0u; }
    if (block_any(top, sm.sc)) {         // B^H == 1: the carry re-enters at 0
#pragma unroll
        for (int i = 0; i < C; i++) t[i] = 0u;
        add_vec<C>(r, t, 1u, sm.sc);     // stays below B^H - 1
    }
#pragma unroll
    for (int i = 0; i < C; i++) Y[px(p0 + i)] = r[i];
    lds_barrier();
#pragma unroll
    for (int i = 0; i < C; i++) {
        const int pos = p0 + i;
        const uint32_t y = Y[px((pos - H) & (M - 1))];
        r[i] = pos < L ? 0u : pos < H ? r[i] : pos <= 2 * L ? y : 0u;
    }
    lds_barrier();
}

// Montgomery: r <- A * B * R^-1 mod N, output in high layout (positions
// [L, 2L]).  A is positions [off, off+L) of r; B as in tproduct.  The high
// half of T = A*B waits in the candidate's scratch area while m and m*N are
// formed, so only two limb vectors are live across a transform.
template <int LOGM>
__device__ __forceinline__ void montmul(Smem<LOGM> &sm, const Twiddles &T, const Consts &K, const Dev &D, int c,
                        uint32_t (&r)[G<LOGM>::C], int off, const uint32_t *Bf)
{
    constexpr int M = G<LOGM>::M, C = G<LOGM>::C;
    const int L = D.Ls[c];
    const int p0 = tid_() * C;
    uint32_t *park = D.scr + (size_t)c * 2 * M + M;            // 2nd half of the scratch
    uint32_t *ysave = D.ysv + (size_t)c * G<LOGM>::YSV;
    const uint32_t *Nl = D.N + (size_t)c * D.stride;
    const uint32_t *Npf = D.Npf + (size_t)c * 2 * M;
    const uint32_t *Nf = D.Nf + (size_t)c * 2 * M;
    if constexpr (G<LOGM>::WRAP) {
        // stage 0: r = T = A*B (< R^2); stage 1: r = m = T_lo N' mod R;
        // stage 2: r = W = m N mod (x^(M/2) - 1) at x = 2^24 (a wrapped
        // product: half the transform work of the full m N, whose low half
        // is known).
        const uint32_t *Nw = G<LOGM>::SPLIT ? Nf : D.Nf2 + (size_t)c * M;
#pragma unroll 1
        for (int stage = 0; stage < 3; stage++) {
            const uint32_t *B = stage == 0 ? Bf : stage == 1 ? Npf : Nw;
            tproduct<LOGM>(sm, T, K, r, stage == 0 ? off : 0, L, B, r, stage == 1 ? L : M, ysave,
                           stage == 2);
            if (stage == 0) save_own<C>(r, park, 0, 2 * L);    // T
        }
        wrap_tail<LOGM>(sm, r, park, L);
    } else {
        bool cbit = false;
        // stage 0: r = T = A*B (< R^2); stage 1: r = m = T_lo N' mod R;
        // stage 2: r = m N.  One product instance keeps the code (and its
        // register allocation) single.
#pragma unroll 1
        for (int stage = 0; stage < 3; stage++) {
            const uint32_t *B = stage == 0 ? Bf : stage == 1 ? Npf : Nf;
            tproduct<LOGM>(sm, T, K, r, stage == 0 ? off : 0, L, B, r, stage == 1 ? L : M, ysave);
            if (stage == 0) {
                bool nz = false;
#pragma unroll
                for (int i = 0; i < C; i++) nz |= (p0 + i < L) && r[i] != 0u;
                cbit = block_any(nz, sm.sc);
                save_own<C>(r, park, L, 2 * L + 1);            // T_hi
            }
        }
        zero_below<C>(r, L);
        {
            uint32_t t[C];
            load_own<C>(t, park, L, 2 * L + 1);
            // + [T_lo != 0] at position L: T_hi's lowest limb is below 2^24, so
            // adding 1 cannot overflow the 32-bit lane before add_vec normalises
#pragma unroll
            for (int i = 0; i < C; i++) t[i] += (cbit && p0 + i == L) ? 1u : 0u;
            add_vec<C>(r, t, 0u, sm.sc);                       // T_hi + (mN)_hi + c
        }
    }
    reduce_once<C>(r, Nl, L, L, sm.sc, D.Nh + (size_t)c * M);
}

// x <- x * s (plain integer product), 1 <= s < 2^24.
template <int C>
__device__ __forceinline__ void mul_small_vec(uint32_t (&x)[C], uint32_t s, Scratch &sc)
{
    const int t = tid_();
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < C; i++) {
        acc += (uint64_t)x[i] * s;
        x[i] = (uint32_t)acc & LIMB_MASK;
        acc >>= LIMB_BITS;
    }
    sc.carry[t] = acc;                  // < 2^24 + 2
    lds_barrier();
    uint64_t cin = t ? sc.carry[t - 1] : 0ull;
    lds_barrier();
#pragma unroll
    for (int i = 0; i < C; i++) {
        cin += x[i];
        x[i] = (uint32_t)cin & LIMB_MASK;
        cin >>= LIMB_BITS;
    }
    bool all1 = true;
#pragma unroll
    for (int i = 0; i < C; i++) all1 &= (x[i] == LIMB_MASK);
    uint32_t cc = block_carry_in(cin != 0ull, all1, sc);
#pragma unroll
    for (int i = 0; i < C; i++) {
        const uint32_t v = x[i] + cc;
        x[i] = v & LIMB_MASK;
        cc = v >> LIMB_BITS;
    }
}

// x <- x * 2^e through the padded LDS buffer X (MP words).
template <int C>
__device__ __forceinline__ void shl_vec(uint32_t (&x)[C], long e, uint32_t *X, int M)
{
    const int p0 = tid_() * C;
#pragma unroll
    for (int i = 0; i < C; i++) X[px(p0 + i)] = x[i];
    lds_barrier();
    const long ls = e / 24;
    const int bs = (int)(e % 24);
#pragma unroll
    for (int i = 0; i < C; i++) {
        const long k = (long)(p0 + i) - ls;
        const uint32_t hi = (k >= 0 && k < M) ? X[px((int)k)] : 0u;
        const uint32_t lo = (k - 1 >= 0 && k - 1 < M) ? X[px((int)k - 1)] : 0u;
        x[i] = bs ? (((hi << bs) | (lo >> (24 - bs))) & LIMB_MASK) : hi;
    }
    lds_barrier();
}

// x <- 5^e.
template <int LOGM>
__device__ __forceinline__ void pow5(Smem<LOGM> &sm, const Twiddles &T, const Consts &K, uint32_t (&x)[G<LOGM>::C],
                     long e, uint32_t *ysave)
{
    constexpr int M = G<LOGM>::M, C = G<LOGM>::C;
#pragma unroll
    for (int i = 0; i < C; i++) x[i] = 0u;
    if (threadIdx.x == 0) x[0] = e ? 5u : 1u;
    if (e <= 1) return;
    int top = 62 - __clzll((long long)e);      // bit below the leading one
    for (int b = top; b >= 0; b--) {
        tproduct<LOGM>(sm, T, K, x, 0, M / 2, nullptr, x, M, ysave);
        if ((e >> b) & 1) mul_small_vec<C>(x, 5u, sm.sc);
    }
}

// N = sum of coef * 10^exp over the candidate's terms; one workgroup each.
// Ls_in (optional, per launch index) sets the slot's limb count.
template <int LOGM>
__global__ __launch_bounds__(THREADS) void build_kernel(Dev D, Twiddles T, Consts K,
                                                        const int64_t *terms, int tmax,
                                                        const int32_t *Ls_in)
{
    constexpr int C = G<LOGM>::C;
    __shared__ Smem<LOGM> sm;
    const int c = slot_id(D);
    uint32_t P[C], Q[C], x[C];
#pragma unroll
    for (int i = 0; i < C; i++) P[i] = Q[i] = 0u;
    const int64_t *tc = terms + (size_t)blockIdx.x * tmax * 2;
    if (Ls_in && threadIdx.x == 0) D.Ls[c] = Ls_in[blockIdx.x];
    for (int t = 0; t < tmax; t++) {
        const int64_t coef = tc[2 * t], ex = tc[2 * t + 1];
        if (coef == 0) continue;
        pow5<LOGM>(sm, T, K, x, (long)ex, D.ysv + (size_t)c * G<LOGM>::YSV);
        const uint32_t mag = (uint32_t)(coef < 0 ? -coef : coef);
        if (mag != 1u) mul_small_vec<C>(x, mag, sm.sc);
        shl_vec<C>(x, (long)ex, sm.X, G<LOGM>::M);
        if (coef > 0) add_vec<C>(P, x, 0u, sm.sc);
        else add_vec<C>(Q, x, 0u, sm.sc);
    }
    sub_vec<C>(x, P, Q, sm.sc);
    store_vec<C>(x, D.N + (size_t)c * D.stride, D.stride, 0);
}

__global__ void fill_i32_kernel(int32_t *p, int n, int v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Admission of candidates restored from a checkpoint: after build, setup
// and a first round start, put back the base in progress and the next
// exponent bit (the residue was copied into D.res beforehand).
// info[3 i ..]: slot, base index, exponent bit.
__global__ void resume_kernel(Dev D, const int32_t *info, int n, const uint32_t *bases)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = info[3 * i], bi = info[3 * i + 1];
    D.bidx[c] = bi;
    D.base[c] = bases[bi];
    D.bitpos[c] = info[3 * i + 2];
    D.alive[c] = 1;
    D.done[c] = 0;
}

__global__ void init_alive_kernel(Dev D)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < D.n) D.alive[c] = D.pre[c] == 2 ? 0 : 1;
}

// ------------------------------------------------------------------ kernels

__global__ void twiddle_kernel(uint32_t *out, int logm, uint32_t w0, uint32_t w1)
{
    const int half = 1 << (logm - 1);
    const int M = 1 << logm;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= half) return;
    for (int q = 0; q < 2; q++) {
        const uint32_t p = prime(q);
        const uint64_t w = q ? w1 : w0;
        auto pw = [&](uint64_t e) {
            uint64_t acc = 1, b = w;
            while (e) { if (e & 1) acc = acc * b % p; b = b * b % p; e >>= 1; }
            return (uint32_t)acc;
        };
        const uint32_t f = pw((uint64_t)i), g = pw((uint64_t)(M - i) % (uint64_t)M);
        uint32_t *base = out + (size_t)q * 4 * half;
        base[i] = f;
        base[half + i] = (uint32_t)(((uint64_t)f << 32) / p);
        base[2 * half + i] = g;
        base[3 * half + i] = (uint32_t)(((uint64_t)g << 32) / p);
    }
}

// words (uint32, n_words per number) -> 24-bit limbs (stride per number).
__global__ void words_to_limbs(const uint32_t *w, int n_words, uint32_t *limbs, int stride, int L)
{
    const int c = blockIdx.y;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < stride; i += gridDim.x * blockDim.x) {
        uint32_t v = 0;
        if (i < L) {
            const long bit = (long)i * 24;
            const long wi = bit >> 5;
            const int sh = (int)(bit & 31);
            uint64_t x = 0;
            if (wi < n_words) x = w[(size_t)c * n_words + wi];
            if (wi + 1 < n_words) x |= (uint64_t)w[(size_t)c * n_words + wi + 1] << 32;
            v = (uint32_t)(x >> sh) & LIMB_MASK;
        }
        limbs[(size_t)c * stride + i] = v;
    }
}

__global__ void limbs_to_words(const uint32_t *limbs, int stride, int L, uint32_t *w, int n_words,
                               size_t w_stride)
{
    const int c = blockIdx.y;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_words; i += gridDim.x * blockDim.x) {
        const long bit = (long)i * 32;
        const long li = bit / 24;
        const int sh = (int)(bit % 24);
        uint64_t x = 0;
        for (int k = 0; k < 3; k++)
            if (li + k < L) x |= (uint64_t)limbs[(size_t)c * stride + li + k] << (24 * k);
        w[(size_t)c * w_stride + i] = (uint32_t)(x >> sh);
    }
}

// Per-candidate constants: nbits, s, NTT(N), N' (Newton), NTT(N'), R mod N,
// (N-1)R mod N.  One workgroup per candidate.
template <int LOGM>
__global__ __launch_bounds__(THREADS) void setup_kernel(Dev D, Twiddles T, Consts K)
{
    constexpr int M = G<LOGM>::M, C = G<LOGM>::C;
    __shared__ Smem<LOGM> sm;
    const int c = slot_id(D);
    const int L = D.Ls[c];
    const uint32_t *Nl = D.N + (size_t)c * D.stride;
    uint32_t nv[C], y[C], e[C];
    load_vec<C>(nv, Nl, L, 0);

    // bit length and two-adic valuation of N - 1
    {
        int top = -1, low = 1 << 30;
        const int p0 = tid_() * C;
#pragma unroll
        for (int i = 0; i < C; i++) {
            if (nv[i]) top = max(top, (p0 + i) * 24 + 31 - __clz((int)nv[i]));
            const uint32_t x = (p0 + i == 0) ? (nv[i] & ~1u) : nv[i];
            if (x) low = min(low, (p0 + i) * 24 + __ffs((int)x) - 1);
        }
        if (threadIdx.x == 0) { sm.bcast[0] = 0; sm.bcast[1] = 0x7FFFFFFF; }
        __syncthreads();
        if (top >= 0) atomicMax(&sm.bcast[0], (uint32_t)(top + 1));
        if (low < (1 << 30)) atomicMin(&sm.bcast[1], (uint32_t)low);
        __syncthreads();
        if (threadIdx.x == 0) {
            const int nb = (int)sm.bcast[0];
            D.nbits[c] = nb;
            D.sval[c] = (int32_t)sm.bcast[1];
            // 2 and 3 are prime, 0, 1 and even numbers are not: no rounds run
            const uint32_t n0 = Nl[0];
            D.pre[c] = (nb <= 2) ? ((n0 == 2u || n0 == 3u) ? 1 : 2) : ((n0 & 1u) ? 0 : 2);
            if (nb > 24 * L - 1) atomicOr(D.err, 1u);
        }
        __syncthreads();
    }
    const int nbits = (int)sm.bcast[0];
    if (D.pre[c] != 0) return;
    {   // N in the high layout, lane-interleaved, for the reductions
        uint32_t h[C];
        load_vec<C>(h, Nl, L, L);
        save_own<C>(h, D.Nh + (size_t)c * M, 0, M);
    }

    uint32_t *Nf = D.Nf + (size_t)c * 2 * M;
    uint32_t *Npf = D.Npf + (size_t)c * 2 * M;
    uint32_t *scr = D.scr + (size_t)c * 2 * M;
    uint32_t *ysv = D.ysv + (size_t)c * G<LOGM>::YSV;
    uint32_t *fsave = ysv + M;                       // split transforms' scratch
    to_fft<LOGM>(sm, T, K, nv, L, Nf, fsave);
    if constexpr (G<LOGM>::HALF) to_fft_half<LOGM>(sm, T, K, nv, L, D.Nf2 + (size_t)c * M);

    // N' = -N^-1 mod R by Newton: y <- y (2 + N y), doubling precision.
    {
        uint32_t n0 = Nl[0];
        uint32_t inv = n0;
        for (int i = 0; i < 5; i++) inv *= 2u - n0 * inv;
#pragma unroll
        for (int i = 0; i < C; i++) y[i] = 0u;
        if (threadIdx.x == 0) y[0] = (0u - inv) & LIMB_MASK;
        for (int prec = 1; prec < L; prec *= 2) {
            tproduct<LOGM>(sm, T, K, y, 0, L, Nf, e, L, ysv);     // e = N y mod R
            uint32_t two[C];
#pragma unroll
            for (int i = 0; i < C; i++) two[i] = (threadIdx.x == 0 && i == 0) ? 2u : 0u;
            add_vec<C>(e, two, 0u, sm.sc);
            {   // truncate to R
                const int p0 = tid_() * C;
#pragma unroll
                for (int i = 0; i < C; i++) if (p0 + i >= L) e[i] = 0u;
            }
            to_fft<LOGM>(sm, T, K, e, L, scr, fsave);
            tproduct<LOGM>(sm, T, K, y, 0, L, scr, y, L, ysv);    // y = y (2 + N y) mod R
        }
    }
    to_fft<LOGM>(sm, T, K, y, L, Npf, fsave);

    // R mod N: 2^nbits - N, then doubled (24 L - nbits) times mod N.
    {
        uint32_t pw[C];
        const int p0 = tid_() * C;
#pragma unroll
        for (int i = 0; i < C; i++)
            pw[i] = (p0 + i == nbits / 24) ? (1u << (nbits % 24)) : 0u;
        sub_vec<C>(y, pw, nv, sm.sc);                  // 2^nbits - N  (< N)
        for (int k = nbits; k < 24 * L; k++) {
            add_vec<C>(y, y, 0u, sm.sc);
            reduce_once<C>(y, Nl, L, 0, sm.sc);
        }
        store_vec<C>(y, D.one + (size_t)c * D.stride, D.stride, 0);
        sub_vec<C>(e, nv, y, sm.sc);                   // N - (R mod N) = (N-1) R mod N
        store_vec<C>(e, D.nm1 + (size_t)c * D.stride, D.stride, 0);
    }
}

// Start a round: r = base * R mod N, bitpos = nbits - 2 (the top bit of d
// is 1 and is consumed by the initial value).
// mode 0: every alive slot, with `base`.
// mode 1 (scheduler, after setup): alive = the number needs rounds; first
//   base of `bases`.
// mode 2 (scheduler, after round_end): next base of `bases` if the slot is
//   alive and one is left, else the verdict is final (done = 1).
// mode 3: every alive slot, with its own base bases[slot].
template <int LOGM>
__global__ __launch_bounds__(THREADS) void round_start_kernel(Dev D, uint32_t base, int mode,
                                                              const uint32_t *bases, int n_bases)
{
    constexpr int C = G<LOGM>::C;
    __shared__ Scratch sc;
    __shared__ uint32_t stash[Lds<LOGM>::MP];
    __shared__ uint32_t go;
    const int c = slot_id(D);
    if (mode == 3) {
        if (!D.alive[c] || D.pre[c]) return;
        base = bases[c];
        if (threadIdx.x == 0) D.base[c] = base;
    } else if (mode != 0) {
        if (threadIdx.x == 0) {
            int bi = 0;
            bool run;
            if (mode == 1) {
                D.alive[c] = D.pre[c] == 2 ? 0 : 1;
                run = D.pre[c] == 0 && n_bases > 0;
            } else {
                bi = D.bidx[c] + 1;
                run = D.alive[c] && bi < n_bases;
            }
            if (run) { D.bidx[c] = bi; D.base[c] = bases[bi]; }
            D.done[c] = run ? 0 : 1;
            go = run;
        }
        __syncthreads();
        if (!go) return;
        base = D.base[c];
    } else {
        if (!D.alive[c] || D.pre[c]) return;
        if (threadIdx.x == 0) D.base[c] = base;
    }
    const int L = D.Ls[c];
    uint32_t r[C];
    load_vec<C>(r, D.one + (size_t)c * D.stride, L, 0);
    if (base > 1u) mulsmall_modN<C>(r, base, D.N + (size_t)c * D.stride, L, 0, stash, sc);
    store_vec<C>(r, D.res + (size_t)c * D.stride, D.stride, 0);
    if (threadIdx.x == 0) D.bitpos[c] = D.nbits[c] - 2;
}

// Up to `steps` exponent bits: r = r^2, times base when the bit of d is set.
// Bit i of d is bit i + s of N (N odd, so N - 1 differs only in bit 0).
template <int LOGM>
__global__ __launch_bounds__(THREADS) void exp_kernel(Dev D, Twiddles T, Consts K, int steps)
{
    constexpr int C = G<LOGM>::C;
    __shared__ Smem<LOGM> sm;
    const int c = slot_id(D);
    if (D.run && !D.run[c]) return;      // (slots under admission are not touched)
    if (!D.alive[c] || D.pre[c]) return;
    const int s = D.sval[c];
    int bp = D.bitpos[c];
    if (bp < s) return;
    const uint32_t base = D.base[c];
    const int L = D.Ls[c];
    const uint32_t *Nl = D.N + (size_t)c * D.stride;
    uint32_t r[C];
    load_vec<C>(r, D.res + (size_t)c * D.stride, L, L);       // high layout
    for (int k = 0; k < steps && bp >= s; k++, bp--) {
        montmul<LOGM>(sm, T, K, D, c, r, L, nullptr);
        const uint32_t bit = (Nl[bp / 24] >> (bp % 24)) & 1u;
        if (bit) mulsmall_modN<C>(r, base, Nl, L, L, sm.X, sm.sc, D.Nh + (size_t)c * G<LOGM>::M);
    }
    store_vec<C>(r, D.res + (size_t)c * D.stride, L, L);
    if (threadIdx.x == 0) D.bitpos[c] = bp;
}

template <int C>
__device__ __forceinline__ bool vec_equal_mem(const uint32_t (&a)[C], const uint32_t *src, int len,
                                              int off, Scratch &sc)
{
    bool ne = false;
    const int p0 = tid_() * C;
#pragma unroll
    for (int i = 0; i < C; i++) {
        const int idx = p0 + i - off;
        ne |= a[i] != ((idx >= 0 && idx < len) ? src[idx] : 0u);
    }
    return !block_any(ne, sc);
}

template <int C>
__device__ __forceinline__ bool vec_equal(const uint32_t (&a)[C], const uint32_t (&b)[C], Scratch &sc)
{
    bool ne = false;
#pragma unroll
    for (int i = 0; i < C; i++) ne |= a[i] != b[i];
    return !block_any(ne, sc);
}

// End of a round: pass iff r == R or r^(2^i) == (N-1) R for some i < s, or
// r == 0, which happens only when the base is a multiple of N (N < 2^24,
// e.g. N = 5 or 7 with base 5 or 7): such a base says nothing about N and
// the round counts as passed, as mpz_probab_prime_p reports 5 and 7 prime
// (miller_rabin.c:91).  Optionally writes r out of Montgomery form (a^d mod
// N) to out.
template <int LOGM>
__global__ __launch_bounds__(THREADS) void round_end_kernel(Dev D, Twiddles T, Consts K,
                                                            uint32_t *out, size_t out_stride)
{
    constexpr int C = G<LOGM>::C;
    __shared__ Smem<LOGM> sm;
    const int c = slot_id(D);
    if (!D.alive[c] || D.pre[c]) return;
    const int L = D.Ls[c];
    const uint32_t *one = D.one + (size_t)c * D.stride;
    const uint32_t *nm1 = D.nm1 + (size_t)c * D.stride;
    uint32_t r[C];
    load_vec<C>(r, D.res + (size_t)c * D.stride, L, L);
    bool pass = vec_equal_mem<C>(r, one, L, L, sm.sc) || vec_equal_mem<C>(r, nm1, L, L, sm.sc) ||
                vec_equal_mem<C>(r, one, 0, L, sm.sc);          // r == 0: base == 0 (mod N)
    if (out) {   // a^d mod N = REDC(r): product with 1
        uint32_t *one_f = D.scr + (size_t)c * 2 * G<LOGM>::M;
        {
            uint32_t unit[C];
#pragma unroll
            for (int i = 0; i < C; i++) unit[i] = (tid_() == 0 && i == 0) ? 1u : 0u;
            to_fft<LOGM>(sm, T, K, unit, L, one_f, D.ysv + (size_t)c * G<LOGM>::YSV + G<LOGM>::M);
        }
        montmul<LOGM>(sm, T, K, D, c, r, L, one_f);
        store_vec<C>(r, out + (size_t)c * out_stride, L, L);
        load_vec<C>(r, D.res + (size_t)c * D.stride, L, L);
    }
    const int s = D.sval[c];
    for (int i = 1; i < s && !pass; i++) {
        montmul<LOGM>(sm, T, K, D, c, r, L, nullptr);
        pass = vec_equal_mem<C>(r, nm1, L, L, sm.sc);
        if (!pass && vec_equal_mem<C>(r, one, L, L, sm.sc)) break;     // reached 1: composite
    }
    if (threadIdx.x == 0 && !pass) D.alive[c] = 0;
}

// Self-check kernels: op 0 product a*b (2L limbs), op 1 a*b*R^-1 mod N,
// op 2 a*s mod N.  a, b, out: [n][stride2] limbs with stride2 >= 2L+1.
template <int LOGM>
__global__ __launch_bounds__(THREADS) void debug_kernel(Dev D, Twiddles T, Consts K, int op,
                                                        const uint32_t *A, const uint32_t *B,
                                                        uint32_t s, uint32_t *out, int stride2)
{
    constexpr int M = G<LOGM>::M, C = G<LOGM>::C;
    __shared__ Smem<LOGM> sm;
    const int c = blockIdx.x;
    const int L = D.Ls[c];
    uint32_t a[C], b[C];
    load_vec<C>(a, A + (size_t)c * stride2, L, 0);
    load_vec<C>(b, B + (size_t)c * stride2, L, 0);
    uint32_t *scr = D.scr + (size_t)c * 2 * M;
    uint32_t *fsave = D.ysv + (size_t)c * G<LOGM>::YSV + M;
    if (op == 0) {
        to_fft<LOGM>(sm, T, K, b, L, scr, fsave);
        tproduct<LOGM>(sm, T, K, a, 0, L, scr, a, M,
                       D.ysv + (size_t)c * G<LOGM>::YSV);
        store_vec<C>(a, out + (size_t)c * stride2, 2 * L, 0);
    } else if (op == 1) {
        to_fft<LOGM>(sm, T, K, b, L, scr, fsave);
        montmul<LOGM>(sm, T, K, D, c, a, 0, scr);
        store_vec<C>(a, out + (size_t)c * stride2, L, L);
    } else {
        mulsmall_modN<C>(a, s, D.N + (size_t)c * D.stride, L, 0, sm.X, sm.sc);
        store_vec<C>(a, out + (size_t)c * stride2, L, 0);
    }
}

// ---- arithmetic on caller-owned device vectors (the reference runner's
// BatchModCtx methods, batch_mod_ctx.cuh:141-161).  A vector of candidate c
// is 24-bit limbs, one per uint64_t, at V + c * ld; a residue is x R mod N
// with R = 2^(24 L_c).  Outputs are written with L_c limbs and zeros up to
// ld.
enum { AR_MUL = 0, AR_SQ = 1, AR_CHECK = 2, AR_R2 = 3, AR_TORES = 4, AR_FROMRES = 5 };

template <int C>
__device__ __forceinline__ void load_vec64(uint32_t (&v)[C], const uint64_t *src, int len, int off)
{
    const int p0 = tid_() * C;
#pragma unroll
    for (int i = 0; i < C; i++) {
        const int idx = p0 + i - off;
        v[i] = (idx >= 0 && idx < len) ? (uint32_t)src[idx] & LIMB_MASK : 0u;
    }
}

// dst[0, ld) = the positions [off, off + len) of v, then zeros.
template <int C>
__device__ __forceinline__ void store_vec64(const uint32_t (&v)[C], uint64_t *dst, int len, int off,
                                            int ld)
{
    const int p0 = tid_() * C;
#pragma unroll
    for (int i = 0; i < C; i++) {
        const int idx = p0 + i - off;
        if (idx >= 0 && idx < len) dst[idx] = v[i];
    }
    for (int idx = len + (int)threadIdx.x; idx < ld; idx += THREADS) dst[idx] = 0u;
}

// op AR_MUL: out = A B R^-1 mod N; AR_SQ: A^2 R^-1; AR_CHECK: flag = (A ==
// R mod N or A == (N-1) R mod N); AR_R2: r2 = R^2 mod N (as 2^(24 L) in
// Montgomery form: left-to-right powering of 2, squarings and doublings);
// AR_TORES: out = A R mod N (A < N, times r2); AR_FROMRES: out = A R^-1 mod
// N.  Slots decided without rounds (even N, N < 4) get zeros and flag 0.
template <int LOGM>
__global__ __launch_bounds__(THREADS) void arith_kernel(Dev D, Twiddles T, Consts K, int op,
                                                        const uint64_t *A, const uint64_t *B,
                                                        int ld, uint64_t *out, uint8_t *flag,
                                                        uint32_t *r2)
{
    constexpr int M = G<LOGM>::M, C = G<LOGM>::C;
    __shared__ Smem<LOGM> sm;
    const int c = blockIdx.x;
    const int L = D.Ls[c];
    uint32_t *scr = D.scr + (size_t)c * 2 * M;
    uint32_t *fsave = D.ysv + (size_t)c * G<LOGM>::YSV + M;
    uint32_t a[C];
    if (D.pre[c] != 0) {
#pragma unroll
        for (int i = 0; i < C; i++) a[i] = 0u;
        if (op == AR_CHECK) {
            if (threadIdx.x == 0) flag[c] = 0;
        } else if (op == AR_R2) {
            store_vec<C>(a, r2 + (size_t)c * D.stride, L, 0);
        } else {
            store_vec64<C>(a, out + (size_t)c * ld, L, 0, ld);
        }
        return;
    }
    if (op == AR_CHECK) {
        load_vec64<C>(a, A + (size_t)c * ld, L, 0);
        const bool pass = vec_equal_mem<C>(a, D.one + (size_t)c * D.stride, L, 0, sm.sc) ||
                          vec_equal_mem<C>(a, D.nm1 + (size_t)c * D.stride, L, 0, sm.sc);
        if (threadIdx.x == 0) flag[c] = pass ? 1 : 0;
        return;
    }
    if (op == AR_R2) {
        const uint32_t *Nl = D.N + (size_t)c * D.stride;
        load_vec<C>(a, D.one + (size_t)c * D.stride, L, L);   // Montgomery form of 1, high layout
        const uint32_t e = 24u * (uint32_t)L;
        for (int i = 31 - __clz((int)e); i >= 0; i--) {
            montmul<LOGM>(sm, T, K, D, c, a, L, nullptr);
            if ((e >> i) & 1u)
                mulsmall_modN<C>(a, 2u, Nl, L, L, sm.X, sm.sc, D.Nh + (size_t)c * M);
        }
        store_vec<C>(a, r2 + (size_t)c * D.stride, L, L);
        return;
    }
    if (op == AR_SQ) {
        load_vec64<C>(a, A + (size_t)c * ld, L, L);           // high layout
        montmul<LOGM>(sm, T, K, D, c, a, L, nullptr);
    } else {
        uint32_t b[C];
        load_vec64<C>(a, A + (size_t)c * ld, L, 0);
        if (op == AR_MUL) {
            load_vec64<C>(b, B + (size_t)c * ld, L, 0);
        } else if (op == AR_TORES) {
            load_vec<C>(b, r2 + (size_t)c * D.stride, L, 0);
        } else {
#pragma unroll
            for (int i = 0; i < C; i++) b[i] = (tid_() == 0 && i == 0) ? 1u : 0u;
        }
        to_fft<LOGM>(sm, T, K, b, L, scr, fsave);
        montmul<LOGM>(sm, T, K, D, c, a, 0, scr);
    }
    store_vec64<C>(a, out + (size_t)c * ld, L, L, ld);
}

} // namespace mr

// ------------------------------------------------------------------ host

namespace {

using namespace mr;

uint64_t h_pow(uint64_t b, uint64_t e, uint64_t p)
{
    uint64_t acc = 1;
    b %= p;
    while (e) {
        if (e & 1) acc = (unsigned __int128)acc * b % p;
        b = (unsigned __int128)b * b % p;
        e >>= 1;
    }
    return acc;
}

uint32_t h_shoupq(uint32_t w, uint32_t p) { return (uint32_t)(((uint64_t)w << 32) / p); }

Consts make_consts(int logm)
{
    Consts K;
    const uint64_t M = 1ull << logm;
    const uint32_t minv0 = (uint32_t)h_pow(M, P0 - 2, P0), minv1 = (uint32_t)h_pow(M, P1 - 2, P1);
    const uint32_t r32_0 = (uint32_t)((1ull << 32) % P0), r32_1 = (uint32_t)((1ull << 32) % P1);
    auto fill = [&](CrtConst &k, uint32_t s0, uint32_t s1) {
        k.s0 = s0; k.s0q = h_shoupq(s0, P0);
        k.c1 = (uint32_t)((uint64_t)s1 * P0_INV_MOD_P1 % P1); k.c1q = h_shoupq(k.c1, P1);
        k.c2 = P0_INV_MOD_P1; k.c2q = h_shoupq(k.c2, P1);
    };
    fill(K.mul, minv0, minv1);
    fill(K.mulh, (uint32_t)h_pow(M / 2, P0 - 2, P0), (uint32_t)h_pow(M / 2, P1 - 2, P1));
    fill(K.sq, (uint32_t)((uint64_t)minv0 * r32_0 % P0), (uint32_t)((uint64_t)minv1 * r32_1 % P1));
    K.r2[0] = (uint32_t)((uint64_t)r32_0 * r32_0 % P0);
    K.r2[1] = (uint32_t)((uint64_t)r32_1 * r32_1 % P1);
    K.np[0] = neg_inv32(P0);
    K.np[1] = neg_inv32(P1);
    return K;
}

constexpr int MIN_LOGM = 11, MAX_LOGM = 15;
int g_min_logm = MIN_LOGM;      // floor for new contexts (mr_set_min_log_len)
constexpr bool split_logm(int logm) { return logm > 14; }
constexpr bool half_logm(int logm) { return logm >= 12 && !split_logm(logm); }
static_assert(split_logm(14) == G<14>::SPLIT && split_logm(15) == G<15>::SPLIT &&
              half_logm(11) == G<11>::HALF && half_logm(12) == G<12>::HALF &&
              half_logm(14) == G<14>::HALF && half_logm(15) == G<15>::HALF,
              "host and device agree on the product geometry");

struct TwCache {
    uint32_t *tw[MAX_LOGM + 1] = {};
    uint32_t *pass[MAX_LOGM + 1] = {};
};
TwCache g_tw[64];

// Per-pass twiddle tables (layout: Twiddles::pass in mr_core.h), built on
// the host: [prime][direction][2M words].
std::vector<uint32_t> pass_tables(int logm)
{
    const int M = 1 << logm, LAST = (logm & 1) ? logm - 1 : logm - 2;
    std::vector<uint32_t> h((size_t)2 * 2 * 2 * M, 0u);
    std::vector<uint32_t> pw(M), pwi(M);
    for (int q = 0; q < 2; q++) {
        const uint32_t p = q ? P1 : P0;
        const uint64_t w = h_pow(q ? G1 : G0, (p - 1) >> logm, p), wi = h_pow(w, p - 2, p);
        pw[0] = pwi[0] = 1u;
        for (int i = 1; i < M; i++) {
            pw[i] = (uint32_t)(pw[i - 1] * w % p);
            pwi[i] = (uint32_t)(pwi[i - 1] * wi % p);
        }
        auto put = [&](uint32_t *blk, int H, int pos, uint32_t a, uint32_t b, uint32_t c) {
            blk[pos] = a; blk[H + pos] = h_shoupq(a, p);
            blk[2 * H + pos] = b; blk[3 * H + pos] = h_shoupq(b, p);
            blk[4 * H + pos] = c; blk[5 * H + pos] = h_shoupq(c, p);
        };
        uint32_t *fwd = h.data() + (size_t)(2 * q) * 2 * M, *inv = fwd + 2 * M;
        for (int S = 0; S < LAST; S += 2) {
            const int H2 = M >> (S + 2);
            for (int pos = 0; pos < H2; pos++)
                put(fwd + fwd_off(logm, S), H2, pos, pw[pos << S], pw[(pos + H2) << S], pw[pos << (S + 1)]);
        }
        for (int T0 = inv_first(logm); T0 + 1 < logm; T0 += 2) {
            const int H = 1 << T0;
            for (int pos = 0; pos < H; pos++)
                put(inv + inv_off(logm, T0), H, pos, pwi[pos << (logm - 1 - T0)], pwi[pos << (logm - 2 - T0)],
                    pwi[(pos + H) << (logm - 2 - T0)]);
        }
    }
    return h;
}

int get_twiddles(int dev, int logm, hipStream_t st, uint32_t **out, uint32_t **pass_out)
{
    if (dev < 0 || dev >= 64) return -1;
    uint32_t *&t = g_tw[dev].tw[logm];
    uint32_t *&pt = g_tw[dev].pass[logm];
    if (!t) {
        const int half = 1 << (logm - 1);
        if (hipMalloc(&t, (size_t)2 * 4 * half * sizeof(uint32_t)) != hipSuccess) { t = nullptr; return -1; }
        const uint32_t w0 = (uint32_t)h_pow(G0, (P0 - 1) >> logm, P0);
        const uint32_t w1 = (uint32_t)h_pow(G1, (P1 - 1) >> logm, P1);
        twiddle_kernel<<<(half + 255) / 256, 256, 0, st>>>(t, logm, w0, w1);
        if (hipStreamSynchronize(st) != hipSuccess) return -1;
    }
    if (!pt) {
        const std::vector<uint32_t> h = pass_tables(logm);
        if (hipMalloc(&pt, h.size() * sizeof(uint32_t)) != hipSuccess) { pt = nullptr; return -1; }
        if (hipMemcpy(pt, h.data(), h.size() * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess) return -1;
    }
    *out = t;
    *pass_out = pt;
    return 0;
}

} // namespace

struct mr_ctx {
    int device = 0;
    int n = 0, n_words = 0, L = 0, logm = 0, stride = 0;
    hipStream_t st = nullptr;
    Dev d{};
    Consts K{};
    Twiddles T{};
    std::vector<void *> allocs;
    std::vector<int32_t> sval;
};

namespace {

const char *kWho = "mr_gpu";

int hip_fail(hipError_t e, const char *what)
{
    fprintf(stderr, "%s: HIP error %s (%s)\n", kWho, hipGetErrorName(e), what);
    return -2;
}

#define MR_CHECK(call)                                        \
    do {                                                      \
        hipError_t e_ = (call);                               \
        if (e_ != hipSuccess) return hip_fail(e_, #call);     \
    } while (0)

template <class F>
int dispatch_logm(int logm, F f)
{
    switch (logm) {
    case 11: return f(std::integral_constant<int, 11>{});
    case 12: return f(std::integral_constant<int, 12>{});
    case 13: return f(std::integral_constant<int, 13>{});
    case 14: return f(std::integral_constant<int, 14>{});
    case 15: return f(std::integral_constant<int, 15>{});
    default: return -1;
    }
}

int ctx_alloc(mr_ctx *x, void **p, size_t bytes)
{
    MR_CHECK(hipMalloc(p, bytes ? bytes : 4));
    x->allocs.push_back(*p);
    MR_CHECK(hipMemsetAsync(*p, 0, bytes ? bytes : 4, x->st));
    return 0;
}

constexpr int TMAX = 24;       // sparse decimal terms per candidate

// Limbs and log2 transform length for numbers of at most `bits` bits:
// 24 L > bits, so that 2N < R = 2^(24 L).
void geometry(long bits, int min_logm, int *L_out, int *logm_out)
{
    int L = (int)(bits / 24 + 1);
    if (L < 4) L = 4;
    int logm = min_logm;
    while ((1 << logm) < 2 * L + 2) logm++;
    *L_out = L;
    *logm_out = logm;
}

// Upper bound on the bit length of a number of `digits` decimal digits.
long digits_to_bits(int64_t digits) { return (long)((double)digits * 3.3219280948873623) + 2; }

// Geometry and device arrays for numbers of at most `bits` bits.
int ctx_alloc_all(mr_ctx *x, long bits, int min_logm)
{
    int L, logm;
    geometry(bits, min_logm, &L, &logm);
    if (logm > MAX_LOGM || L > MAX_LIMBS) {
        fprintf(stderr, "%s: numbers of %ld bits are beyond this build's transform sizes\n", kWho, bits);
        return -1;
    }
    x->L = L;
    x->logm = logm;
    x->stride = (1 << logm);   // room for 2L+1 limbs in debug products
    const int M = 1 << logm;
    const size_t n = (size_t)x->n;
    Dev &d = x->d;
    d.L = L;
    d.stride = x->stride;
    d.n = x->n;
    void *p;
#define MR_ALLOC(field, type, count)                                   \
    if (ctx_alloc(x, &p, sizeof(type) * (size_t)(count)) != 0) return -2; \
    d.field = (type *)p;
    MR_ALLOC(N, uint32_t, n * x->stride);
    MR_ALLOC(Nh, uint32_t, n * M);
    MR_ALLOC(Nf, uint32_t, n * 2 * M);
    MR_ALLOC(Npf, uint32_t, n * 2 * M);
    MR_ALLOC(Nf2, uint32_t, half_logm(logm) ? n * M : 1);
    MR_ALLOC(one, uint32_t, n * x->stride);
    MR_ALLOC(nm1, uint32_t, n * x->stride);
    MR_ALLOC(res, uint32_t, n * x->stride);
    MR_ALLOC(scr, uint32_t, n * 2 * M);
    MR_ALLOC(ysv, uint32_t, split_logm(logm) ? n * (M + M / 2) : 1);
    MR_ALLOC(nbits, int32_t, n);
    MR_ALLOC(sval, int32_t, n);
    MR_ALLOC(bitpos, int32_t, n);
    MR_ALLOC(alive, uint8_t, n);
    MR_ALLOC(pre, uint8_t, n);
    MR_ALLOC(err, uint32_t, 1);
    MR_ALLOC(Ls, int32_t, n);
    MR_ALLOC(base, uint32_t, n);
    MR_ALLOC(bidx, int32_t, n);
    MR_ALLOC(done, uint8_t, n);
#undef MR_ALLOC
    d.list = nullptr;
    fill_i32_kernel<<<(x->n + 255) / 256, 256, 0, x->st>>>(d.Ls, x->n, L);
    // split products (2^15) transform halves of 2^14 and twist with the
    // 2^15 table
    const int tl = split_logm(logm) ? 14 : logm;
    uint32_t *tw = nullptr, *ptw = nullptr, *btw = nullptr, *bptw = nullptr;
    if (get_twiddles(x->device, tl, x->st, &tw, &ptw) != 0 ||
        get_twiddles(x->device, logm, x->st, &btw, &bptw) != 0)
        return hip_fail(hipErrorUnknown, "twiddles");
    x->T.base = tw;
    x->T.pass = ptw;
    x->T.big = btw;
    x->T.big_half = M / 2;
    x->T.half = (1 << tl) / 2;
    x->T.hbase = tw;
    x->T.hpass = ptw;
    x->T.hhalf = x->T.half;
    if (half_logm(logm)) {      // wrapped products of M/2 points
        if (get_twiddles(x->device, logm - 1, x->st, &tw, &ptw) != 0)
            return hip_fail(hipErrorUnknown, "twiddles");
        x->T.hbase = tw;
        x->T.hpass = ptw;
        x->T.hhalf = M / 4;
    }
    x->K = make_consts(logm);
    return 0;
}

// Per-candidate constants once the moduli are resident in d.N.
int ctx_finish(mr_ctx *x)
{
    const size_t n = (size_t)x->n;
    int rc = dispatch_logm(x->logm, [&](auto lg) {
        constexpr int LG = decltype(lg)::value;
        setup_kernel<LG><<<x->n, THREADS, 0, x->st>>>(x->d, x->T, x->K);
        return 0;
    });
    if (rc) return rc;
    MR_CHECK(hipGetLastError());
    uint32_t err = 0;
    MR_CHECK(hipMemcpyAsync(&err, x->d.err, 4, hipMemcpyDeviceToHost, x->st));
    x->sval.resize(n);
    MR_CHECK(hipMemcpyAsync(x->sval.data(), x->d.sval, 4 * n, hipMemcpyDeviceToHost, x->st));
    MR_CHECK(hipStreamSynchronize(x->st));
    if (err) {
        fprintf(stderr, "%s: a modulus does not fit the context\n", kWho);
        return -1;
    }
    return 0;
}

int ctx_init(mr_ctx *x, const uint32_t *h_words)
{
    int rc = ctx_alloc_all(x, (long)x->n_words * 32, g_min_logm);
    if (rc) return rc;
    uint32_t *dw = nullptr;
    const size_t n = (size_t)x->n;
    if (ctx_alloc(x, (void **)&dw, sizeof(uint32_t) * n * x->n_words) != 0) return -2;
    MR_CHECK(hipMemcpyAsync(dw, h_words, sizeof(uint32_t) * n * x->n_words, hipMemcpyHostToDevice, x->st));
    words_to_limbs<<<dim3(8, x->n), 256, 0, x->st>>>(dw, x->n_words, x->d.N, x->stride, x->L);
    return ctx_finish(x);
}

mr_ctx *ctx_new(int n_words, int n_batch, int device)
{
    if (n_words < 1 || n_batch < 1) return nullptr;
    if (device < 0 && hipGetDevice(&device) != hipSuccess) return nullptr;
    mr_ctx *x = new (std::nothrow) mr_ctx;
    if (!x) return nullptr;
    x->device = device;
    x->n = n_batch;
    x->n_words = n_words;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&x->st, hipStreamNonBlocking) != hipSuccess) {
        fprintf(stderr, "%s: cannot use HIP device %d\n", kWho, device);
        delete x;
        return nullptr;
    }
    return x;
}

} // namespace

extern "C" void mr_ctx_destroy(mr_ctx_t *x)
{
    if (!x) return;
    (void)hipSetDevice(x->device);
    if (x->st) (void)hipStreamSynchronize(x->st);
    for (void *p : x->allocs) (void)hipFree(p);
    if (x->st) (void)hipStreamDestroy(x->st);
    delete x;
}

extern "C" mr_ctx_t *mr_ctx_create(const uint32_t *N_words, int n_words, int n_batch, int device)
{
    if (!N_words) return nullptr;
    mr_ctx *x = ctx_new(n_words, n_batch, device);
    if (!x) return nullptr;
    if (ctx_init(x, N_words) != 0) {
        mr_ctx_destroy(x);
        return nullptr;
    }
    return x;
}

extern "C" mr_ctx_t *mr_ctx_create_limbs16(const uint64_t *N_all, int n_limbs16, int n_batch, int device)
{
    if (!N_all || n_limbs16 < 1 || n_batch < 1) return nullptr;
    const int n_words = (n_limbs16 + 1) / 2;
    std::vector<uint32_t> w((size_t)n_batch * n_words, 0u);
    for (int c = 0; c < n_batch; c++)
        for (int i = 0; i < n_limbs16; i++) {
            const uint64_t v = N_all[(size_t)c * n_limbs16 + i];
            if (v > 0xFFFFu) return nullptr;
            w[(size_t)c * n_words + i / 2] |= (uint32_t)v << (16 * (i & 1));
        }
    return mr_ctx_create(w.data(), n_words, n_batch, device);
}

extern "C" int mr_ctx_info(const mr_ctx_t *x, int *n_words, int *limbs, int *log_len, int *s_out)
{
    if (!x) return -1;
    if (n_words) *n_words = x->n_words;
    if (limbs) *limbs = x->L;
    if (log_len) *log_len = x->logm;
    if (s_out) memcpy(s_out, x->sval.data(), sizeof(int32_t) * x->n);
    return 0;
}

extern "C" int mr_ctx_moduli(const mr_ctx_t *x, uint32_t *out)
{
    if (!x || !out) return -1;
    (void)hipSetDevice(x->device);
    uint32_t *dw = nullptr;
    const size_t bytes = sizeof(uint32_t) * (size_t)x->n * x->n_words;
    MR_CHECK(hipMalloc(&dw, bytes));
    limbs_to_words<<<dim3(8, x->n), 256, 0, x->st>>>(x->d.N, x->stride, x->L, dw, x->n_words, x->n_words);
    hipError_t e = hipMemcpyAsync(out, dw, bytes, hipMemcpyDeviceToHost, x->st);
    if (e == hipSuccess) e = hipStreamSynchronize(x->st);
    (void)hipFree(dw);
    return e == hipSuccess ? 0 : hip_fail(e, "mr_ctx_moduli");
}

namespace {

// Rounds of every base in order over the whole context.  all_bases = false:
// a candidate retires at its first failing base (its later rounds are not
// run); true: every base is run on every candidate, as the reference runner
// does for each witness (miller_rabin_runner.cu:305-354), and alive_out is
// the AND over the bases.  pass_out (nullable, [n][n_bases]): 1 where the
// round of that base passed (retiring: where the candidate was still alive
// after it).
// per_slot: one round in which slot c uses base bases[c] (n_bases == n).
int run_rounds(mr_ctx *x, const uint32_t *bases, int n_bases, bool all_bases,
               uint8_t *alive_out, uint8_t *pass_out, uint32_t *residues_out,
               bool per_slot = false)
{
    if (!x || !bases || n_bases < 0 || !alive_out) return -1;
    if (per_slot && n_bases != x->n) return -1;
    for (int i = 0; i < n_bases; i++)
        if (bases[i] < 2u || bases[i] >= (1u << 24)) return -1;
    MR_CHECK(hipSetDevice(x->device));
    const int n = x->n;
    uint32_t *dbases = nullptr;
    if (per_slot) {
        MR_CHECK(hipMalloc(&dbases, sizeof(uint32_t) * (size_t)n));
        MR_CHECK(hipMemcpyAsync(dbases, bases, sizeof(uint32_t) * (size_t)n,
                                hipMemcpyHostToDevice, x->st));
        n_bases = 1;     // one round
    }
    init_alive_kernel<<<(n + 255) / 256, 256, 0, x->st>>>(x->d);
    uint32_t *dres = nullptr;
    uint32_t *dlimbs = nullptr;
    uint8_t *dpass = nullptr;
    const size_t rbytes = sizeof(uint32_t) * (size_t)n * n_bases * x->n_words;
    if (residues_out) {
        MR_CHECK(hipMalloc(&dres, rbytes ? rbytes : 4));
        MR_CHECK(hipMemsetAsync(dres, 0, rbytes ? rbytes : 4, x->st));
        MR_CHECK(hipMalloc(&dlimbs, sizeof(uint32_t) * (size_t)n * x->stride));
    }
    if (pass_out || all_bases) MR_CHECK(hipMalloc(&dpass, (size_t)n * (n_bases ? n_bases : 1)));
    int max_bits = 0;
    {
        std::vector<int32_t> nb(n);
        MR_CHECK(hipMemcpyAsync(nb.data(), x->d.nbits, 4 * n, hipMemcpyDeviceToHost, x->st));
        MR_CHECK(hipStreamSynchronize(x->st));
        for (int v : nb) max_bits = v > max_bits ? v : max_bits;
    }
    const int CHUNK = 256;
    int rc = 0;
    for (int bi = 0; bi < n_bases && rc == 0; bi++) {
        const uint32_t base = bases[bi];
        if (all_bases && bi > 0) init_alive_kernel<<<(n + 255) / 256, 256, 0, x->st>>>(x->d);
        rc = dispatch_logm(x->logm, [&](auto lg) {
            constexpr int LG = decltype(lg)::value;
            if (per_slot)
                round_start_kernel<LG><<<n, THREADS, 0, x->st>>>(x->d, 0u, 3, dbases, n);
            else
                round_start_kernel<LG><<<n, THREADS, 0, x->st>>>(x->d, base, 0, nullptr, 0);
            for (int done = 0; done < max_bits; done += CHUNK)
                exp_kernel<LG><<<n, THREADS, 0, x->st>>>(x->d, x->T, x->K, CHUNK);
            if (residues_out) {
                MR_CHECK(hipMemsetAsync(dlimbs, 0, sizeof(uint32_t) * (size_t)n * x->stride, x->st));
                round_end_kernel<LG><<<n, THREADS, 0, x->st>>>(x->d, x->T, x->K, dlimbs, x->stride);
                limbs_to_words<<<dim3(8, n), 256, 0, x->st>>>(
                    dlimbs, x->stride, x->L, dres + (size_t)bi * x->n_words, x->n_words,
                    (size_t)n_bases * x->n_words);
            } else {
                round_end_kernel<LG><<<n, THREADS, 0, x->st>>>(x->d, x->T, x->K, nullptr, 0);
            }
            if (dpass)
                MR_CHECK(hipMemcpyAsync(dpass + (size_t)bi * n, x->d.alive, n,
                                        hipMemcpyDeviceToDevice, x->st));
            return 0;
        });
        if (rc == 0) {
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) rc = hip_fail(e, "mr_run launch");
        }
    }
    std::vector<uint8_t> hpass;
    if (rc == 0) {
        hipError_t e = hipMemcpyAsync(alive_out, x->d.alive, n, hipMemcpyDeviceToHost, x->st);
        if (e == hipSuccess && residues_out)
            e = hipMemcpyAsync(residues_out, dres, rbytes, hipMemcpyDeviceToHost, x->st);
        if (e == hipSuccess && dpass && n_bases > 0) {
            hpass.resize((size_t)n * n_bases);
            e = hipMemcpyAsync(hpass.data(), dpass, hpass.size(), hipMemcpyDeviceToHost, x->st);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(x->st);
        if (e != hipSuccess) rc = hip_fail(e, "mr_run");
    }
    if (rc == 0 && !hpass.empty()) {
        for (int c = 0; c < n; c++) {
            bool all = true;
            for (int bi = 0; bi < n_bases; bi++) {
                const uint8_t v = hpass[(size_t)bi * n + c];
                all = all && v;
                if (pass_out) pass_out[(size_t)c * n_bases + bi] = v ? 1 : 0;
            }
            if (all_bases) alive_out[c] = all ? 1 : 0;
        }
    }
    if (dres) (void)hipFree(dres);
    if (dlimbs) (void)hipFree(dlimbs);
    if (dpass) (void)hipFree(dpass);
    if (dbases) (void)hipFree(dbases);
    return rc;
}

} // namespace

extern "C" int mr_run(mr_ctx_t *x, const uint32_t *bases, int n_bases, uint8_t *alive_out,
                      uint32_t *residues_out)
{
    return run_rounds(x, bases, n_bases, false, alive_out, nullptr, residues_out);
}

extern "C" int mr_run_per_slot(mr_ctx_t *x, const uint32_t *bases, uint8_t *pass_out,
                               uint32_t *residues_out)
{
    if (!x || !pass_out) return -1;
    return run_rounds(x, bases, x->n, false, pass_out, nullptr, residues_out, true);
}

extern "C" int mr_run_ex(mr_ctx_t *x, const uint32_t *bases, int n_bases, int flags,
                         uint8_t *alive_out, uint8_t *pass_out, uint32_t *residues_out)
{
    if (flags & ~MR_RUN_ALL_BASES) return -1;
    return run_rounds(x, bases, n_bases, (flags & MR_RUN_ALL_BASES) != 0, alive_out, pass_out,
                      residues_out);
}

namespace {

int debug_op(mr_ctx *x, int op, const uint32_t *a, const uint32_t *b, uint32_t s, uint32_t *out)
{
    if (!x || !a || !out || (op != 2 && !b)) return -1;
    MR_CHECK(hipSetDevice(x->device));
    const int n = x->n, nw = x->n_words;
    const size_t wb = sizeof(uint32_t) * (size_t)n * nw;
    const int out_words = op == 0 ? 2 * nw : nw;
    uint32_t *da = nullptr, *db = nullptr, *la = nullptr, *lb = nullptr, *lo = nullptr, *dout = nullptr;
    const size_t lb_bytes = sizeof(uint32_t) * (size_t)n * x->stride;
    int rc = 0;
    hipError_t e = hipSuccess;
    if ((e = hipMalloc(&da, wb)) != hipSuccess || (e = hipMalloc(&db, wb)) != hipSuccess ||
        (e = hipMalloc(&la, lb_bytes)) != hipSuccess || (e = hipMalloc(&lb, lb_bytes)) != hipSuccess ||
        (e = hipMalloc(&lo, lb_bytes)) != hipSuccess ||
        (e = hipMalloc(&dout, sizeof(uint32_t) * (size_t)n * out_words)) != hipSuccess) {
        rc = hip_fail(e, "debug alloc");
    }
    if (rc == 0) {
        (void)hipMemcpyAsync(da, a, wb, hipMemcpyHostToDevice, x->st);
        (void)hipMemcpyAsync(db, op == 2 ? a : b, wb, hipMemcpyHostToDevice, x->st);
        (void)hipMemsetAsync(lo, 0, lb_bytes, x->st);
        words_to_limbs<<<dim3(8, n), 256, 0, x->st>>>(da, nw, la, x->stride, x->L);
        words_to_limbs<<<dim3(8, n), 256, 0, x->st>>>(db, nw, lb, x->stride, x->L);
        rc = dispatch_logm(x->logm, [&](auto lg) {
            constexpr int LG = decltype(lg)::value;
            debug_kernel<LG><<<n, THREADS, 0, x->st>>>(x->d, x->T, x->K, op, la, lb, s, lo, x->stride);
            return 0;
        });
        limbs_to_words<<<dim3(8, n), 256, 0, x->st>>>(lo, x->stride, op == 0 ? 2 * x->L : x->L,
                                                        dout, out_words, out_words);
        e = hipMemcpyAsync(out, dout, sizeof(uint32_t) * (size_t)n * out_words, hipMemcpyDeviceToHost, x->st);
        if (e == hipSuccess) e = hipStreamSynchronize(x->st);
        if (e != hipSuccess) rc = hip_fail(e, "debug op");
    }
    (void)hipFree(da); (void)hipFree(db); (void)hipFree(la); (void)hipFree(lb); (void)hipFree(lo);
    (void)hipFree(dout);
    return rc;
}

} // namespace

extern "C" int mr_debug_product(mr_ctx_t *x, const uint32_t *a, const uint32_t *b, uint32_t *out)
{
    return debug_op(x, 0, a, b, 0, out);
}

extern "C" int mr_debug_montmul(mr_ctx_t *x, const uint32_t *a, const uint32_t *b, uint32_t *out)
{
    return debug_op(x, 1, a, b, 0, out);
}

extern "C" int mr_debug_mulsmall(mr_ctx_t *x, const uint32_t *a, uint32_t s, uint32_t *out)
{
    if (s < 1u || s >= (1u << 24)) return -1;
    return debug_op(x, 2, a, nullptr, s, out);
}

namespace {

int arith_launch(mr_ctx *x, int op, const uint64_t *A, const uint64_t *B, int ld, uint64_t *out,
                 uint8_t *flag, uint32_t *r2, hipStream_t st)
{
    MR_CHECK(hipSetDevice(x->device));
    const int rc = dispatch_logm(x->logm, [&](auto lg) {
        constexpr int LG = decltype(lg)::value;
        arith_kernel<LG><<<x->n, THREADS, 0, st>>>(x->d, x->T, x->K, op, A, B, ld, out, flag, r2);
        return 0;
    });
    if (rc) return rc;
    MR_CHECK(hipGetLastError());
    return 0;
}

// 16-bit limbs (one per uint64_t) <-> 24-bit limbs (one per uint64_t).
void limbs16_to_24(const uint64_t *x16, int n16, uint64_t *x24, int n24)
{
    for (int i = 0; i < n24; i++) x24[i] = 0u;
    for (int i = 0; i < n16; i++) {
        const uint64_t v = x16[i] & 0xFFFFu;
        const long bit = 16L * i;
        const long w = bit / 24, sh = bit % 24;
        if (w < n24) x24[w] |= (v << sh) & 0xFFFFFFu;
        if (sh > 8 && w + 1 < n24) x24[w + 1] |= v >> (24 - sh);
    }
}

void limbs24_to_16(const uint64_t *x24, int n24, uint64_t *x16, int n16)
{
    for (int i = 0; i < n16; i++) x16[i] = 0u;
    for (int i = 0; i < n24; i++) {
        const uint64_t v = x24[i] & 0xFFFFFFu;
        const long bit = 24L * i;
        for (long b = 0; b < 24; b += 8) {          // byte by byte
            const long at = bit + b;
            const long w = at / 16, sh = at % 16;
            if (w < n16) x16[w] |= ((v >> b) & 0xFFu) << sh;
        }
    }
}

// a < b for little-endian 24-bit limb vectors of length n.
bool limbs_less(const uint64_t *a, const uint32_t *b, int n)
{
    for (int i = n - 1; i >= 0; i--) {
        if (a[i] != b[i]) return a[i] < b[i];
    }
    return false;
}

} // namespace

extern "C" int mr_ctx_modmul_dev(mr_ctx_t *x, const uint64_t *d_a, const uint64_t *d_b,
                                 uint64_t *d_out, int ld, void *stream)
{
    if (!x || !d_a || !d_b || !d_out || ld < x->L) return -1;
    return arith_launch(x, AR_MUL, d_a, d_b, ld, d_out, nullptr, nullptr, (hipStream_t)stream);
}

extern "C" int mr_ctx_modsq_dev(mr_ctx_t *x, const uint64_t *d_a, uint64_t *d_out, int ld,
                                void *stream)
{
    if (!x || !d_a || !d_out || ld < x->L) return -1;
    return arith_launch(x, AR_SQ, d_a, nullptr, ld, d_out, nullptr, nullptr, (hipStream_t)stream);
}

extern "C" int mr_ctx_check_passed_dev(mr_ctx_t *x, const uint64_t *d_r, int ld, uint8_t *d_passed,
                                       void *stream)
{
    if (!x || !d_r || !d_passed || ld < x->L) return -1;
    return arith_launch(x, AR_CHECK, d_r, nullptr, ld, nullptr, d_passed, nullptr,
                        (hipStream_t)stream);
}

extern "C" int mr_ctx_to_residue(mr_ctx_t *x, const uint64_t *x16, int n_limbs16, uint64_t *out,
                                 int ld)
{
    if (!x || !x16 || !out || n_limbs16 < 1 || ld < x->L) return -1;
    MR_CHECK(hipSetDevice(x->device));
    const int n = x->n, L = x->L;
    std::vector<uint64_t> h((size_t)n * ld, 0u);
    std::vector<uint32_t> N((size_t)n * x->stride);
    MR_CHECK(hipMemcpyAsync(N.data(), x->d.N, sizeof(uint32_t) * N.size(), hipMemcpyDeviceToHost, x->st));
    MR_CHECK(hipStreamSynchronize(x->st));
    for (int c = 0; c < n; c++) {
        limbs16_to_24(x16 + (size_t)c * n_limbs16, n_limbs16, h.data() + (size_t)c * ld, ld);
        for (int i = L; i < ld; i++)
            if (h[(size_t)c * ld + i]) {
                fprintf(stderr, "%s: to_residue: value %d is not below its modulus\n", kWho, c);
                return -1;
            }
        if (!limbs_less(h.data() + (size_t)c * ld, N.data() + (size_t)c * x->stride, L)) {
            fprintf(stderr, "%s: to_residue: value %d is not below its modulus\n", kWho, c);
            return -1;
        }
    }
    uint64_t *d_in = nullptr, *d_out = nullptr;
    uint32_t *d_r2 = nullptr;
    int rc = 0;
    hipError_t e;
    if ((e = hipMalloc(&d_in, sizeof(uint64_t) * h.size())) != hipSuccess ||
        (e = hipMalloc(&d_out, sizeof(uint64_t) * h.size())) != hipSuccess ||
        (e = hipMalloc(&d_r2, sizeof(uint32_t) * (size_t)n * x->stride)) != hipSuccess)
        rc = hip_fail(e, "to_residue alloc");
    if (rc == 0 && (e = hipMemcpyAsync(d_in, h.data(), sizeof(uint64_t) * h.size(),
                                       hipMemcpyHostToDevice, x->st)) != hipSuccess)
        rc = hip_fail(e, "to_residue upload");
    if (rc == 0) rc = arith_launch(x, AR_R2, nullptr, nullptr, ld, nullptr, nullptr, d_r2, x->st);
    if (rc == 0) rc = arith_launch(x, AR_TORES, d_in, nullptr, ld, d_out, nullptr, d_r2, x->st);
    if (rc == 0 && ((e = hipMemcpyAsync(out, d_out, sizeof(uint64_t) * h.size(),
                                        hipMemcpyDeviceToHost, x->st)) != hipSuccess ||
                    (e = hipStreamSynchronize(x->st)) != hipSuccess))
        rc = hip_fail(e, "to_residue");
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    (void)hipFree(d_r2);
    return rc;
}

extern "C" int mr_ctx_from_residue_dev(mr_ctx_t *x, const uint64_t *d_x, int ld, uint64_t *out16,
                                       int n_limbs16)
{
    if (!x || !d_x || !out16 || n_limbs16 < 1 || ld < x->L) return -1;
    MR_CHECK(hipSetDevice(x->device));
    const int n = x->n;
    std::vector<uint64_t> h((size_t)n * ld);
    uint64_t *d_out = nullptr;
    hipError_t e = hipMalloc(&d_out, sizeof(uint64_t) * h.size());
    if (e != hipSuccess) return hip_fail(e, "from_residue alloc");
    int rc = arith_launch(x, AR_FROMRES, d_x, nullptr, ld, d_out, nullptr, nullptr, x->st);
    if (rc == 0 && ((e = hipMemcpyAsync(h.data(), d_out, sizeof(uint64_t) * h.size(),
                                        hipMemcpyDeviceToHost, x->st)) != hipSuccess ||
                    (e = hipStreamSynchronize(x->st)) != hipSuccess))
        rc = hip_fail(e, "from_residue");
    (void)hipFree(d_out);
    if (rc == 0)
        for (int c = 0; c < n; c++)
            limbs24_to_16(h.data() + (size_t)c * ld, ld, out16 + (size_t)c * n_limbs16, n_limbs16);
    return rc;
}

namespace {

__host__ __device__ int ndig_u(uint64_t v)
{
    int n = 1;
    while (v >= 10u) { v /= 10u; n++; }
    return n;
}

__host__ __device__ uint64_t rev_u(uint64_t v, int len)
{
    uint64_t r = 0;
    for (int i = 0; i < len; i++) { r = r * 10u + v % 10u; v /= 10u; }
    return r;
}

// value = sum coef * 10^exp (exponents distinct) as TMAX (coef, exp) pairs
// in terms[2 TMAX] (unused pairs zero), and the decimal length of the value
// in *digits.  False for an invalid candidate.  The same code builds the
// terms on the host (candidates submitted from the host) and on the device
// (survivor lists submitted in place, surv_terms_kernel).
__host__ __device__ bool cand_terms(int64_t a, int64_t b, int64_t c, uint64_t k, uint64_t j,
                                    int which, int64_t *terms, int64_t *digits)
{
    if (a < 1 || b < 0 || c < 0 || k < 1 || j < 1 || k > 1000000u || j > 1000000u) return false;
    int64_t raw[2 * TMAX];
    int nr = 0;
    if (which == 2) {
        const int lk = ndig_u(k), lj = ndig_u(j);
        const int64_t eb = a - b - lk, ec = a - c - lj;
        if (eb < 0 || ec < 0 || a >= 2000000) return false;
        const int64_t r4[8] = {1, a, -(int64_t)rev_u(k, lk), eb, -(int64_t)rev_u(j, lj), ec, -1, 0};
        for (int i = 0; i < 8; i++) raw[i] = r4[i];
        nr = 4;
        *digits = a;
    } else if (which == 0 || which == 1) {
        // S = k 10^b + j 10^c as sparse decimal digits; N = (10^a - 1) - S
        // has digits 9 - s_i with no borrow, when S < 10^a - 1.
        int64_t sp[20];                 // positions of S's nonzero digits
        int sd[20];                     // and the digits (at most 19: S < 10^19 below)
        int ns = 0;
        const int64_t lo = b < c ? b : c, hi = b < c ? c : b;
        const uint64_t clo = b < c ? k : j, chi = b < c ? j : k;
        if (hi - lo <= 12) {
            uint64_t x = chi;
            for (int64_t i = 0; i < hi - lo; i++) x *= 10u;
            x += clo;
            for (int64_t q = lo; x; q++, x /= 10u)
                if (x % 10u) { sp[ns] = q; sd[ns] = (int)(x % 10u); ns++; }
        } else {
            uint64_t x = clo;
            for (int64_t q = lo; x; q++, x /= 10u)
                if (x % 10u) { sp[ns] = q; sd[ns] = (int)(x % 10u); ns++; }
            x = chi;
            for (int64_t q = hi; x; q++, x /= 10u)
                if (x % 10u) { sp[ns] = q; sd[ns] = (int)(x % 10u); ns++; }
        }
        for (int i = 0; i < ns; i++) if (sp[i] >= a) return false;          // S >= 10^a
        // decimal length D of N: drop leading digits 9 - 9 = 0
        int64_t D = a;
        for (;;) {
            if (D == 0) return false;
            bool nine = false;
            for (int i = 0; i < ns; i++) if (sp[i] == D - 1 && sd[i] == 9) nine = true;
            if (!nine) break;
            D--;
        }
        if (a >= 2000000) return false;
        raw[0] = 1;
        raw[1] = which == 0 ? a : D;
        raw[2] = -1;
        raw[3] = 0;
        nr = 2;
        for (int i = 0; i < ns; i++) {
            if (which == 0) {
                raw[2 * nr] = -(int64_t)sd[i];
                raw[2 * nr + 1] = sp[i];
                nr++;
            } else if (sp[i] < D) {
                raw[2 * nr] = -(int64_t)sd[i];
                raw[2 * nr + 1] = D - 1 - sp[i];
                nr++;
            }
        }
        *digits = which == 0 ? a : D;
    } else {
        return false;
    }
    int nt = 0;
    for (int r = 0; r < nr; r++) {
        bool merged = false;
        for (int t = 0; t < nt; t++)
            if (terms[2 * t + 1] == raw[2 * r + 1]) { terms[2 * t] += raw[2 * r]; merged = true; }
        if (merged) continue;
        if (nt == TMAX) return false;
        terms[2 * nt] = raw[2 * r];
        terms[2 * nt + 1] = raw[2 * r + 1];
        nt++;
    }
    for (int t = nt; t < TMAX; t++) terms[2 * t] = terms[2 * t + 1] = 0;
    return true;
}

// Where an admitted job's terms come from: the host (list == nullptr, terms
// copied with the admission) or entry idx of a device survivor list.
struct SurvJob {
    const uint32_t *list;
    int64_t a, b, c, digits;
    uint32_t idx;
    int32_t which;
};

// Terms of admitted jobs whose (k, j) is an entry of a device survivor list
// ((k << 16) | j, as trial_div_survivors writes it): one thread per job, the
// row of the others (host-built, list == nullptr) left as copied.  A
// candidate that is invalid or has another decimal length than its slot
// was sized for sets bit 2 of the error latch.
__global__ void surv_terms_kernel(const SurvJob *jobs, int n, int64_t *terms, uint32_t *err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const SurvJob jb = jobs[i];
    if (!jb.list) return;
    const uint32_t v = jb.list[jb.idx];
    int64_t *t = terms + (size_t)i * TMAX * 2;
    int64_t dg = 0;
    if (!cand_terms(jb.a, jb.b, jb.c, v >> 16, v & 0xFFFFu, jb.which, t, &dg) || dg != jb.digits) {
        for (int q = 0; q < 2 * TMAX; q++) t[q] = 0;
        atomicOr(err, 2u);
    }
}

// cand_terms for a host candidate, as (coef, exp) pairs.
bool candidate_terms(const mr_candidate_t &cd, std::vector<std::pair<int64_t, int64_t>> &terms,
                     int64_t *digits)
{
    int64_t t[2 * TMAX];
    if (!cand_terms(cd.a, cd.b, cd.c, cd.k, cd.j, cd.which, t, digits)) return false;
    terms.clear();
    for (int i = 0; i < TMAX; i++)
        if (t[2 * i] != 0 || t[2 * i + 1] != 0) terms.push_back({t[2 * i], t[2 * i + 1]});
    return true;
}

} // namespace

extern "C" mr_ctx_t *mr_ctx_create_sparse(const mr_candidate_t *cands, int n_batch, int device)
{
    if (!cands || n_batch < 1) return nullptr;
    std::vector<int64_t> terms((size_t)n_batch * TMAX * 2, 0);
    int64_t max_digits = 1;
    for (int i = 0; i < n_batch; i++) {
        std::vector<std::pair<int64_t, int64_t>> t;
        int64_t dg = 0;
        if (!candidate_terms(cands[i], t, &dg) || (int)t.size() > TMAX) {
            fprintf(stderr, "%s: invalid candidate %d (a=%lld b=%lld c=%lld k=%u j=%u which=%d)\n",
                    kWho, i, (long long)cands[i].a, (long long)cands[i].b, (long long)cands[i].c,
                    cands[i].k, cands[i].j, cands[i].which);
            return nullptr;
        }
        for (size_t q = 0; q < t.size(); q++) {
            terms[((size_t)i * TMAX + q) * 2] = t[q].first;
            terms[((size_t)i * TMAX + q) * 2 + 1] = t[q].second;
        }
        if (dg > max_digits) max_digits = dg;
    }
    const long bits = digits_to_bits(max_digits);
    mr_ctx *x = ctx_new((int)((bits + 31) / 32), n_batch, device);
    if (!x) return nullptr;
    int rc = ctx_alloc_all(x, (long)x->n_words * 32, g_min_logm);
    int64_t *dterms = nullptr;
    if (rc == 0 && ctx_alloc(x, (void **)&dterms, sizeof(int64_t) * terms.size()) != 0) rc = -2;
    if (rc == 0 && hipMemcpyAsync(dterms, terms.data(), sizeof(int64_t) * terms.size(),
                                  hipMemcpyHostToDevice, x->st) != hipSuccess) rc = -2;
    if (rc == 0)
        rc = dispatch_logm(x->logm, [&](auto lg) {
            constexpr int LG = decltype(lg)::value;
            build_kernel<LG><<<x->n, THREADS, 0, x->st>>>(x->d, x->T, x->K, dterms, TMAX, nullptr);
            return 0;
        });
    if (rc == 0) rc = ctx_finish(x);
    if (rc != 0) {
        mr_ctx_destroy(x);
        return nullptr;
    }
    return x;
}

extern "C" int mr_set_min_log_len(int log_len)
{
    if (log_len < MIN_LOGM || log_len > MAX_LOGM) return -1;
    g_min_logm = log_len;
    return 0;
}

extern "C" int mr_time_squarings(mr_ctx_t *x, uint32_t base, int n_bits, float *ms)
{
    if (!x || !ms || n_bits < 1 || base < 2u || base >= (1u << 24)) return -1;
    MR_CHECK(hipSetDevice(x->device));
    hipEvent_t e0, e1;
    MR_CHECK(hipEventCreate(&e0));
    MR_CHECK(hipEventCreate(&e1));
    const int n = x->n;
    init_alive_kernel<<<(n + 255) / 256, 256, 0, x->st>>>(x->d);
    int rc = dispatch_logm(x->logm, [&](auto lg) {
        constexpr int LG = decltype(lg)::value;
        round_start_kernel<LG><<<n, THREADS, 0, x->st>>>(x->d, base, 0, nullptr, 0);
        (void)hipEventRecord(e0, x->st);
        const int CHUNK = 256;
        for (int done = 0; done < n_bits; done += CHUNK) {
            const int steps = n_bits - done < CHUNK ? n_bits - done : CHUNK;
            exp_kernel<LG><<<n, THREADS, 0, x->st>>>(x->d, x->T, x->K, steps);
        }
        (void)hipEventRecord(e1, x->st);
        return 0;
    });
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    return e == hipSuccess ? 0 : hip_fail(e, "mr_time_squarings");
}

// ------------------------------------------------------------ scheduler
//
// Slot state machine (host side; one entry per slot of a pool):
//   FREE -> ADM (admin stream: build, setup, first round start)
//   ADM  -> RUN (rounds to run) | FREE (decided without rounds)
//   RUN  -> END (chunk finished the exponent: bitpos < s)
//   END  -> ADM (admin stream: round end, next base or verdict)
// A slot is touched by one stream at a time; chunks carry only RUN slots.

namespace {

enum : uint8_t { SL_FREE = 0, SL_ADM = 1, SL_RUN = 2, SL_END = 3 };

double mono_now()
{
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

struct Req {
    mr_candidate_t cand;
    int64_t tag;
    double t_submit, t_admit;
    int L, logm;
    int64_t digits = 0;             // decimal length (sets L and logm)
    int rounds;
    int64_t terms[TMAX * 2];
    // restored from a checkpoint after admission: the base in progress, the
    // next exponent bit and the residue (L limbs) to resume from
    int resume_bidx = -1, resume_bitpos = 0;
    std::vector<uint32_t> resume_res;
    // submitted in place from a device survivor list (mr_sched_submit_survivors):
    // (k, j) is entry dev_idx of dev_list, (k << 16) | j, read by the device
    // when the job is admitted; cand.k = cand.j = 0 until then
    const uint32_t *dev_list = nullptr;
    uint32_t dev_idx = 0;
};


// Bytes of device memory per slot of a pool of transform length 2^logm
// (the arrays of ctx_alloc_all).
int64_t slot_bytes(int logm)
{
    const int64_t M = 1ll << logm;
    return 4 * (4 * M + 7 * M + (split_logm(logm) ? M + M / 2 : 0) + (half_logm(logm) ? M : 0)) + 32;
}

struct Pool {
    int logm = 0, n = 0;
    int chunk = 0;                  // squaring steps per chunk (graph launch)
    mr_ctx *x = nullptr;            // x->st is the admin stream
    hipStream_t st_run = nullptr;
    hipEvent_t ev_run = nullptr, ev_adm = nullptr;
    int32_t *d_list_adm = nullptr, *d_Ls = nullptr;
    int64_t *d_terms = nullptr;
    uint32_t *d_bases = nullptr;
    int32_t *d_resume = nullptr;    // [n][3] slot, base index, exponent bit
    uint8_t *d_run = nullptr;       // [n] slots that square in the next chunk
    uint8_t *h_run = nullptr;
    hipGraph_t graph = nullptr;     // one chunk: run flags in, squarings, exponent bits out
    hipGraphExec_t gexec = nullptr;
    int32_t *h_resume = nullptr;
    uint32_t *h_resres = nullptr;   // [n][stride] residues to resume (pinned, on first use)
    // pinned host mirrors
    int32_t *h_list_adm = nullptr, *h_Ls = nullptr;
    int32_t *h_bitpos = nullptr, *h_sval = nullptr;
    int64_t *h_terms = nullptr;
    SurvJob *h_sjobs = nullptr, *d_sjobs = nullptr;   // [n] terms sources of an admission
    uint8_t *h_alive = nullptr, *h_done = nullptr;
    uint32_t *h_err = nullptr;
    std::vector<uint8_t> state;
    std::vector<int> req;          // slot -> request index
    std::vector<int> run_list, adm_list;
    int adm_n_end = 0;
    int busy_slots = 0;            // slots holding an admitted job (ADM, RUN, END)
    bool run_busy = false, adm_busy = false;
};

} // namespace

struct mr_sched {
    int device = 0, n_bases = 0, max_active = 0, chunk = 256;
    int cus = 0;                   // compute units of the device
    int pool_slots = 0;            // slots of a pool without a bucket choice
    bool auto_active = false;      // max_active defaulted: follows the buckets' slots
    // per-bucket choices (0: the scheduler-wide chunk / max_active), set
    // before the bucket's pool exists (mr_sched_set_bucket, a tuning table)
    int bucket_chunk[MAX_LOGM + 1] = {};
    int bucket_slots[MAX_LOGM + 1] = {};
    int64_t budget = 0;
    int64_t hbm = 0;               // device memory held by the pools (all slots)
    int64_t job_bytes = 0;         // footprint of the admitted jobs (their slots)
    int64_t peak_job_bytes = 0;
    int64_t reclaimed = 0;         // idle pools released to make room
    std::vector<uint32_t> bases;
    std::vector<Req> reqs;
    std::vector<mr_result_t> done;
    size_t done_head = 0;
    Pool *pools[MAX_LOGM + 1] = {};
    // queued request indices per transform bucket, FIFO
    std::vector<int> queue[MAX_LOGM + 1];
    size_t qhead[MAX_LOGM + 1] = {};
    int active = 0;
    int64_t n_done = 0, chunks = 0, admins = 0, squarings = 0;
    bool failed = false;
};

namespace {

void pool_free(Pool *p)
{
    if (!p) return;
    if (p->st_run) (void)hipStreamSynchronize(p->st_run);
    if (p->x && p->x->st) (void)hipStreamSynchronize(p->x->st);
    if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->ev_run) (void)hipEventDestroy(p->ev_run);
    if (p->ev_adm) (void)hipEventDestroy(p->ev_adm);
    if (p->st_run) (void)hipStreamDestroy(p->st_run);
    void *hp[] = {p->h_list_adm, p->h_Ls, p->h_bitpos, p->h_sval, p->h_terms, p->h_sjobs,
                  p->h_alive, p->h_done, p->h_err, p->h_resume, p->h_resres, p->h_run};
    for (void *q : hp) if (q) (void)hipHostFree(q);
    if (p->x) mr_ctx_destroy(p->x);       // frees the device arrays it owns
    delete p;
}

int pool_new(mr_sched *S, int logm, Pool **out)
{
    const int64_t per = slot_bytes(logm);
    int64_t room = (S->budget - S->hbm) / per;
    int n = std::min(S->bucket_slots[logm] ? S->bucket_slots[logm] : S->pool_slots, S->max_active);
    if (room < n) n = (int)room;
    if (n < 1) return 1;                 // no room under the budget now
    Pool *p = new (std::nothrow) Pool;
    if (!p) return -2;
    p->logm = logm;
    p->n = n;
    p->chunk = S->bucket_chunk[logm] ? S->bucket_chunk[logm] : S->chunk;
    p->x = ctx_new(1, n, S->device);
    int rc = p->x ? 0 : -2;
    // numbers of up to (2^logm - 2) / 2 limbs (24 lmax - 1 bits): the pool's
    // transform length
    const long lmax = std::min<long>(((1l << logm) - 2) / 2, MAX_LIMBS);
    if (rc == 0) rc = ctx_alloc_all(p->x, 24l * lmax - 1, logm);
    if (rc == 0 && p->x->logm != logm) rc = -1;
    auto dev = [&](void **q, size_t b) { if (rc == 0 && ctx_alloc(p->x, q, b) != 0) rc = -2; };
    dev((void **)&p->d_list_adm, 4 * (size_t)n);
    dev((void **)&p->d_Ls, 4 * (size_t)n);
    dev((void **)&p->d_terms, sizeof(int64_t) * TMAX * 2 * (size_t)n);
    dev((void **)&p->d_sjobs, sizeof(SurvJob) * (size_t)n);
    dev((void **)&p->d_bases, 4 * (size_t)S->n_bases);
    dev((void **)&p->d_resume, 12 * (size_t)n);
    dev((void **)&p->d_run, (size_t)n);
    auto host = [&](void **q, size_t b) {
        if (rc == 0 && hipHostMalloc(q, b ? b : 4, hipHostMallocDefault) != hipSuccess) rc = -2;
    };
    host((void **)&p->h_list_adm, 4 * (size_t)n);
    host((void **)&p->h_Ls, 4 * (size_t)n);
    host((void **)&p->h_bitpos, 4 * (size_t)n);
    host((void **)&p->h_sval, 4 * (size_t)n);
    host((void **)&p->h_terms, sizeof(int64_t) * TMAX * 2 * (size_t)n);
    host((void **)&p->h_sjobs, sizeof(SurvJob) * (size_t)n);
    host((void **)&p->h_alive, (size_t)n);
    host((void **)&p->h_done, (size_t)n);
    host((void **)&p->h_err, 4);
    host((void **)&p->h_resume, 12 * (size_t)n);
    host((void **)&p->h_run, (size_t)n);
    if (rc == 0 && (hipStreamCreateWithFlags(&p->st_run, hipStreamNonBlocking) != hipSuccess ||
                    hipEventCreateWithFlags(&p->ev_run, hipEventDisableTiming) != hipSuccess ||
                    hipEventCreateWithFlags(&p->ev_adm, hipEventDisableTiming) != hipSuccess))
        rc = -2;
    if (rc == 0 && hipMemcpyAsync(p->d_bases, S->bases.data(), 4 * (size_t)S->n_bases,
                                  hipMemcpyHostToDevice, p->x->st) != hipSuccess)
        rc = -2;
    if (rc == 0 && hipStreamSynchronize(p->x->st) != hipSuccess) rc = -2;
    if (rc != 0) {
        pool_free(p);
        return rc;
    }
    p->state.assign(n, SL_FREE);
    p->req.assign(n, -1);
    S->hbm += per * n;
    *out = p;
    return 0;
}

// Release pools that hold no admitted job while another bucket waits for
// room under the budget (their memory goes back to the budget).
void reclaim_idle(mr_sched *S, int except)
{
    for (int lg = 0; lg <= MAX_LOGM; lg++) {
        Pool *&p = S->pools[lg];
        if (!p || lg == except || p->busy_slots || p->run_busy || p->adm_busy) continue;
        if (S->qhead[lg] < S->queue[lg].size()) continue;      // its own jobs are waiting
        S->hbm -= slot_bytes(lg) * p->n;
        pool_free(p);
        p = nullptr;
        S->reclaimed++;
    }
}

// A pool for every bucket with queued jobs, as the budget allows: a pool
// takes min(max_active, room / slot bytes) slots; a bucket that finds no
// room releases idle pools of other buckets first, else its jobs wait in
// the queue until room frees up.
int ensure_pools(mr_sched *S)
{
    for (int lg = 0; lg <= MAX_LOGM; lg++) {
        if (S->pools[lg] || S->qhead[lg] >= S->queue[lg].size()) continue;
        int rc = pool_new(S, lg, &S->pools[lg]);
        if (rc == 1) {
            reclaim_idle(S, lg);
            rc = pool_new(S, lg, &S->pools[lg]);
        }
        if (rc == 1) { S->pools[lg] = nullptr; continue; }
        if (rc) { S->pools[lg] = nullptr; return rc; }
    }
    return 0;
}

void finish(mr_sched *S, int r, int verdict)
{
    Req &q = S->reqs[r];
    mr_result_t o;
    o.tag = q.tag;
    o.verdict = verdict;
    o.rounds = q.rounds;
    o.t_submit = q.t_submit;
    o.t_admit = q.t_admit;
    o.t_done = mono_now();
    S->done.push_back(o);
    S->n_done++;
}

// The pool's chunk as a HIP graph, captured once: copy the run flags in,
// one squaring kernel over every slot of the pool (slots not running return
// at once), copy the exponent bits out.  Fixed pointers and chunk length,
// so each chunk is one hipGraphLaunch; no memset nodes.
int pool_graph(mr_sched *S, Pool *p)
{
    hipStream_t st = p->st_run;
    Dev Dg = p->x->d;
    Dg.list = nullptr;
    Dg.run = p->d_run;
    const int chunk = p->chunk, n = p->n;
    MR_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipMemcpyAsync(p->d_run, p->h_run, (size_t)n, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        (void)dispatch_logm(p->logm, [&](auto lg) {
            constexpr int LG = decltype(lg)::value;
            exp_kernel<LG><<<n, THREADS, 0, st>>>(Dg, p->x->T, p->x->K, chunk);
            return 0;
        });
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(p->h_bitpos, p->x->d.bitpos, 4 * (size_t)n, hipMemcpyDeviceToHost, st);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(st, &g);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0);
    p->graph = g;
    if (e != hipSuccess) { p->gexec = nullptr; return hip_fail(e, "chunk graph"); }
    return 0;
}

// Launch one squaring chunk over the pool's RUN slots (one graph launch).
int launch_chunk(mr_sched *S, Pool *p)
{
    p->run_list.clear();
    for (int c = 0; c < p->n; c++) {
        const bool run = p->state[c] == SL_RUN;
        p->h_run[c] = run ? 1 : 0;
        if (run) p->run_list.push_back(c);
    }
    if (p->run_list.empty()) return 0;
    const int nr = (int)p->run_list.size();
    if (!p->gexec) { int rc = pool_graph(S, p); if (rc) return rc; }
    MR_CHECK(hipGraphLaunch(p->gexec, p->st_run));
    MR_CHECK(hipEventRecord(p->ev_run, p->st_run));
    const int chunk = p->chunk;
    p->run_busy = true;
    S->chunks++;
    S->squarings += (int64_t)nr * chunk;
    return 0;
}

void chunk_done(Pool *p)
{
    for (int c : p->run_list)
        if (p->h_bitpos[c] < p->h_sval[c]) p->state[c] = SL_END;
    p->run_busy = false;
}

// Admin batch: end the rounds of END slots, admit pending requests into
// FREE slots (within the scheduler's active cap).
int launch_admin(mr_sched *S, Pool *p)
{
    p->adm_list.clear();
    for (int c = 0; c < p->n; c++)
        if (p->state[c] == SL_END) p->adm_list.push_back(c);
    const int n_end = (int)p->adm_list.size();
    int n_res = 0;          // restored candidates among the admitted
    int n_dev = 0;          // admitted from device survivor lists
    std::vector<int> &queue = S->queue[p->logm];
    size_t &qh = S->qhead[p->logm];
    const int64_t per = slot_bytes(p->logm);
    // a job is admitted when a slot of its bucket is free, the active cap
    // allows, and its footprint (its slot's device state) fits the budget
    for (int c = 0; c < p->n && qh < queue.size() && S->active < S->max_active &&
                    S->job_bytes + per <= S->budget; c++) {
        if (p->state[c] != SL_FREE) continue;
        const int r = queue[qh++];
        Req &q = S->reqs[r];
        const int i = (int)p->adm_list.size() - n_end;
        memcpy(p->h_terms + (size_t)i * TMAX * 2, q.terms, sizeof(q.terms));
        p->h_sjobs[i] = SurvJob{q.dev_list, q.cand.a, q.cand.b, q.cand.c, q.digits, q.dev_idx,
                                q.cand.which};
        n_dev += q.dev_list ? 1 : 0;
        p->h_Ls[i] = q.L;
        if (q.resume_bidx >= 0) {
            if (!p->h_resres &&
                hipHostMalloc(&p->h_resres, 4 * (size_t)p->n * p->x->stride, hipHostMallocDefault) != hipSuccess) {
                p->h_resres = nullptr;
                return hip_fail(hipErrorOutOfMemory, "resume staging");
            }
            memcpy(p->h_resres + (size_t)n_res * p->x->stride, q.resume_res.data(), 4 * (size_t)q.L);
            p->h_resume[3 * n_res] = c;
            p->h_resume[3 * n_res + 1] = q.resume_bidx;
            p->h_resume[3 * n_res + 2] = q.resume_bitpos;
            n_res++;
        }
        q.t_admit = mono_now();
        p->req[c] = r;
        p->adm_list.push_back(c);
        S->active++;
        S->job_bytes += per;
        S->peak_job_bytes = std::max(S->peak_job_bytes, S->job_bytes);
        p->busy_slots++;
    }
    if (qh == queue.size()) { queue.clear(); qh = 0; }
    const int n_all = (int)p->adm_list.size(), n_new = n_all - n_end;
    if (n_all == 0) return 0;
    for (int c : p->adm_list) p->state[c] = SL_ADM;
    p->adm_n_end = n_end;
    hipStream_t st = p->x->st;
    memcpy(p->h_list_adm, p->adm_list.data(), 4 * (size_t)n_all);
    MR_CHECK(hipMemcpyAsync(p->d_list_adm, p->h_list_adm, 4 * (size_t)n_all, hipMemcpyHostToDevice, st));
    if (n_new) {
        MR_CHECK(hipMemcpyAsync(p->d_terms, p->h_terms, sizeof(int64_t) * TMAX * 2 * (size_t)n_new,
                                hipMemcpyHostToDevice, st));
        MR_CHECK(hipMemcpyAsync(p->d_Ls, p->h_Ls, 4 * (size_t)n_new, hipMemcpyHostToDevice, st));
        if (n_dev) {        // terms of jobs from device survivor lists, built in place
            MR_CHECK(hipMemcpyAsync(p->d_sjobs, p->h_sjobs, sizeof(SurvJob) * (size_t)n_new,
                                    hipMemcpyHostToDevice, st));
            surv_terms_kernel<<<(n_new + 63) / 64, 64, 0, st>>>(p->d_sjobs, n_new, p->d_terms,
                                                                 p->x->d.err);
            MR_CHECK(hipGetLastError());
        }
    }
    Dev De = p->x->d, Dn = p->x->d;
    De.list = p->d_list_adm;
    Dn.list = p->d_list_adm + n_end;
    const Twiddles &T = p->x->T;
    const Consts &K = p->x->K;
    const uint32_t *bases = p->d_bases;
    const int nb = S->n_bases;
    int rc = dispatch_logm(p->logm, [&](auto lg) {
        constexpr int LG = decltype(lg)::value;
        if (n_end) {
            round_end_kernel<LG><<<n_end, THREADS, 0, st>>>(De, T, K, nullptr, 0);
            round_start_kernel<LG><<<n_end, THREADS, 0, st>>>(De, 0u, 2, bases, nb);
        }
        if (n_new) {
            build_kernel<LG><<<n_new, THREADS, 0, st>>>(Dn, T, K, p->d_terms, TMAX, p->d_Ls);
            setup_kernel<LG><<<n_new, THREADS, 0, st>>>(Dn, T, K);
            round_start_kernel<LG><<<n_new, THREADS, 0, st>>>(Dn, 0u, 1, bases, nb);
        }
        return 0;
    });
    if (rc) return rc;
    MR_CHECK(hipGetLastError());
    if (n_res) {        // restored candidates: residue and position after the round start
        for (int i = 0; i < n_res; i++) {
            const int c = p->h_resume[3 * i];
            MR_CHECK(hipMemcpyAsync(p->x->d.res + (size_t)c * p->x->stride,
                                    p->h_resres + (size_t)i * p->x->stride,
                                    4 * (size_t)S->reqs[p->req[c]].L, hipMemcpyHostToDevice, st));
        }
        MR_CHECK(hipMemcpyAsync(p->d_resume, p->h_resume, 12 * (size_t)n_res, hipMemcpyHostToDevice, st));
        resume_kernel<<<(n_res + 63) / 64, 64, 0, st>>>(p->x->d, p->d_resume, n_res, bases);
        MR_CHECK(hipGetLastError());
    }
    const size_t n = (size_t)p->n;
    MR_CHECK(hipMemcpyAsync(p->h_alive, p->x->d.alive, n, hipMemcpyDeviceToHost, st));
    MR_CHECK(hipMemcpyAsync(p->h_done, p->x->d.done, n, hipMemcpyDeviceToHost, st));
    MR_CHECK(hipMemcpyAsync(p->h_sval, p->x->d.sval, 4 * n, hipMemcpyDeviceToHost, st));
    MR_CHECK(hipMemcpyAsync(p->h_err, p->x->d.err, 4, hipMemcpyDeviceToHost, st));
    MR_CHECK(hipEventRecord(p->ev_adm, st));
    p->adm_busy = true;
    S->admins++;
    return 0;
}

int admin_done(mr_sched *S, Pool *p)
{
    p->adm_busy = false;
    if (*p->h_err) {
        fprintf(stderr, "%s: a candidate does not fit its slot (or a survivor-list entry is "
                        "not a valid (k, j))\n", kWho);
        return -2;
    }
    for (int i = 0; i < (int)p->adm_list.size(); i++) {
        const int c = p->adm_list[i];
        const int r = p->req[c];
        if (i < p->adm_n_end) S->reqs[r].rounds++;     // the round that just ended
        if (p->h_done[c]) {
            finish(S, r, p->h_alive[c] ? 1 : 0);
            p->state[c] = SL_FREE;
            p->req[c] = -1;
            S->active--;
            S->job_bytes -= slot_bytes(p->logm);
            p->busy_slots--;
        } else {
            p->state[c] = SL_RUN;
        }
    }
    return 0;
}

int sched_pending(const mr_sched *S)
{
    return (int)((int64_t)S->reqs.size() - S->n_done);
}

} // namespace

extern "C" mr_sched_t *mr_sched_create(const uint32_t *bases, int n_bases, int device,
                                       int max_active, int chunk, int64_t hbm_budget)
{
    if (!bases || n_bases < 1 || chunk < 0 || max_active < 0 || hbm_budget < 0) return nullptr;
    for (int i = 0; i < n_bases; i++)
        if (bases[i] < 2u || bases[i] >= (1u << 24)) return nullptr;
    if (device < 0 && hipGetDevice(&device) != hipSuccess) return nullptr;
    if (hipSetDevice(device) != hipSuccess) return nullptr;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return nullptr;
    mr_sched *S = new (std::nothrow) mr_sched;
    if (!S) return nullptr;
    S->auto_active = max_active == 0;
    if (max_active == 0) max_active = cus;
    S->cus = cus;
    S->pool_slots = max_active;
    S->device = device;
    S->n_bases = n_bases;
    S->bases.assign(bases, bases + n_bases);
    S->max_active = max_active;
    S->chunk = chunk ? chunk : 256;
    S->budget = hbm_budget ? hbm_budget : (64ll << 30);
    return S;
}

namespace {

// A request for candidate cd: sparse terms and geometry.  False (with a
// message) when the candidate cannot be formed or is too large.
bool make_req(const mr_candidate_t &cd, int64_t tag, Req &q)
{
    std::vector<std::pair<int64_t, int64_t>> t;
    int64_t dg = 0;
    if (!candidate_terms(cd, t, &dg) || (int)t.size() > TMAX) {
        fprintf(stderr, "%s: invalid candidate (a=%lld b=%lld c=%lld k=%u j=%u which=%d)\n", kWho,
                (long long)cd.a, (long long)cd.b, (long long)cd.c, cd.k, cd.j, cd.which);
        return false;
    }
    geometry(digits_to_bits(dg), MIN_LOGM, &q.L, &q.logm);
    q.digits = dg;
    if (q.logm > MAX_LOGM || q.L > MAX_LIMBS) {
        fprintf(stderr, "%s: candidate of %lld digits is beyond this build's transform sizes\n", kWho,
                (long long)dg);
        return false;
    }
    memset(q.terms, 0, sizeof(q.terms));
    for (size_t k = 0; k < t.size(); k++) {
        q.terms[2 * k] = t[k].first;
        q.terms[2 * k + 1] = t[k].second;
    }
    q.cand = cd;
    q.tag = tag;
    q.rounds = 0;
    q.t_admit = 0.0;
    return true;
}

// Queue validated requests in their bucket (pools are created by the pump
// as the budget allows).  A job whose footprint alone exceeds the budget
// could never be admitted and is rejected with the whole call.
int enqueue(mr_sched *S, std::vector<Req> &add)
{
    const double now = mono_now();
    if (hipSetDevice(S->device) != hipSuccess) return -2;
    for (const Req &q : add)
        if (slot_bytes(q.logm) > S->budget) {
            fprintf(stderr, "%s: a candidate needs %lld bytes of device state, more than the "
                            "HBM budget of %lld\n", kWho, (long long)slot_bytes(q.logm),
                    (long long)S->budget);
            return -1;
        }
    for (Req &q : add) {
        q.t_submit = now;
        S->queue[q.logm].push_back((int)S->reqs.size());
        S->reqs.push_back(std::move(q));
    }
    return 0;
}

} // namespace

extern "C" int mr_sched_submit(mr_sched_t *S, const mr_candidate_t *cands, const int64_t *tags, int n)
{
    if (!S || (!cands && n) || n < 0) return -1;
    try {
        std::vector<Req> add((size_t)n);
        for (int i = 0; i < n; i++)
            if (!make_req(cands[i], tags ? tags[i] : (int64_t)(S->reqs.size() + i), add[i]))
                return -1;
        // longest first within one call (the classical rule for finishing a
        // batch early on identical units: the longest tests start while the
        // short ones fill the gaps at the end); calls keep their order
        std::stable_sort(add.begin(), add.end(),
                         [](const Req &x, const Req &y) { return x.L > y.L; });
        return enqueue(S, add);
    } catch (...) {             // allocation failure: nothing escapes the C ABI
        return -2;
    }
}

namespace {

int submit_survivors(mr_sched *S, long long a, long long b, long long c, const uint32_t *d_list,
                     uint32_t n, int which, int64_t tag_base)
{
    std::vector<Req> add((size_t)n);
    if (std::max(b, c) > a - 7) {
        // k 10^b + j 10^c may reach the leading digit, so the decimal length
        // (and the slot size) can differ between survivors: the host path
        std::vector<uint32_t> h(n);
        const hipError_t e = hipMemcpy(h.data(), d_list, 4 * (size_t)n, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(e, "survivor list");
        for (uint32_t i = 0; i < n; i++) {
            mr_candidate_t cd{a, b, c, h[i] >> 16, h[i] & 0xFFFFu, which};
            if (!make_req(cd, tag_base + i, add[i])) return -1;
        }
        return enqueue(S, add);
    }
    // k, j < 2^16 < 10^5, so k 10^b + j 10^c < 10^(a-1): every survivor has a
    // decimal digits, and one representative fixes the slot size
    Req proto;
    if (!make_req(mr_candidate_t{a, b, c, 1u, 1u, which}, 0, proto)) return -1;
    for (uint32_t i = 0; i < n; i++) {
        Req &q = add[i];
        q = proto;
        memset(q.terms, 0, sizeof(q.terms));
        q.cand.k = q.cand.j = 0u;
        q.tag = tag_base + i;
        q.dev_list = d_list;
        q.dev_idx = i;
    }
    return enqueue(S, add);
}

} // namespace

extern "C" int mr_sched_submit_survivors(mr_sched_t *S, long long a, long long b, long long c,
                                         const uint32_t *d_list, uint32_t n, int which,
                                         int64_t tag_base)
{
    if (!S || (!d_list && n) || which < 0 || which > 2 || a < 1 || b < 0 || c < 0) return -1;
    if (n == 0) return 0;
    if (hipSetDevice(S->device) != hipSuccess) return -2;
    try {
        return submit_survivors(S, a, b, c, d_list, n, which, tag_base);
    } catch (...) {             // allocation failure: nothing escapes the C ABI
        return -2;
    }
}


namespace {

int sched_checkpoint(mr_sched *S, const char *path)
{
    if (S->failed) return -2;
    if (hipSetDevice(S->device) != hipSuccess) return -2;
    // 1. let the work in flight finish and take its results
    for (Pool *p : S->pools) {
        if (!p) continue;
        if (p->run_busy) {
            hipError_t e = hipEventSynchronize(p->ev_run);
            if (e != hipSuccess) { S->failed = true; return hip_fail(e, "checkpoint"); }
            chunk_done(p);
        }
        if (p->adm_busy) {
            hipError_t e = hipEventSynchronize(p->ev_adm);
            if (e != hipSuccess) { S->failed = true; return hip_fail(e, "checkpoint"); }
            if (admin_done(S, p) != 0) { S->failed = true; return -2; }
        }
    }
    // 2. per undecided request: admitted (slot state) or still queued
    struct Row { int r, bidx, bitpos; std::vector<uint32_t> res; };
    std::vector<Row> rows;
    for (Pool *p : S->pools) {
        if (!p) continue;
        std::vector<int32_t> bidx((size_t)p->n), bitpos((size_t)p->n);
        MR_CHECK(hipMemcpy(bidx.data(), p->x->d.bidx, 4 * (size_t)p->n, hipMemcpyDeviceToHost));
        MR_CHECK(hipMemcpy(bitpos.data(), p->x->d.bitpos, 4 * (size_t)p->n, hipMemcpyDeviceToHost));
        for (int c = 0; c < p->n; c++) {
            if (p->state[c] != SL_RUN && p->state[c] != SL_END) continue;
            Row w{p->req[c], bidx[c], bitpos[c], {}};
            w.res.resize((size_t)S->reqs[w.r].L);
            MR_CHECK(hipMemcpy(w.res.data(), p->x->d.res + (size_t)c * p->x->stride, 4 * w.res.size(),
                               hipMemcpyDeviceToHost));
            rows.push_back(std::move(w));
        }
    }
    for (int lg = 0; lg <= MAX_LOGM; lg++)
        for (size_t i = S->qhead[lg]; i < S->queue[lg].size(); i++)
            rows.push_back(Row{S->queue[lg][i], -1, 0, {}});
    std::sort(rows.begin(), rows.end(), [](const Row &x, const Row &y) { return x.r < y.r; });
    // jobs submitted from device survivor lists: read their (k, j) now
    for (const Row &w : rows) {
        Req &q = S->reqs[w.r];
        if (!q.dev_list || q.cand.k) continue;
        uint32_t v = 0;
        MR_CHECK(hipMemcpy(&v, q.dev_list + q.dev_idx, 4, hipMemcpyDeviceToHost));
        q.cand.k = v >> 16;
        q.cand.j = v & 0xFFFFu;
    }
    // 3. the file (written whole, then renamed over path)
    const std::string tmp = std::string(path) + ".tmp";
    FILE *f = fopen(tmp.c_str(), "w");
    if (!f) return -1;
    fprintf(f, "mr-sched-checkpoint 1\nbases %d", S->n_bases);
    for (uint32_t b : S->bases) fprintf(f, " %u", b);
    fprintf(f, "\ncandidates %zu\n", rows.size());
    for (const Row &w : rows) {
        const Req &q = S->reqs[w.r];
        fprintf(f, "%lld %lld %lld %lld %u %u %d %d %d %d %zu", (long long)q.tag, (long long)q.cand.a,
                (long long)q.cand.b, (long long)q.cand.c, q.cand.k, q.cand.j, q.cand.which, q.rounds,
                w.bidx, w.bitpos, w.res.size());
        for (uint32_t v : w.res) fprintf(f, " %x", v);
        fputc('\n', f);
    }
    const bool ok = !ferror(f);
    if (fclose(f) != 0 || !ok || rename(tmp.c_str(), path) != 0) {
        (void)remove(tmp.c_str());
        return -1;
    }
    return (int)rows.size();
}

} // namespace

extern "C" int mr_sched_checkpoint(mr_sched_t *S, const char *path)
{
    if (!S || !path) return -1;
    try {
        return sched_checkpoint(S, path);
    } catch (...) {             // allocation failure: nothing escapes the C ABI
        return -2;
    }
}

namespace {

int sched_restore(mr_sched *S, const char *path)
{
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    std::vector<Req> add;
    auto bad = [&](const char *what) {
        fprintf(stderr, "%s: checkpoint %s: %s\n", kWho, path, what);
        fclose(f);
        return -1;
    };
    int version = 0, nb = 0;
    if (fscanf(f, "mr-sched-checkpoint %d bases %d", &version, &nb) != 2 || version != 1)
        return bad("not a checkpoint of this format");
    if (nb != S->n_bases) return bad("another list of bases");
    for (int i = 0; i < nb; i++) {
        unsigned b = 0;
        if (fscanf(f, "%u", &b) != 1 || b != S->bases[i]) return bad("another list of bases");
    }
    long long count = 0;
    if (fscanf(f, " candidates %lld", &count) != 1 || count < 0 || count > (1ll << 30))
        return bad("bad candidate count");
    add.reserve((size_t)std::min(count, 1ll << 16));   // the lines, not the header, set the size
    for (long long i = 0; i < count; i++) {
        long long tag, a, b, c;
        unsigned k, j;
        int which, rounds, bidx, bitpos;
        size_t L;
        if (fscanf(f, "%lld %lld %lld %lld %u %u %d %d %d %d %zu", &tag, &a, &b, &c, &k, &j, &which, &rounds,
                   &bidx, &bitpos, &L) != 11)
            return bad("truncated candidate line");
        mr_candidate_t cd{};
        cd.a = a; cd.b = b; cd.c = c; cd.k = k; cd.j = j; cd.which = which;
        add.emplace_back();
        Req &q = add.back();
        if (!make_req(cd, tag, q)) return bad("invalid candidate");
        if (rounds < 0 || rounds > nb || bidx < -1 || bidx >= nb) return bad("bad round state");
        q.rounds = rounds;
        if (bidx >= 0) {
            if (L != (size_t)q.L || bitpos < -1 || bitpos > 24 * q.L) return bad("bad residue state");
            q.resume_bidx = bidx;
            q.resume_bitpos = bitpos;
            q.resume_res.resize(L);
            for (size_t w = 0; w < L; w++) {
                unsigned v = 0;
                if (fscanf(f, "%x", &v) != 1 || v > LIMB_MASK) return bad("bad residue");
                q.resume_res[w] = v;
            }
        } else if (L != 0) {
            return bad("residue on a queued candidate");
        }
    }
    fclose(f);
    const int rc = enqueue(S, add);
    return rc ? rc : (int)count;
}

} // namespace

extern "C" int mr_sched_restore(mr_sched_t *S, const char *path)
{
    if (!S || !path) return -1;
    try {
        return sched_restore(S, path);
    } catch (...) {             // allocation failure: nothing escapes the C ABI
        return -2;
    }
}

namespace {
int sched_pump(mr_sched *S, int block);
}

extern "C" int mr_sched_pump(mr_sched_t *S, int block)
{
    if (!S) return -1;
    try {
        return sched_pump(S, block);
    } catch (...) {             // allocation failure: nothing escapes the C ABI
        S->failed = true;
        return -2;
    }
}

namespace {

int sched_pump(mr_sched *S, int block)
{
    if (S->failed) return -2;
    if (hipSetDevice(S->device) != hipSuccess) return -2;
    auto fail = [&](int rc) { S->failed = true; return rc < 0 ? rc : -2; };
    for (;;) {
        // 1. collect finished work
        for (Pool *p : S->pools) {
            if (!p) continue;
            if (p->run_busy) {
                hipError_t e = hipEventQuery(p->ev_run);
                if (e == hipSuccess) chunk_done(p);
                else if (e != hipErrorNotReady) return fail(hip_fail(e, "chunk"));
            }
            if (p->adm_busy) {
                hipError_t e = hipEventQuery(p->ev_adm);
                if (e == hipSuccess) { int rc = admin_done(S, p); if (rc) return fail(rc); }
                else if (e != hipErrorNotReady) return fail(hip_fail(e, "admission"));
            }
        }
        // 2. pools for the buckets with queued jobs (as the budget allows);
        //    launch admissions oldest-first over the pools, then chunks
        { int rc = ensure_pools(S); if (rc) return fail(rc); }
        Pool *order[MAX_LOGM + 1];
        int np = 0;
        for (Pool *p : S->pools) if (p) order[np++] = p;
        for (int a = 1; a < np; a++)       // by submit time of the head request,
            for (int b = a; b > 0; b--) {  // the longer transform first on a tie
                auto head = [&](Pool *p) {
                    const int lg = p->logm;
                    return S->qhead[lg] < S->queue[lg].size()
                               ? S->reqs[S->queue[lg][S->qhead[lg]]].t_submit : 1e300;
                };
                const double hb = head(order[b]), ha = head(order[b - 1]);
                if (hb < ha || (hb == ha && order[b]->logm > order[b - 1]->logm))
                    std::swap(order[b], order[b - 1]);
                else
                    break;
            }
        for (int i = 0; i < np; i++) {
            Pool *p = order[i];
            if (!p->adm_busy) { int rc = launch_admin(S, p); if (rc) return fail(rc); }
        }
        for (int i = 0; i < np; i++) {
            Pool *p = order[i];
            if (!p->run_busy) { int rc = launch_chunk(S, p); if (rc) return fail(rc); }
        }
        const int left = sched_pending(S);
        bool busy = false;
        for (Pool *p : S->pools) if (p && (p->run_busy || p->adm_busy)) busy = true;
        if (!busy && left > 0) {       // every undecided candidate is queued or in flight
            fprintf(stderr, "%s: scheduler stalled with %d candidates left\n", kWho, left);
            return fail(-2);
        }
        if (!block || !busy) return left;
        // 3. wait for the first chunk in flight (else an admission batch)
        hipEvent_t ev = nullptr;
        for (Pool *p : S->pools) if (p && p->run_busy) { ev = p->ev_run; break; }
        if (!ev)
            for (Pool *p : S->pools) if (p && p->adm_busy) { ev = p->ev_adm; break; }
        hipError_t e = hipEventSynchronize(ev);
        if (e != hipSuccess) return fail(hip_fail(e, "wait"));
        block = 0;      // one wait per call; loop once more to act on it
    }
}

} // namespace

extern "C" int mr_sched_results(mr_sched_t *S, mr_result_t *out, int cap)
{
    if (!S || (!out && cap) || cap < 0) return -1;
    int k = 0;
    while (k < cap && S->done_head < S->done.size()) out[k++] = S->done[S->done_head++];
    if (S->done_head == S->done.size()) { S->done.clear(); S->done_head = 0; }
    return k;
}

extern "C" int mr_sched_set_bucket(mr_sched_t *S, int log_len, int chunk, int slots)
{
    if (!S || log_len < MIN_LOGM || log_len > MAX_LOGM || chunk < 0 || chunk > (1 << 20) ||
        slots < 0 || slots > (1 << 20))
        return -1;
    if (S->pools[log_len]) return -1;   // the pool is built: its graph holds the chunk
    S->bucket_chunk[log_len] = chunk;
    S->bucket_slots[log_len] = slots;
    if (S->auto_active && slots > S->max_active) S->max_active = slots;
    return 0;
}

namespace {

// The "buckets" array of a tuning table: objects of integer fields only,
// e.g. {"log_len": 14, "chunk": 256, "slots_per_cu": 1, ...}.  Other
// top-level keys (measurements, device) are skipped.  Returns the number of
// buckets, or -1 when the array is missing or malformed.
int parse_tuning(const std::string &t, int *log_len, int *chunk, int *spc, int cap)
{
    const size_t key = t.find("\"buckets\"");
    if (key == std::string::npos) return -1;
    size_t i = t.find_first_not_of(" \t\r\n", key + 9);
    if (i == std::string::npos || t[i] != ':') return -1;
    i = t.find_first_not_of(" \t\r\n", i + 1);
    if (i == std::string::npos || t[i] != '[') return -1;
    int n = 0;
    for (i++;;) {
        i = t.find_first_not_of(" \t\r\n,", i);
        if (i == std::string::npos) return -1;
        if (t[i] == ']') return n;
        if (t[i] != '{') return -1;
        const size_t end = t.find('}', i);
        if (end == std::string::npos) return -1;
        const std::string obj = t.substr(i + 1, end - i - 1);
        if (obj.find_first_of("{[") != std::string::npos) return -1;
        long v[3] = {-1, -1, -1};
        const char *names[3] = {"\"log_len\"", "\"chunk\"", "\"slots_per_cu\""};
        for (int f = 0; f < 3; f++) {
            const size_t k = obj.find(names[f]);
            if (k == std::string::npos) return -1;
            const size_t c = obj.find(':', k);
            if (c == std::string::npos) return -1;
            char *stop = nullptr;
            v[f] = strtol(obj.c_str() + c + 1, &stop, 10);
            if (stop == obj.c_str() + c + 1) return -1;
        }
        if (v[0] < MIN_LOGM || v[0] > MAX_LOGM || v[1] < 0 || v[1] > (1 << 20) || v[2] < 0 || v[2] > 64)
            return -1;
        if (n < cap) {
            log_len[n] = (int)v[0];
            chunk[n] = (int)v[1];
            spc[n] = (int)v[2];
        }
        n++;
        i = end + 1;
    }
}

} // namespace

extern "C" int mr_tuning_read(const char *path, int *log_len, int *chunk, int *slots_per_cu,
                              int cap)
{
    if (!path || cap < 0 || (cap > 0 && (!log_len || !chunk || !slots_per_cu))) return -1;
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    std::string t;
    char buf[4096];
    size_t got;
    while ((got = fread(buf, 1, sizeof buf, f)) > 0 && t.size() < (1u << 22)) t.append(buf, got);
    fclose(f);
    return parse_tuning(t, log_len, chunk, slots_per_cu, cap);
}

extern "C" int mr_sched_load_tuning(mr_sched_t *S, const char *path)
{
    if (!S) return -1;
    int lg[MAX_LOGM + 1], ch[MAX_LOGM + 1], spc[MAX_LOGM + 1];
    const int n = mr_tuning_read(path, lg, ch, spc, MAX_LOGM + 1);
    if (n < 0 || n > MAX_LOGM + 1) return -1;
    for (int i = 0; i < n; i++)
        if (S->pools[lg[i]]) return -1;          // nothing set: a pool is built already
    for (int i = 0; i < n; i++)
        if (mr_sched_set_bucket(S, lg[i], ch[i], spc[i] * S->cus) != 0) return -1;
    return n;
}

extern "C" int mr_sched_stats(const mr_sched_t *S, mr_sched_stats_t *o)
{
    if (!S || !o) return -1;
    o->submitted = (int64_t)S->reqs.size();
    o->completed = S->n_done;
    o->active = S->active;
    o->pending = o->submitted - o->completed - o->active;
    o->chunks = S->chunks;
    o->admin_batches = S->admins;
    o->squarings = S->squarings;
    o->hbm_bytes = S->hbm;
    o->max_active = S->max_active;
    int np = 0;
    for (Pool *p : S->pools) np += p != nullptr;
    o->n_pools = np;
    o->job_bytes = S->job_bytes;
    o->peak_job_bytes = S->peak_job_bytes;
    o->reclaimed_pools = S->reclaimed;
    return 0;
}

extern "C" int64_t mr_sched_job_bytes(int64_t bits)
{
    if (bits < 1) return -1;
    int L, logm;
    geometry((long)bits, MIN_LOGM, &L, &logm);
    if (logm > MAX_LOGM || L > MAX_LIMBS) return -1;
    return slot_bytes(logm);
}

extern "C" void mr_sched_destroy(mr_sched_t *S)
{
    if (!S) return;
    (void)hipSetDevice(S->device);
    for (Pool *&p : S->pools) { pool_free(p); p = nullptr; }
    delete S;
}

extern "C" int mr_test_batch(const mr_candidate_t *cands, int n, const uint32_t *bases, int n_bases,
                             int device, uint8_t *verdict, int32_t *rounds)
{
    if (!cands || n < 1 || !verdict) return -1;
    mr_sched_t *S = mr_sched_create(bases, n_bases, device, 0, 0, 0);
    if (!S) return -1;
    int rc = 0;
    try {
        std::vector<int64_t> tags((size_t)n);
        for (int i = 0; i < n; i++) tags[i] = i;
        rc = mr_sched_submit(S, cands, tags.data(), n);
        while (rc == 0) {
            const int left = mr_sched_pump(S, 1);
            if (left < 0) { rc = left; break; }
            if (left == 0) break;
        }
        if (rc == 0) {
            std::vector<mr_result_t> res((size_t)n);
            const int k = mr_sched_results(S, res.data(), n);
            for (int i = 0; i < k; i++) {
                const int64_t t = res[i].tag;
                verdict[t] = (uint8_t)res[i].verdict;
                if (rounds) rounds[t] = res[i].rounds;
            }
            if (k != n) rc = -2;
        }
    } catch (...) {             // allocation failure: nothing escapes the C ABI
        rc = -2;
    }
    mr_sched_destroy(S);
    return rc;
}
