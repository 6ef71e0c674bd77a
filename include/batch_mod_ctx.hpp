/*
 * batch_mod_ctx.hpp -- the reference's C++ surface of the GPU Miller-Rabin
 * runner, header-only over the C ABI of mr_gpu.h (link libpfamd.so).
 *
 * Replaces:
 *   BatchModCtx(N_all, n_limbs, n_batch, device)     batch_mod_ctx.cuh:132-133
 *   BatchModCtx(std::vector<mpz_t*>, device)         batch_mod_ctx.cuh:137
 *     (declared when <gmp.h> is included first)
 *   to_residue_batch, from_residue_batch,
 *   check_passed, modmul_batch, modsq_batch          batch_mod_ctx.cuh:141-155
 *   gpu_miller_rabin_s1(...), gpu_miller_rabin(...)  miller_rabin_runner.cuh:28-49
 *   DEFAULT_WITNESSES                                miller_rabin_runner.cuh:12-13
 * so that callers such as bench_mr_gpu.cu:60-64 and the arithmetic checks of
 * correctness_tests.cuh compile against it.  N_all holds 16-bit limbs, one
 * per uint64_t, [n_batch][n_limbs] (candidate.cuh:62-70).
 *
 * Differences, by design:
 *  - The working form of a residue is this library's: x R mod N with
 *    R = 2^(24 L), stored as L 24-bit limbs, one per uint64_t, in the
 *    caller's [n_batch][n_limbs] buffer (zeros above L; L < n_limbs for
 *    every modulus above 48 bits).  to_residue_batch / from_residue_batch
 *    convert between it and 16-bit-limb values, so a caller that only
 *    converts, multiplies and checks never sees the difference; one that
 *    builds working-form values by hand does.
 *  - d and s are derived from N on the device.  exp_all and Nm1_all, when
 *    not empty, and s are checked against them and a mismatch throws
 *    (the reference ignores Nm1_all, miller_rabin_runner.cu:284,368).
 *  - A candidate retires at its first failing witness (same verdicts).
 *  - Streams are passed as void* (a hipStream_t; nullptr = null stream).
 *  - Errors throw std::runtime_error, as the reference's CU macros do.
 */
#pragma once
#include "mr_gpu.h"

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

inline const std::vector<uint32_t> DEFAULT_WITNESSES = {
    2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53};

struct BatchModCtx {
    int n_limbs = 0, n_batch = 0, device_id = 0;
    mr_ctx_t *ctx = nullptr;
    std::vector<uint64_t> N_host;      // the moduli as given (16-bit limbs), for argument checks

    explicit BatchModCtx(const std::vector<uint64_t> &N_all, int n_limbs_, int n_batch_,
                         int device_id_ = 0)
    {
        init(N_all, n_limbs_, n_batch_, device_id_);
    }

#ifdef __GNU_MP_VERSION
    // From GMP numbers: n_limbs = 16-bit limbs of the largest one.
    explicit BatchModCtx(const std::vector<mpz_t *> &numbers, int device_id_ = 0)
    {
        size_t bits = 1;
        for (const mpz_t *z : numbers) bits = std::max(bits, mpz_sizeinbase(*z, 2));
        const int nl = (int)((bits + 15) / 16);
        std::vector<uint64_t> all(numbers.size() * (size_t)nl, 0u);
        for (size_t i = 0; i < numbers.size(); i++) {
            std::vector<uint16_t> w((size_t)nl, 0u);
            size_t cnt = 0;
            mpz_export(w.data(), &cnt, -1, sizeof(uint16_t), 0, 0, *numbers[i]);
            for (size_t k = 0; k < cnt; k++) all[i * (size_t)nl + k] = w[k];
        }
        init(all, nl, (int)numbers.size(), device_id_);
    }
#endif

    ~BatchModCtx() { mr_ctx_destroy(ctx); }
    BatchModCtx(const BatchModCtx &) = delete;
    BatchModCtx &operator=(const BatchModCtx &) = delete;

    // x_all (host, 16-bit limbs, each x < N) -> working form (host).
    void to_residue_batch(const std::vector<uint64_t> &x_all, std::vector<uint64_t> &out_all) const
    {
        if (x_all.size() < size())
            throw std::runtime_error("to_residue_batch: x_all smaller than n_batch * n_limbs");
        out_all.assign(size(), 0u);
        if (mr_ctx_to_residue(ctx, x_all.data(), n_limbs, out_all.data(), n_limbs) != 0)
            throw std::runtime_error("to_residue_batch failed (a value not below its modulus?)");
    }

    // d_x (device, working form) -> values x R^-1 mod N (host, 16-bit limbs).
    void from_residue_batch(const uint64_t *d_x, std::vector<uint64_t> &out_all) const
    {
        out_all.assign(size(), 0u);
        if (mr_ctx_from_residue_dev(ctx, d_x, n_limbs, out_all.data(), n_limbs) != 0)
            throw std::runtime_error("from_residue_batch failed");
    }

    // d_passed[c] = 1 if d_r_mont[c] is the working form of 1 or of N - 1.
    void check_passed(const uint64_t *d_r_mont, uint8_t *d_passed, void *s = nullptr) const
    {
        if (mr_ctx_check_passed_dev(ctx, d_r_mont, n_limbs, d_passed, s) != 0)
            throw std::runtime_error("check_passed failed");
    }

    // d_out = mont_mul(d_A, d_B) for every candidate.
    void modmul_batch(const uint64_t *d_A, const uint64_t *d_B, uint64_t *d_out, void *s = nullptr)
    {
        if (mr_ctx_modmul_dev(ctx, d_A, d_B, d_out, n_limbs, s) != 0)
            throw std::runtime_error("modmul_batch failed");
    }

    // d_out = mont_sq(d_A) for every candidate.
    void modsq_batch(const uint64_t *d_A, uint64_t *d_out, void *s = nullptr)
    {
        if (mr_ctx_modsq_dev(ctx, d_A, d_out, n_limbs, s) != 0)
            throw std::runtime_error("modsq_batch failed");
    }

    size_t size() const { return (size_t)n_limbs * (size_t)n_batch; }

    // s per candidate as derived on the device (N - 1 = 2^s d).
    std::vector<int32_t> s_values() const
    {
        std::vector<int32_t> s((size_t)n_batch, 0);
        if (mr_ctx_info(ctx, nullptr, nullptr, nullptr, s.data()) != 0)
            throw std::runtime_error("mr_ctx_info failed");
        return s;
    }

private:
    void init(const std::vector<uint64_t> &N_all, int n_limbs_, int n_batch_, int device_id_)
    {
        n_limbs = n_limbs_;
        n_batch = n_batch_;
        device_id = device_id_;
        if (n_limbs_ < 1 || n_batch_ < 1 || N_all.size() < size())
            throw std::runtime_error("BatchModCtx: N_all smaller than n_limbs * n_batch");
        N_host.assign(N_all.begin(), N_all.begin() + (long)size());
        ctx = mr_ctx_create_limbs16(N_all.data(), n_limbs_, n_batch_, device_id_);
        if (!ctx) throw std::runtime_error("BatchModCtx: context creation failed");
        int L = 0;
        if (mr_ctx_info(ctx, nullptr, &L, nullptr, nullptr) != 0 || L > n_limbs) {
            mr_ctx_destroy(ctx);
            ctx = nullptr;
            throw std::runtime_error("BatchModCtx: moduli too small for the uint64 working form "
                                     "(24-bit limb count exceeds n_limbs)");
        }
    }
};

namespace pfamd_detail {

// N - 1 and (N - 1) >> s of candidate i, 16-bit limbs.
inline void nm1_and_d(const BatchModCtx &m, int i, int s, std::vector<uint64_t> &nm1,
                      std::vector<uint64_t> &d)
{
    const size_t nl = (size_t)m.n_limbs;
    nm1.assign(m.N_host.begin() + (long)(i * nl), m.N_host.begin() + (long)((i + 1) * nl));
    for (size_t k = 0; k < nl; k++) {                 // N odd: no borrow past limb 0
        if (nm1[k] > 0) { nm1[k] -= 1; break; }
        nm1[k] = 0xFFFF;
    }
    d.assign(nl, 0u);
    const size_t q = (size_t)s / 16, r = (size_t)s % 16;
    for (size_t k = 0; k + q < nl; k++) {
        uint64_t v = nm1[k + q] >> r;
        if (r && k + q + 1 < nl) v |= (nm1[k + q + 1] << (16 - r)) & 0xFFFF;
        d[k] = v & 0xFFFF;
    }
}

// The caller's s, d (exp_all) and N - 1 (Nm1_all) must be those of N.
inline void check_args(const BatchModCtx &m, const std::vector<uint64_t> &exp_all,
                       const std::vector<uint64_t> &Nm1_all, int s, const char *who)
{
    const std::vector<int32_t> sv = m.s_values();
    const size_t nl = (size_t)m.n_limbs;
    if (!exp_all.empty() && exp_all.size() < m.size())
        throw std::runtime_error(std::string(who) + ": exp_all smaller than n_batch * n_limbs");
    if (!Nm1_all.empty() && Nm1_all.size() < m.size())
        throw std::runtime_error(std::string(who) + ": Nm1_all smaller than n_batch * n_limbs");
    std::vector<uint64_t> nm1, d;
    for (int i = 0; i < m.n_batch; i++) {
        bool tiny = m.N_host[(size_t)i * nl] <= 3u;              // N < 4 or even: no rounds run
        for (size_t k = 1; k < nl && tiny; k++) tiny = m.N_host[(size_t)i * nl + k] == 0u;
        if (tiny || (m.N_host[(size_t)i * nl] & 1u) == 0u) continue;
        if (sv[(size_t)i] != s)
            throw std::runtime_error(std::string(who) + ": s = " + std::to_string(s) +
                                     " but candidate " + std::to_string(i) + " has s = " +
                                     std::to_string(sv[(size_t)i]));
        nm1_and_d(m, i, s, nm1, d);
        for (size_t k = 0; k < nl; k++) {
            if (!exp_all.empty() && exp_all[(size_t)i * nl + k] != d[k])
                throw std::runtime_error(std::string(who) + ": exp_all of candidate " +
                                         std::to_string(i) + " is not (N - 1) / 2^s");
            if (!Nm1_all.empty() && Nm1_all[(size_t)i * nl + k] != nm1[k])
                throw std::runtime_error(std::string(who) + ": Nm1_all of candidate " +
                                         std::to_string(i) + " is not N - 1");
        }
    }
}

inline std::vector<bool> run_rounds(BatchModCtx &mont, int n_total,
                                    const std::vector<uint32_t> &witnesses, const char *label,
                                    bool show_report)
{
    if (n_total != mont.n_batch)
        throw std::runtime_error("gpu_miller_rabin: n_total differs from the context's batch");
    std::vector<uint8_t> alive((size_t)n_total, 0);
    const auto t0 = std::chrono::steady_clock::now();
    if (mr_run(mont.ctx, witnesses.data(), (int)witnesses.size(), alive.data(), nullptr) != 0)
        throw std::runtime_error("gpu_miller_rabin: mr_run failed");
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::vector<bool> out((size_t)n_total);
    int passed = 0;
    for (int i = 0; i < n_total; i++) {
        out[(size_t)i] = alive[(size_t)i] != 0;
        passed += out[(size_t)i] ? 1 : 0;
    }
    if (show_report)
        std::printf("[%s] %d candidates, %zu witnesses: %d probable primes, %.3f s\n",
                    label ? label : "mr", n_total, witnesses.size(), passed, s);
    return out;
}

} // namespace pfamd_detail

// N - 1 = 2 d (s = 1 for every candidate).
inline std::vector<bool> gpu_miller_rabin_s1(BatchModCtx &mont, const std::vector<uint64_t> &exp_all,
                                             const std::vector<uint64_t> &Nm1_all, int n_total,
                                             const std::vector<uint32_t> &witnesses,
                                             const char *label, bool show_report = false,
                                             bool show_progress = false)
{
    (void)show_progress;
    pfamd_detail::check_args(mont, exp_all, Nm1_all, 1, "gpu_miller_rabin_s1");
    return pfamd_detail::run_rounds(mont, n_total, witnesses, label, show_report);
}

// N - 1 = 2^s d with one s for the whole batch (bench_mr_gpu.cu:59).
inline std::vector<bool> gpu_miller_rabin(BatchModCtx &mont, const std::vector<uint64_t> &exp_all,
                                          const std::vector<uint64_t> &Nm1_all, int s, int n_total,
                                          const std::vector<uint32_t> &witnesses, const char *label,
                                          bool show_report = false, bool show_progress = false)
{
    (void)show_progress;
    pfamd_detail::check_args(mont, exp_all, Nm1_all, s, "gpu_miller_rabin");
    return pfamd_detail::run_rounds(mont, n_total, witnesses, label, show_report);
}
