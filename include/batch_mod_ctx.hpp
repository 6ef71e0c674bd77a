/*
 * batch_mod_ctx.hpp -- the reference's C++ surface of the GPU Miller-Rabin
 * runner, header-only over the C ABI of mr_gpu.h (link libpfamd.so).
 *
 * Replaces:
 *   BatchModCtx(N_all, n_limbs, n_batch, device)   batch_mod_ctx.cuh:132-133
 *   gpu_miller_rabin_s1(...), gpu_miller_rabin(...) miller_rabin_runner.cuh:28-49
 *   DEFAULT_WITNESSES                              miller_rabin_runner.cuh:12-13
 * so that a caller such as bench_mr_gpu.cu:60-64 compiles unchanged against
 * it.  N_all holds 16-bit limbs, one per uint64_t, [n_batch][n_limbs]
 * (candidate.cuh:62-70).
 *
 * Differences, by design:
 *  - d and s come from N on the device, per candidate: exp_all, Nm1_all and
 *    the s argument are accepted for signature compatibility and not read
 *    (the reference's runners ignore Nm1_all too, miller_rabin_runner.cu:284,368).
 *  - a candidate retires at its first failing witness (same verdicts).
 *  - the arithmetic working form is this library's (24-bit limbs,
 *    R = 2^(24 L)); the reference's device-pointer operations
 *    (modmul_batch, modsq_batch, to_/from_residue_batch, check_passed on
 *    Data64 buffers) are not carried over.
 *  - errors throw std::runtime_error, as the reference's CU macros do.
 */
#pragma once
#include "mr_gpu.h"

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <vector>

inline const std::vector<uint32_t> DEFAULT_WITNESSES = {
    2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53};

struct BatchModCtx {
    int n_limbs = 0, n_batch = 0, device_id = 0;
    mr_ctx_t *ctx = nullptr;

    explicit BatchModCtx(const std::vector<uint64_t> &N_all, int n_limbs_, int n_batch_,
                         int device_id_ = 0)
        : n_limbs(n_limbs_), n_batch(n_batch_), device_id(device_id_)
    {
        if (n_limbs_ < 1 || n_batch_ < 1 || N_all.size() < (size_t)n_limbs_ * (size_t)n_batch_)
            throw std::runtime_error("BatchModCtx: N_all smaller than n_limbs * n_batch");
        ctx = mr_ctx_create_limbs16(N_all.data(), n_limbs_, n_batch_, device_id_);
        if (!ctx) throw std::runtime_error("BatchModCtx: context creation failed");
    }
    ~BatchModCtx() { mr_ctx_destroy(ctx); }
    BatchModCtx(const BatchModCtx &) = delete;
    BatchModCtx &operator=(const BatchModCtx &) = delete;
};

namespace pfamd_detail {

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

// N - 1 = 2 d (s = 1).
inline std::vector<bool> gpu_miller_rabin_s1(BatchModCtx &mont, const std::vector<uint64_t> &exp_all,
                                             const std::vector<uint64_t> &Nm1_all, int n_total,
                                             const std::vector<uint32_t> &witnesses,
                                             const char *label, bool show_report = false,
                                             bool show_progress = false)
{
    (void)exp_all;
    (void)Nm1_all;
    (void)show_progress;
    return pfamd_detail::run_rounds(mont, n_total, witnesses, label, show_report);
}

// N - 1 = 2^s d, any s >= 1 (s is derived per candidate).
inline std::vector<bool> gpu_miller_rabin(BatchModCtx &mont, const std::vector<uint64_t> &exp_all,
                                          const std::vector<uint64_t> &Nm1_all, int s, int n_total,
                                          const std::vector<uint32_t> &witnesses, const char *label,
                                          bool show_report = false, bool show_progress = false)
{
    (void)exp_all;
    (void)Nm1_all;
    (void)s;
    (void)show_progress;
    return pfamd_detail::run_rounds(mont, n_total, witnesses, label, show_report);
}
