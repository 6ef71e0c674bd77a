// mr_cpp_check -- exercises include/batch_mod_ctx.hpp the way the
// reference's C++ callers drive its runner: 16-bit limbs in uint64_t, one
// context per batch.  Built beside the library; used by the GPU tests.
//
//   mr_cpp_check verdicts   stdin: n_limbs n_batch, then the moduli's limbs;
//                           bench_mr_gpu.cu:44-64: s and d derived from N on
//                           the host and passed in, DEFAULT_WITNESSES;
//                           prints "<index> <verdict>" per candidate.
//   mr_cpp_check arith      stdin: n_limbs n_batch, the moduli's limbs, then
//                           x and y (each < N) as limbs; the arithmetic of
//                           correctness_tests.cuh:217-218,318,353,387:
//                           to_residue_batch, modmul_batch, modsq_batch,
//                           check_passed, from_residue_batch; prints, per
//                           candidate, "<index> xy x^2 passed(R) passed(x)"
//                           with the values in hex.
//   mr_cpp_check badargs    stdin as "verdicts": a wrong s and a wrong
//                           exp_all must throw; prints "ok".
#include "batch_mod_ctx.hpp"
#include "pf_runtime.h"

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

bool read_limbs(std::vector<uint64_t> &v, size_t n)
{
    v.resize(n);
    for (auto &x : v)
        if (std::scanf("%lu", &x) != 1) return false;
    return true;
}

// (N - 1) >> s for every candidate, 16-bit limbs (the caller's own d).
std::vector<uint64_t> host_d(const std::vector<uint64_t> &N, int nl, int nb, int s)
{
    std::vector<uint64_t> d((size_t)nl * nb, 0u);
    for (int i = 0; i < nb; i++) {
        std::vector<uint64_t> t(N.begin() + (long)i * nl, N.begin() + (long)(i + 1) * nl);
        t[0] -= 1;   // odd N
        for (int k = 0; k < nl; k++) {
            const int q = s / 16, r = s % 16;
            if (k + q >= nl) break;
            uint64_t v = t[(size_t)(k + q)] >> r;
            if (r && k + q + 1 < nl) v |= (t[(size_t)(k + q + 1)] << (16 - r)) & 0xFFFF;
            d[(size_t)i * nl + k] = v & 0xFFFF;
        }
    }
    return d;
}

void print_hex(const uint64_t *v, int nl)
{
    int top = nl - 1;
    while (top > 0 && v[top] == 0) top--;
    std::printf(" %lx", (unsigned long)v[top]);
    for (int k = top - 1; k >= 0; k--) std::printf("%04lx", (unsigned long)v[k]);
}

struct DevBuf {
    void *p = nullptr;
    explicit DevBuf(size_t bytes) : p(pf_malloc(bytes))
    {
        if (!p) throw std::runtime_error("pf_malloc failed");
    }
    ~DevBuf() { pf_free(p); }
};

}  // namespace

int main(int argc, char **argv)
{
    const std::string mode = argc > 1 ? argv[1] : "verdicts";
    int n_limbs = 0, n_batch = 0;
    if (std::scanf("%d %d", &n_limbs, &n_batch) != 2 || n_limbs < 1 || n_batch < 1) return 2;
    std::vector<uint64_t> N;
    if (!read_limbs(N, (size_t)n_limbs * n_batch)) return 2;
    try {
        BatchModCtx mont(N, n_limbs, n_batch);
        if (mode == "verdicts") {
            const std::vector<int32_t> sv = mont.s_values();
            const int s = sv[0];
            const auto d = host_d(N, n_limbs, n_batch, s);
            const auto v = gpu_miller_rabin(mont, d, std::vector<uint64_t>(), s, n_batch,
                                            DEFAULT_WITNESSES, "cpp", true);
            for (int i = 0; i < n_batch; i++) std::printf("%d %d\n", i, v[(size_t)i] ? 1 : 0);
        } else if (mode == "badargs") {
            const int s = mont.s_values()[0];
            int thrown = 0;
            try {
                (void)gpu_miller_rabin(mont, {}, {}, s + 1, n_batch, DEFAULT_WITNESSES, "bad");
            } catch (const std::runtime_error &) { thrown++; }
            auto d = host_d(N, n_limbs, n_batch, s);
            d[0] ^= 2u;
            try {
                (void)gpu_miller_rabin(mont, d, {}, s, n_batch, DEFAULT_WITNESSES, "bad");
            } catch (const std::runtime_error &) { thrown++; }
            std::printf(thrown == 2 ? "ok\n" : "not thrown\n");
        } else if (mode == "arith") {
            std::vector<uint64_t> X, Y;
            if (!read_limbs(X, mont.size()) || !read_limbs(Y, mont.size())) return 2;
            std::vector<uint64_t> xr, yr, one((size_t)mont.size(), 0u), oner;
            for (int i = 0; i < n_batch; i++) one[(size_t)i * n_limbs] = 1u;
            mont.to_residue_batch(X, xr);
            mont.to_residue_batch(Y, yr);
            mont.to_residue_batch(one, oner);
            const size_t bytes = sizeof(uint64_t) * mont.size();
            DevBuf dx(bytes), dy(bytes), dxy(bytes), dsq(bytes), d1(bytes), flags(2 * (size_t)n_batch);
            if (pf_memcpy_htod(dx.p, xr.data(), bytes) || pf_memcpy_htod(dy.p, yr.data(), bytes) ||
                pf_memcpy_htod(d1.p, oner.data(), bytes))
                throw std::runtime_error("upload failed");
            mont.modmul_batch((const uint64_t *)dx.p, (const uint64_t *)dy.p, (uint64_t *)dxy.p);
            mont.modsq_batch((const uint64_t *)dx.p, (uint64_t *)dsq.p);
            mont.check_passed((const uint64_t *)d1.p, (uint8_t *)flags.p);
            mont.check_passed((const uint64_t *)dx.p, (uint8_t *)flags.p + n_batch);
            if (pf_sync()) throw std::runtime_error("sync failed");
            std::vector<uint64_t> xy, sq;
            mont.from_residue_batch((const uint64_t *)dxy.p, xy);
            mont.from_residue_batch((const uint64_t *)dsq.p, sq);
            std::vector<uint8_t> f(2 * (size_t)n_batch);
            if (pf_memcpy_dtoh(f.data(), flags.p, f.size())) throw std::runtime_error("download failed");
            for (int i = 0; i < n_batch; i++) {
                std::printf("%d", i);
                print_hex(xy.data() + (size_t)i * n_limbs, n_limbs);
                print_hex(sq.data() + (size_t)i * n_limbs, n_limbs);
                std::printf(" %d %d\n", f[(size_t)i], f[(size_t)n_batch + i]);
            }
        } else {
            return 2;
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "mr_cpp_check: %s\n", e.what());
        return 1;
    }
    return 0;
}
