// mr_cpp_check -- exercises include/batch_mod_ctx.hpp the way the
// reference's bench_mr_gpu.cu:44-64 drives its runner: 16-bit limbs in
// uint64_t, one context per batch, verdicts per candidate.  Prints one line
// per candidate: "<index> <verdict>".  Input on stdin: n_limbs n_batch,
// then n_batch * n_limbs limbs.  Built beside the library; used by the
// GPU tests.
#include "batch_mod_ctx.hpp"

#include <cstdio>
#include <vector>

int main()
{
    int n_limbs = 0, n_batch = 0;
    if (std::scanf("%d %d", &n_limbs, &n_batch) != 2 || n_limbs < 1 || n_batch < 1) return 2;
    std::vector<uint64_t> N((size_t)n_limbs * n_batch);
    for (auto &x : N)
        if (std::scanf("%lu", &x) != 1) return 2;
    try {
        BatchModCtx mont(N, n_limbs, n_batch);
        std::vector<uint64_t> unused;
        const auto v = gpu_miller_rabin(mont, unused, unused, 1, n_batch, DEFAULT_WITNESSES, "cpp", true);
        for (int i = 0; i < n_batch; i++) std::printf("%d %d\n", i, v[(size_t)i] ? 1 : 0);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "mr_cpp_check: %s\n", e.what());
        return 1;
    }
    return 0;
}
