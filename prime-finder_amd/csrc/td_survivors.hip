// td_survivors.hip -- survivor extraction on the device: the pairs (k, j)
// whose bit of a trial-division bitmap is clear, in the lexicographic order
// of worker.c:45-54 (k outer, j inner), so that Phase-1 survivors can feed
// Phase 2 without a host pass over the bitmap.
#include "pf_hip.h"
#include "../../include/trial_div.h"

#include <hipcub/hipcub.hpp>
#include <stdio.h>

namespace {

__device__ __forceinline__ uint32_t surv_mask(uint32_t w, uint64_t total)
{
    const uint64_t lo = (uint64_t)w * 32u;
    if (lo + 32u <= total) return 0xFFFFFFFFu;
    const uint32_t n = (uint32_t)(total - lo);      // 1..31
    return (1u << n) - 1u;
}

__global__ void surv_count_kernel(const uint32_t *__restrict__ bits, uint64_t total, uint32_t n_words,
                                  uint32_t *__restrict__ cnt)
{
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w < n_words) cnt[w] = (uint32_t)__popc(~bits[w] & surv_mask(w, total));
}

__global__ void surv_write_kernel(const uint32_t *__restrict__ bits, uint64_t total, uint32_t n_words,
                                  uint32_t J, const uint32_t *__restrict__ off,
                                  uint32_t *__restrict__ out, uint32_t cap)
{
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n_words) return;
    uint32_t m = ~bits[w] & surv_mask(w, total);
    uint32_t o = off[w];
    while (m) {
        const uint32_t i = (uint32_t)__ffs((int)m) - 1u;
        m &= m - 1u;
        const uint32_t b = w * 32u + i;
        if (o < cap) out[o] = ((b / J + 1u) << 16) | (b % J + 1u);
        o++;
    }
}

int surv_fail(hipError_t e, const char *what)
{
    fprintf(stderr, "trial_div_survivors: HIP error %s (%s)\n", hipGetErrorName(e), what);
    return -1;
}

} // namespace

extern "C" int trial_div_survivors(const uint32_t *d_bits, uint32_t K, uint32_t J, uint32_t *d_out,
                                   uint32_t cap, uint32_t *count, void *stream)
{
    if (!d_bits || !count || K < 1u || J < 1u || K > 9999u || J > 9999u || (cap && !d_out)) return -1;
    hipStream_t st = (hipStream_t)stream;
    const uint64_t total = (uint64_t)K * J;
    const uint32_t n_words = (uint32_t)((total + 31u) / 32u);
    uint32_t *cnt = nullptr, *off = nullptr;
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    hipError_t e = hipMalloc(&cnt, 2 * sizeof(uint32_t) * n_words);
    if (e != hipSuccess) return surv_fail(e, "alloc");
    off = cnt + n_words;
    int rc = 0;
    const uint32_t grid = (n_words + 255u) / 256u;
    surv_count_kernel<<<grid, 256, 0, st>>>(d_bits, total, n_words, cnt);
    if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, cnt, off, (int)n_words, st)) != hipSuccess ||
        (e = hipMalloc(&tmp, tmp_bytes ? tmp_bytes : 4)) != hipSuccess ||
        (e = hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, cnt, off, (int)n_words, st)) != hipSuccess) {
        rc = surv_fail(e, "scan");
    }
    if (rc == 0) {
        surv_write_kernel<<<grid, 256, 0, st>>>(d_bits, total, n_words, J, off, d_out, cap);
        uint32_t last[2] = {0, 0};
        if ((e = hipGetLastError()) != hipSuccess ||
            (e = hipMemcpyAsync(&last[0], off + n_words - 1, 4, hipMemcpyDeviceToHost, st)) != hipSuccess ||
            (e = hipMemcpyAsync(&last[1], cnt + n_words - 1, 4, hipMemcpyDeviceToHost, st)) != hipSuccess ||
            (e = hipStreamSynchronize(st)) != hipSuccess)
            rc = surv_fail(e, "write");
        else
            *count = last[0] + last[1];
    }
    if (tmp) (void)hipFree(tmp);
    (void)hipFree(cnt);
    return rc;
}
