/*
 * miller_rabin_gpu.c -- miller_rabin_test (include/miller_rabin.h) over the
 * batched Miller-Rabin C ABI: one context per number, N first and rev(N)
 * only when N passes, as miller_rabin.c:76-118 orders the two tests.
 */
#define _POSIX_C_SOURCE 199309L
#include <stdio.h>
#include <time.h>

#include "miller_rabin.h"
#include "mr_gpu.h"

static const uint32_t kBases[4] = {2, 3, 5, 7};

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* One number: build on the current device, run the four bases.
 * Returns 1 = passed, 0 = composite, -1 = error. */
static int test_one(long long a, long long b, long long c, int k, int j, int which,
                    double *t_build, double *t_run)
{
    mr_candidate_t cd = {a, b, c, (uint32_t)k, (uint32_t)j, which};
    double t0 = now_s();
    mr_ctx_t *ctx = mr_ctx_create_sparse(&cd, 1, -1);
    double t1 = now_s();
    *t_build = t1 - t0;
    if (!ctx) return -1;
    uint8_t alive = 0;
    int rc = mr_run(ctx, kBases, 4, &alive, NULL);
    mr_ctx_destroy(ctx);
    *t_run = now_s() - t1;
    return rc != 0 ? -1 : (alive ? 1 : 0);
}

int miller_rabin_test(long long a, long long b, long long c, int k, int j,
                      mr_timing_t *timing)
{
    mr_timing_t t = {0.0, 0.0, 0.0, 0.0};
    int code;
    if (k < 1 || j < 1) {
        code = -1;
    } else {
        int rn = test_one(a, b, c, k, j, 0, &t.build_n, &t.mr_n);
        if (rn <= 0) {
            code = rn;
        } else {
            int rr = test_one(a, b, c, k, j, 2, &t.build_revn, &t.mr_revn);
            code = rr < 0 ? -1 : 1 + rr;
        }
    }
    if (code < 0)
        fprintf(stderr, "miller_rabin_test: cannot test a=%lld b=%lld c=%lld k=%d j=%d\n",
                a, b, c, k, j);
    if (timing) *timing = t;
    return code;
}
