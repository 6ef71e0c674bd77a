// td_work.cpp -- the algorithmic work of one trial-division launch, counted
// on the host, for rating the kernels against the INT32 roof.  The lattice
// path is counted by running its per-prime routine (td_lattice.h, the code
// the lattice kernel runs) with counting hooks: Montgomery multiplications,
// first-hit queries, continued-fraction levels, hits, the division steps that
// compute the return times, the return-time steps between hits and the
// direct steps of the one-digit rev class.  The k-loop paths (primes from 2^16 below the
// lattice range, or every prime from 2^16 when the j digit groups do not
// merge) are counted as prime x k steps.  No GPU call.
#include <stdint.h>
#include <string.h>

#include <thread>
#include <vector>

namespace {
struct TdWork {
    uint64_t prime, modmul, query, level, hit, step, gapstep, gapcf, modmul_lazy;
};
thread_local TdWork *td_work_cur;
}  // namespace

#define PF_MONT_COUNT() (++td_work_cur->modmul)
#define PF_MONT_COUNT_LAZY() (++td_work_cur->modmul_lazy)
#define TD_LAT_COUNT(what) (++td_work_cur->what)
#include "td_lattice.h"
#include "../../include/trial_div.h"

namespace {
struct NoMark {
    void operator()(uint32_t) const {}
};

template <int DJ>
void td_work_range(const uint32_t *primes, size_t lo, size_t hi, size_t stride, uint32_t lmin,
                   long long a, long long b, long long c, uint32_t K, uint32_t J, TdWork &w)
{
    td_work_cur = &w;
    NoMark mark;
    const uint32_t wtop = td_top_window(K);
    for (size_t i = lo; i < hi; i += stride) {
        const uint32_t p = primes[i];
        if (p < lmin || p >= 0x80000000u) continue;
        w.prime++;
        td_lattice_prime<DJ>(p, a, b, c, K, J, wtop, mark);
    }
    td_work_cur = nullptr;
}
}  // namespace

extern "C" int trial_div_work(const uint32_t *primes, uint32_t n_primes, long long a,
                              long long b, long long c, uint32_t K, uint32_t J,
                              uint32_t threads, uint64_t out[TD_WORK_FIELDS])
{
    if (!out || (n_primes && !primes) || K < 1 || K > 9999 || J < 1 || J > 9999 || a < 1 ||
        b < 0 || c < 0 || b >= a || c >= a)
        return -1;
    memset(out, 0, TD_WORK_FIELDS * sizeof(uint64_t));
    const int DJ = J >= 1000u ? 4 : J >= 100u ? 3 : J >= 10u ? 2 : 1;
    const bool merged = a - c >= DJ;
    const uint32_t lmin = merged ? trial_div_lattice_min() : 0x80000000u;
    uint64_t ksteps = 0;
    for (uint32_t i = 0; i < n_primes; i++)
        if (primes[i] >= (1u << 16) && primes[i] < lmin) ksteps += K;
    out[TD_WORK_KSTEP] = ksteps;
    if (!merged) return 0;
    const unsigned nt = threads ? threads : 1u;
    std::vector<TdWork> part(nt, TdWork{0, 0, 0, 0, 0, 0, 0, 0, 0});
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; t++)
        pool.emplace_back([&, t] {
            switch (DJ) {
            case 1: td_work_range<1>(primes, t, n_primes, nt, lmin, a, b, c, K, J, part[t]); break;
            case 2: td_work_range<2>(primes, t, n_primes, nt, lmin, a, b, c, K, J, part[t]); break;
            case 3: td_work_range<3>(primes, t, n_primes, nt, lmin, a, b, c, K, J, part[t]); break;
            default: td_work_range<4>(primes, t, n_primes, nt, lmin, a, b, c, K, J, part[t]); break;
            }
        });
    for (auto &th : pool) th.join();
    for (const TdWork &w : part) {
        out[TD_WORK_PRIME] += w.prime;
        out[TD_WORK_MODMUL] += w.modmul;
        out[TD_WORK_QUERY] += w.query;
        out[TD_WORK_LEVEL] += w.level;
        out[TD_WORK_HIT] += w.hit;
        out[TD_WORK_STEP] += w.step;
        out[TD_WORK_GAPSTEP] += w.gapstep;
        out[TD_WORK_GAPCF] += w.gapcf;
        out[TD_WORK_MODMUL_LAZY] += w.modmul_lazy;
    }
    return 0;
}
