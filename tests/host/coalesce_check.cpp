/* Host-only check of the single-call coalescer (gw_analysis_tools_amd/csrc/gwh_coalesce.h), built with the thread
 * sanitizer (and separately with the address / undefined-behaviour sanitizers) by tests/test_host_sanitizers.py.
 *
 * T threads submit K items each, in two groups; run() stands in for a launch: it checks that no lane runs two
 * batches at once and that a batch holds one group, drops every 97th item (as the engine drops calls whose event
 * changed), computes each remaining item's result and sleeps a little.  Every item must come back done with its
 * own result, and concurrent calls must have shared batches.  Exit status 0 on success. */
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#include "gwh_coalesce.h"

struct Item : gwh::CoalesceItem {
    long long value = 0, out = -2;
};

int main(int argc, char **argv)
{
    const int T = argc > 1 ? atoi(argv[1]) : 48, K = argc > 2 ? atoi(argv[2]) : 300;
    gwh::Coalescer<2> co(64, std::chrono::microseconds(50));
    std::atomic<int> in_run[2];
    in_run[0] = 0;
    in_run[1] = 0;
    std::atomic<int> failures{0};
    auto run = [&](int lane, std::vector<gwh::CoalesceItem *> &batch) {
        if (lane < 0 || lane > 1 || in_run[lane].fetch_add(1) != 0) failures++;
        std::vector<gwh::CoalesceItem *> kept;
        for (auto *it : batch) {
            Item *q = static_cast<Item *>(it);
            if (q->group != batch[0]->group || q->status.load() != 1) failures++;
            if (q->value % 97 == 0) {
                q->out = -1; /* dropped: not launched */
                continue;
            }
            q->out = q->value * q->value + q->group;
            kept.push_back(it);
        }
        batch.swap(kept);
        std::this_thread::sleep_for(std::chrono::microseconds(20 + (int)(batch.size() % 7) * 5));
        in_run[lane].fetch_sub(1);
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < T; t++)
        pool.emplace_back([&, t] {
            for (int k = 0; k < K; k++) {
                Item me;
                me.value = (long long)t * 100000 + k;
                me.group = (t + k) % 2;
                co.submit(me, run);
                const long long want = (me.value % 97 == 0) ? -1 : me.value * me.value + me.group;
                if (me.status.load() != 2 || me.out != want) failures++;
            }
        });
    for (auto &th : pool) th.join();
    long long launches, items, most;
    int queued;
    co.stats(&launches, &items, &most, &queued);
    printf("threads %d items %d: %lld launches, %lld items launched, at most %lld per launch, %d queued, %d failures\n",
           T, T * K, launches, items, most, queued, failures.load());
    if (failures.load() != 0 || queued != 0 || most < 2) return 1;
    return 0;
}
