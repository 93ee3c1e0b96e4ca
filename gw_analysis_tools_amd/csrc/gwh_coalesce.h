/* Coalescing of concurrent single requests into shared batches (host code only, no HIP).
 *
 * Each calling thread queues one item and returns when the item is done.  A caller that finds one of the
 * Lanes free leads: it takes every queued item of the queue head's group (in arrival order, up to max_batch)
 * and runs them as one batch on that lane through the caller's run(lane, batch) with no lock held, then marks
 * them done.  Callers still queued poll their item (and the lanes) for `spin`, then block on a condition
 * variable; only blocked callers are woken.  A finished lane hands over to the queue head, so some caller leads
 * whenever items are queued and a lane is free.
 *
 * The likelihood engine uses it for the reference-signature single calls, which samplers make from many
 * threads at once (mcmc_sampler.cpp:4585-4600); tests/host/coalesce_check.cpp runs it under the thread
 * sanitizer with a host-only run(). */
#pragma once

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

namespace gwh {

struct CoalesceItem {
    int group = 0;              /* items of one group may share a batch */
    std::atomic<int> status{0}; /* 0 queued, 1 in a batch, 2 done (the item's results are written before) */
    bool sleeping = false;      /* blocked on cv (coalescer lock held when read or written) */
    std::condition_variable cv;
};

template <int Lanes> class Coalescer {
  public:
    using Batch = std::vector<CoalesceItem *>;
    explicit Coalescer(int max_batch = 4096, std::chrono::microseconds spin = std::chrono::microseconds(200))
        : max_batch_(max_batch), spin_(spin)
    {
        free_lanes_.store(Lanes);
    }

    /* queue `me`, lead batches while a lane is free, return once `me` is done.  run(lane, batch) runs one
     * batch (it may shrink `batch` to the items it ran; every item taken is marked done afterwards). */
    template <class Run> void submit(CoalesceItem &me, Run &&run)
    {
        std::unique_lock<std::mutex> lk(m_);
        queue_.push_back(&me);
        queued_.store((int)queue_.size(), std::memory_order_relaxed);
        for (;;) {
            if (me.status.load(std::memory_order_acquire) == 2) break;
            const int lane = free_lane();
            if (lane >= 0 && !queue_.empty()) {
                set_busy(lane, true);
                Batch batch;
                const int group = queue_.front()->group;
                for (auto it = queue_.begin(); it != queue_.end();) {
                    if ((*it)->group == group && (int)batch.size() < max_batch_) {
                        (*it)->status.store(1, std::memory_order_relaxed);
                        batch.push_back(*it);
                        it = queue_.erase(it);
                    } else {
                        ++it;
                    }
                }
                queued_.store((int)queue_.size(), std::memory_order_relaxed);
                if (free_lane() >= 0 && !queue_.empty()) wake(queue_.front()); /* another group: another lane */
                const Batch taken = batch;
                lk.unlock();
                run(lane, batch);
                lk.lock();
                set_busy(lane, false);
                launches_++;
                items_ += (long long)batch.size();
                if ((long long)batch.size() > max_merged_) max_merged_ = (long long)batch.size();
                for (CoalesceItem *q : taken) {
                    const bool blocked = q->sleeping; /* read before the release: its owner may return at once */
                    q->status.store(2, std::memory_order_release);
                    if (blocked && q != &me) q->cv.notify_one();
                }
                if (!queue_.empty()) wake(queue_.front()); /* a lane is free: the next head leads */
                continue;
            }
            /* poll without the lock first: no system call on either side while batches turn over quickly */
            lk.unlock();
            const auto until = std::chrono::steady_clock::now() + spin_;
            bool again = false;
            while (std::chrono::steady_clock::now() < until) {
                if (me.status.load(std::memory_order_acquire) == 2 ||
                    (free_lanes_.load(std::memory_order_acquire) > 0 && queued_.load(std::memory_order_relaxed) > 0)) {
                    again = true;
                    break;
                }
                std::this_thread::yield();
            }
            lk.lock();
            if (again || me.status.load(std::memory_order_acquire) == 2 || (free_lane() >= 0 && !queue_.empty()))
                continue;
            me.sleeping = true;
            me.cv.wait(lk);
            me.sleeping = false;
        }
    }

    /* launches, items launched, most items in one launch, items queued now */
    void stats(long long *launches, long long *items, long long *max_merged, int *queued)
    {
        std::lock_guard<std::mutex> lk(m_);
        *launches = launches_;
        *items = items_;
        *max_merged = max_merged_;
        *queued = (int)queue_.size();
    }

  private:
    int free_lane() const
    {
        for (int l = 0; l < Lanes; l++)
            if (!busy_[l]) return l;
        return -1;
    }
    void set_busy(int lane, bool b)
    {
        busy_[lane] = b;
        int f = 0;
        for (int l = 0; l < Lanes; l++) f += busy_[l] ? 0 : 1;
        free_lanes_.store(f, std::memory_order_release);
    }
    static void wake(CoalesceItem *q)
    {
        if (q->sleeping) q->cv.notify_one();
    }

    std::mutex m_;
    std::deque<CoalesceItem *> queue_;
    bool busy_[Lanes] = {};
    std::atomic<int> free_lanes_{0};
    std::atomic<int> queued_{0};
    const int max_batch_;
    const std::chrono::microseconds spin_;
    long long launches_ = 0, items_ = 0, max_merged_ = 0;
};

} // namespace gwh
