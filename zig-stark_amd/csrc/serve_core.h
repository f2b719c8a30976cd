// serve_core.h -- the host-side core of the serving loop: a request queue,
// admission against a device-memory budget, a first-fit arena over that budget,
// worker threads, tickets answered once, and a shutdown that serves everything
// queued and lets blocked waiters leave before returning.
//
// It holds no HIP code: a request's device work is a callback run on a worker
// thread with the request's arena offset, so the same code serves the GPU
// (job.cpp: zs_server_*) and a CPU stress test that runs it under
// ThreadSanitizer and AddressSanitizer (tests/c/serve_core_stress.cpp).
#pragma once
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace zs {

inline double steady_seconds() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over one device allocation (offsets, 256-byte granules);
// free blocks are kept coalesced.
class Arena {
  public:
    static uint64_t round(uint64_t bytes) { return (bytes + 255) & ~uint64_t(255); }
    void reset(uint64_t size) {
        free_.clear();
        if (size) free_[0] = size;
    }
    bool alloc(uint64_t bytes, uint64_t *off) {
        bytes = round(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            uint64_t rest = it->second - bytes, at = it->first + bytes;
            free_.erase(it);
            if (rest) free_[at] = rest;
            return true;
        }
        return false;
    }
    bool fits(uint64_t bytes) const {
        bytes = round(bytes);
        for (auto &f : free_)
            if (f.second >= bytes) return true;
        return false;
    }
    void release(uint64_t off, uint64_t bytes) {
        bytes = round(bytes);
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

  private:
    std::map<uint64_t, uint64_t> free_;
};

// Admission order: arrival order, except that a request may pass earlier ones
// whose memory is not free yet; each earlier request can be passed at most
// kMaxBypass times, after which nothing behind it is admitted until it fits, so
// a large request cannot starve under a stream of small ones.
constexpr uint32_t kMaxBypass = 8;

template <class T>
class ServeCore {
  public:
    struct Item {
        uint64_t id = 0, bytes = 0, off = 0;
        uint32_t bypassed = 0;  // later requests admitted while this one did not fit
        bool done = false;
        bool claimed = false;  // a wait or poll has taken this ticket
        double t_submit = 0, t_admit = 0, t_done = 0;
        T data;
    };
    // the request's work, on worker `worker`, with its memory at [off, off + bytes)
    using Run = std::function<void(size_t worker, Item &)>;

    ServeCore(uint64_t budget, Run run) : budget_(Arena::round(budget)), run_(std::move(run)) { arena_.reset(budget_); }
    ~ServeCore() { stop(); }
    ServeCore(const ServeCore &) = delete;
    ServeCore &operator=(const ServeCore &) = delete;

    uint64_t budget() const { return budget_; }

    // n_workers threads, each calling on_start(index) before it serves
    void start(size_t n_workers, std::function<void(size_t)> on_start) {
        for (size_t i = 0; i < n_workers; i++)
            threads_.emplace_back([this, i, on_start] {
                if (on_start) on_start(i);
                loop(i);
            });
    }

    // false: the request alone exceeds the budget (it is dropped)
    bool submit(std::unique_ptr<Item> it, uint64_t *ticket) {
        if (Arena::round(it->bytes) > budget_) return false;
        it->t_submit = steady_seconds();
        {
            std::lock_guard<std::mutex> lk(mu_);
            it->id = next_id_++;
            *ticket = it->id;
            queue_.push_back(it.get());
            items_[it->id] = std::move(it);
        }
        cv_work_.notify_all();
        return true;
    }

    // Block until the request is answered and take it.  A ticket is answered once:
    // the first wait or poll claims it; any later or concurrent one gets false.
    bool wait(uint64_t ticket, std::unique_ptr<Item> *out) {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = items_.find(ticket);
        if (it == items_.end() || it->second->claimed) return false;
        Item *q = it->second.get();
        q->claimed = true;
        waiters_++;
        cv_done_.wait(lk, [&] { return q->done; });
        *out = std::move(items_[ticket]);
        items_.erase(ticket);
        waiters_--;
        if (stopping_) cv_done_.notify_all();  // stop() waits for the last waiter to leave
        return true;
    }

    // Non-blocking: *ready and the request (taken) if answered.  false: unknown or claimed ticket.
    bool poll(uint64_t ticket, std::unique_ptr<Item> *out, bool *ready) {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = items_.find(ticket);
        if (it == items_.end() || it->second->claimed) return false;
        *ready = it->second->done;
        if (*ready) {
            *out = std::move(it->second);
            items_.erase(it);
        }
        return true;
    }

    void memory(uint64_t *in_use, uint64_t *peak) {
        std::lock_guard<std::mutex> lk(mu_);
        if (in_use) *in_use = in_use_;
        if (peak) *peak = peak_;
    }

    // Serve every queued request, join the workers, and return once no thread is
    // inside wait().  No call may start after stop().
    void stop() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (stopped_) return;
            stopping_ = true;
        }
        cv_work_.notify_all();
        for (auto &t : threads_) t.join();
        threads_.clear();
        std::unique_lock<std::mutex> lk(mu_);
        cv_done_.notify_all();
        cv_done_.wait(lk, [&] { return waiters_ == 0; });
        stopped_ = true;
    }

  private:
    // the queued request to admit next, or queue_.end(); called with mu_ held
    typename std::deque<Item *>::iterator admissible() {
        for (auto it = queue_.begin(); it != queue_.end(); ++it) {
            if (arena_.fits((*it)->bytes)) return it;
            if ((*it)->bypassed >= kMaxBypass) break;  // this one goes next; nothing passes it
        }
        return queue_.end();
    }

    void loop(size_t worker) {
        for (;;) {
            Item *q = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_work_.wait(lk, [&] { return (stopping_ && queue_.empty()) || admissible() != queue_.end(); });
                if (stopping_ && queue_.empty()) return;
                auto it = admissible();
                if (it == queue_.end() || !arena_.alloc((*it)->bytes, &(*it)->off)) continue;
                for (auto e = queue_.begin(); e != it; ++e) (*e)->bypassed++;
                q = *it;
                queue_.erase(it);
                in_use_ += Arena::round(q->bytes);
                peak_ = std::max(peak_, in_use_);
                q->t_admit = steady_seconds();
            }
            run_(worker, *q);  // the item stays owned by items_ and unclaimed-until-done for waiters
            {
                std::lock_guard<std::mutex> lk(mu_);
                q->t_done = steady_seconds();
                q->done = true;
                in_use_ -= Arena::round(q->bytes);
                arena_.release(q->off, q->bytes);
            }
            cv_done_.notify_all();
            cv_work_.notify_all();
        }
    }

    const uint64_t budget_;
    Run run_;
    Arena arena_;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<Item *> queue_;
    std::map<uint64_t, std::unique_ptr<Item>> items_;
    std::vector<std::thread> threads_;
    uint64_t next_id_ = 1, in_use_ = 0, peak_ = 0;
    uint32_t waiters_ = 0;
    bool stopping_ = false, stopped_ = false;
};

}  // namespace zs
