// Stress test of the serving loop's host core (zig-stark_amd/csrc/serve_core.h),
// the code zs_server_* runs, on the CPU with a stand-in for the device work.
// tests/test_serve_core.py builds it with ThreadSanitizer and with
// AddressSanitizer + UndefinedBehaviorSanitizer and runs it.  Checks:
//   * every admitted request's [off, off + bytes) lies in the budget and overlaps
//     no other request in flight, and the admitted bytes never exceed the budget;
//   * every ticket is answered once, with its own result, through wait or poll,
//     also when two threads wait on the same ticket at once;
//   * a large request behind a stream of small ones is passed at most kMaxBypass times;
//   * stop() serves what is queued and returns only after blocked waiters left.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "../../zig-stark_amd/csrc/serve_core.h"

using zs::Arena;
using zs::ServeCore;

namespace {

struct Payload {
    uint64_t seed = 0, result = 0;
    uint32_t sleep_us = 0;
};
using Core = ServeCore<Payload>;

std::atomic<int> failures{0};
#define CHECK(c)                                                          \
    do {                                                                  \
        if (!(c)) {                                                       \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            failures++;                                                   \
        }                                                                 \
    } while (0)

uint64_t mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// the memory ranges of the requests in flight
struct Ranges {
    std::mutex mu;
    std::map<uint64_t, uint64_t> live;  // off -> end
    uint64_t budget;
    void enter(uint64_t off, uint64_t bytes) {
        std::lock_guard<std::mutex> lk(mu);
        const uint64_t end = off + Arena::round(bytes);
        CHECK(end <= budget);
        auto nx = live.lower_bound(off);
        if (nx != live.end()) CHECK(nx->first >= end);
        if (nx != live.begin()) CHECK(std::prev(nx)->second <= off);
        live[off] = end;
    }
    void leave(uint64_t off) {
        std::lock_guard<std::mutex> lk(mu);
        CHECK(live.erase(off) == 1);
    }
};

Core::Run runner(Ranges &rg) {
    return [&rg](size_t, Core::Item &it) {
        rg.enter(it.off, it.bytes);
        it.data.result = mix(it.data.seed);
        if (it.data.sleep_us) std::this_thread::sleep_for(std::chrono::microseconds(it.data.sleep_us));
        rg.leave(it.off);
    };
}

std::unique_ptr<Core::Item> item(uint64_t bytes, uint64_t seed, uint32_t sleep_us) {
    auto it = std::make_unique<Core::Item>();
    it->bytes = bytes;
    it->data.seed = seed;
    it->data.sleep_us = sleep_us;
    return it;
}

void random_load() {
    const uint64_t budget = 1 << 20;
    Ranges rg;
    rg.budget = Arena::round(budget);
    Core core(budget, runner(rg));
    core.start(4, nullptr);
    std::vector<std::thread> th;
    for (int t = 0; t < 4; t++) {
        th.emplace_back([&, t] {
            std::mt19937_64 rng(t + 1);
            std::vector<std::pair<uint64_t, uint64_t>> mine;  // ticket, seed
            for (int k = 0; k < 300; k++) {
                const uint64_t seed = (uint64_t(t) << 32) | k, ticket_bytes = 256 + rng() % 300000;
                uint64_t ticket = 0;
                CHECK(core.submit(item(ticket_bytes, seed, rng() % 50), &ticket));
                mine.push_back({ticket, seed});
                uint64_t in_use = 0, peak = 0;  // read while the workers admit and release
                core.memory(&in_use, &peak);
                CHECK(in_use <= peak && peak <= Arena::round(budget));
            }
            uint64_t refused_ticket = 0;
            CHECK(!core.submit(item(budget + 4096, 0, 0), &refused_ticket));  // more than the budget
            for (size_t k = 0; k < mine.size(); k++) {
                std::unique_ptr<Core::Item> got;
                if (k % 2 == 0) {
                    CHECK(core.wait(mine[k].first, &got));
                } else {
                    bool ready = false;
                    for (;;) {
                        CHECK(core.poll(mine[k].first, &got, &ready));
                        if (ready) break;
                        std::this_thread::yield();
                    }
                }
                CHECK(got && got->data.result == mix(mine[k].second));
                CHECK(got->t_submit <= got->t_admit && got->t_admit <= got->t_done);
                std::unique_ptr<Core::Item> again;
                bool ready = false;
                CHECK(!core.wait(mine[k].first, &again));  // answered once
                CHECK(!core.poll(mine[k].first, &again, &ready));
            }
        });
    }
    for (auto &x : th) x.join();
    uint64_t in_use = 1, peak = 0;
    core.memory(&in_use, &peak);
    CHECK(in_use == 0);
    CHECK(peak <= Arena::round(budget) && peak > 0);
}

void concurrent_waiters() {
    Ranges rg;
    rg.budget = 1 << 16;
    Core core(1 << 16, runner(rg));
    core.start(1, nullptr);
    for (int round = 0; round < 50; round++) {
        uint64_t ticket = 0;
        CHECK(core.submit(item(4096, round, 200), &ticket));
        std::atomic<int> ok{0}, refused{0};
        std::vector<std::thread> th;
        for (int k = 0; k < 3; k++)
            th.emplace_back([&] {
                std::unique_ptr<Core::Item> got;
                if (core.wait(ticket, &got)) {
                    CHECK(got->data.result == mix(uint64_t(round)));
                    ok++;
                } else {
                    refused++;
                }
            });
        for (auto &x : th) x.join();
        CHECK(ok == 1 && refused == 2);
    }
}

void no_starvation() {
    const uint64_t small = 4096, large = 60000;
    const uint64_t budget = Arena::round(large) + Arena::round(small) - 256;  // the large one needs every small one done
    Ranges rg;
    rg.budget = budget;
    Core core(budget, runner(rg));
    core.start(2, nullptr);
    std::vector<uint64_t> before(2), after(40);
    uint64_t big = 0;
    for (auto &t : before) CHECK(core.submit(item(small, 1, 2000), &t));
    CHECK(core.submit(item(large, 2, 2000), &big));
    for (auto &t : after) CHECK(core.submit(item(small, 3, 2000), &t));
    std::unique_ptr<Core::Item> rb, x;
    CHECK(core.wait(big, &rb));
    int passed = 0, later = 0;
    for (uint64_t t : after) {
        CHECK(core.wait(t, &x));
        if (x->t_admit < rb->t_admit) passed++;
        else later++;
    }
    for (uint64_t t : before) CHECK(core.wait(t, &x));
    std::fprintf(stderr, "starvation: %d of %zu later small requests passed the large one\n", passed, after.size());
    CHECK(passed <= (int)zs::kMaxBypass);
    CHECK(later > 0);
}

void stop_with_blocked_waiters() {
    Ranges rg;
    rg.budget = 1 << 20;
    auto core = std::make_unique<Core>(1 << 20, runner(rg));
    core->start(2, nullptr);
    std::vector<uint64_t> tickets(8);
    for (size_t k = 0; k < tickets.size(); k++) CHECK(core->submit(item(8192, k, 3000), &tickets[k]));
    std::atomic<int> answered{0};
    std::vector<std::thread> th;
    for (uint64_t t : tickets)
        th.emplace_back([&, t] {
            std::unique_ptr<Core::Item> got;
            if (core->wait(t, &got)) answered++;
        });
    std::this_thread::sleep_for(std::chrono::milliseconds(2));
    core->stop();  // serves the queue, returns after every waiter inside wait() left
    for (auto &x : th) x.join();
    CHECK(answered == (int)tickets.size());
    uint64_t in_use = 1;
    core->memory(&in_use, nullptr);
    CHECK(in_use == 0);
    core.reset();
}

}  // namespace

int main() {
    random_load();
    concurrent_waiters();
    no_starvation();
    stop_with_blocked_waiters();
    if (failures) {
        std::fprintf(stderr, "%d checks failed\n", failures.load());
        return 1;
    }
    std::printf("ok\n");
    return 0;
}
