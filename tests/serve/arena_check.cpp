// CPU check of the serving runtime's HBM arena and admission policy
// (process_amd/csrc/psx_arena.h) against a byte-map model: blocks never
// overlap or leave the budget, freed memory coalesces back into one range,
// admission is in arrival order and never exceeds the budget, and a stream of
// requests served by several lanes finishing in random order always drains.
#include <cstdio>
#include <cmath>
#include <cstdlib>
#include <random>

#include "psx_arena.h"

using psxh::arena;

#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);    \
            std::exit(1);                                               \
        }                                                               \
    } while (0)

static void random_take_give(unsigned seed) {
    std::mt19937_64 rng(seed);
    const size_t total = 1 << 20;
    arena a(total);
    std::vector<unsigned char> owner(total / arena::ALIGN, 0);  // 1 = in use
    std::vector<std::pair<size_t, size_t>> live;
    for (int step = 0; step < 20000; ++step) {
        if (live.empty() || rng() % 3) {
            const size_t n = 1 + rng() % 40000;
            const size_t off = a.take(n);
            const size_t len = arena::round(n);
            if (off == arena::NONE) {
                // no free run of len bytes may exist in the model
                size_t run = 0, best = 0;
                for (unsigned char o : owner) { run = o ? 0 : run + 1; best = run > best ? run : best; }
                CHECK(best * arena::ALIGN < len);
                continue;
            }
            CHECK(off % arena::ALIGN == 0 && off + len <= total);
            for (size_t u = off / arena::ALIGN; u < (off + len) / arena::ALIGN; ++u) {
                CHECK(!owner[u]);
                owner[u] = 1;
            }
            live.push_back({off, len});
        } else {
            const size_t k = rng() % live.size();
            CHECK(a.give(live[k].first));
            CHECK(!a.give(live[k].first));  // a block is returned once
            for (size_t u = live[k].first / arena::ALIGN; u < (live[k].first + live[k].second) / arena::ALIGN; ++u)
                owner[u] = 0;
            live[k] = live.back();
            live.pop_back();
        }
        size_t used = 0;
        for (auto& l : live) used += l.second;
        CHECK(a.in_use() == used && a.peak() >= used && a.peak() <= total);
    }
    for (auto& l : live) CHECK(a.give(l.first));
    CHECK(a.in_use() == 0 && a.free_ranges() == 1 && a.largest_free() == total);
}

struct req {
    int id;
    size_t bytes;
    size_t off;
};

static void serving_drains(unsigned seed) {
    std::mt19937_64 rng(seed);
    const size_t budget = 64u << 20;
    arena a(budget);
    std::deque<req*> pending;
    std::vector<req> reqs(600);
    for (int i = 0; i < (int)reqs.size(); ++i) {
        // log-uniform 16 KiB .. 48 MiB outputs, all within the budget
        const double x = std::uniform_real_distribution<double>(std::log(16384.0), std::log(48.0 * (1 << 20)))(rng);
        reqs[i] = req{i, (size_t)std::exp(x), arena::NONE};
    }
    size_t next_arrival = 0;
    int admitted_upto = -1;
    std::vector<std::vector<req*>> lanes(3);
    size_t done = 0;
    for (int tick = 0; done < reqs.size(); ++tick) {
        CHECK(tick < 1000000);
        for (int k = 0; k < 3 && next_arrival < reqs.size(); ++k) pending.push_back(&reqs[next_arrival++]);
        for (auto& L : lanes) {
            if (!L.empty()) continue;
            psxh::admit_fifo(pending, a, 8, L, [](req* q) { return q->bytes; }, [](req* q, size_t o) { q->off = o; });
            for (req* q : L) {
                CHECK(q->id == admitted_upto + 1);  // arrival order
                admitted_upto = q->id;
            }
            CHECK(a.in_use() <= budget);
        }
        // a random lane finishes and returns its memory
        auto& L = lanes[rng() % lanes.size()];
        for (req* q : L) {
            CHECK(a.give(q->off));
            ++done;
        }
        L.clear();
    }
    CHECK(pending.empty() && a.in_use() == 0 && a.peak() <= budget);
}

int main() {
    for (unsigned s = 1; s <= 5; ++s) random_take_give(s);
    for (unsigned s = 1; s <= 5; ++s) serving_drains(s);
    std::printf("arena ok\n");
    return 0;
}
