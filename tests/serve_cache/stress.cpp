// CPU stress test of the request server's memory bookkeeping
// (magma_amd/csrc/serve_cache.h): random admissions into up to four in-flight
// batches, completions in random order with random success, factor keys that
// repeat (hits, pins, keeps, evictions) and budgets small enough to force
// eviction and re-carving.  Built with -fsanitize=address,undefined by
// tests/test_serve_cache.py.
//
// After every operation:
//   * every live block (in-flight grants and residents) lies inside the arena
//     and no two overlap; each holds a tag written at admission, so a block
//     handed out twice is also caught by its tag;
//   * the arena's used bytes equal the sum of the live blocks' sizes;
//   * each resident's pin count equals the in-flight hits on its key.
// Exit status 0 on success; the first violated check prints and aborts.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <random>
#include <vector>

#include "serve_cache.h"

using magma_amd::serve::Arena;
using magma_amd::serve::FactorCache;
using magma_amd::serve::Grant;

namespace {

#define CHECK(c)                                                              \
    do {                                                                      \
        if (!(c)) {                                                           \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            std::abort();                                                     \
        }                                                                     \
    } while (0)

struct Req {
    long long key;
    int n, kind;
    size_t abytes, fbytes;
    Grant g;
    unsigned long long tag;
};

struct Batch {
    std::vector<Req> reqs;
};

size_t abytes_of(int n) { return sizeof(double) * (size_t)n * n; }
size_t fbytes_of(int n) { return sizeof(float) * ((size_t)n * n + 2 * (size_t)n + 1024 * (size_t)((n + 31) / 32) + 4); }

void put_tag(void *p, unsigned long long t) { std::memcpy(p, &t, sizeof t); }
unsigned long long get_tag(const void *p)
{
    unsigned long long t;
    std::memcpy(&t, p, sizeof t);
    return t;
}

int run(unsigned seed, size_t budget, int steps, int max_batch)
{
    std::mt19937_64 rng(seed);
    std::vector<char> mem(budget);
    Arena arena;
    arena.init(mem.data(), budget);
    FactorCache cache(arena);
    std::vector<Batch> inflight;
    std::vector<Req> pending;
    std::map<const void *, unsigned long long> resident_tag;   // tag of each resident's F block
    unsigned long long next_tag = 1;
    long long admitted = 0, completed = 0;

    auto check = [&]() {
        struct Blk { const char *p; size_t b; };
        std::vector<Blk> live;
        std::map<long long, int> hits;
        for (const Batch &bt : inflight) {
            for (const Req &r : bt.reqs) {
                live.push_back({static_cast<const char *>(r.g.A), Arena::block(r.abytes)});
                CHECK(get_tag(r.g.A) == r.tag);
                if (r.g.hit) {
                    hits[r.key]++;
                } else {
                    live.push_back({static_cast<const char *>(r.g.F), Arena::block(r.fbytes)});
                    CHECK(get_tag(r.g.F) == r.tag);
                }
            }
        }
        cache.visit([&](long long key, const void *A, size_t ab, const void *F, size_t fb, int pins) {
            live.push_back({static_cast<const char *>(A), Arena::block(ab)});
            live.push_back({static_cast<const char *>(F), Arena::block(fb)});
            CHECK(pins == hits[key]);
            CHECK(resident_tag.count(F) && get_tag(F) == resident_tag[F]);
        });
        for (auto &kv : hits) {
            bool found = false;
            cache.visit([&](long long key, const void *, size_t, const void *, size_t, int) { found |= key == kv.first; });
            CHECK(found);
        }
        size_t sum = 0;
        for (const Blk &b : live) {
            CHECK(b.p >= mem.data() && b.p + b.b <= mem.data() + budget);
            sum += b.b;
        }
        CHECK(sum == arena.used());
        std::sort(live.begin(), live.end(), [](const Blk &x, const Blk &y) { return x.p < y.p; });
        for (size_t i = 1; i < live.size(); ++i) CHECK(live[i - 1].p + live[i - 1].b <= live[i].p);
    };

    for (int step = 0; step < steps; ++step) {
        // new arrivals: orders 16..512 (log-uniform), a third keyed from a small pool
        const int arrivals = (int)(rng() % 4);
        for (int a = 0; a < arrivals; ++a) {
            Req r;
            const double u = (rng() >> 11) * 0x1p-53;
            r.n = (int)std::exp(std::log(16.0) + u * (std::log(513.0) - std::log(16.0)));
            r.kind = (int)(rng() % 2);
            r.key = (rng() % 3 == 0) ? (long long)(rng() % 12) : -1;
            if (r.key >= 0) {
                // a key always names one system (order and kind), as callers use it
                r.n = 16 + (int)(r.key * 41 % 497);
                r.kind = (int)(r.key % 2);
            }
            r.abytes = abytes_of(r.n);
            r.fbytes = fbytes_of(r.n);
            if (Arena::block(r.abytes) + Arena::block(r.fbytes) > budget) continue;   // rejected at submission
            pending.push_back(r);
        }
        const bool can_launch = inflight.size() < 4 && !pending.empty();
        if (can_launch && (inflight.empty() || rng() % 2)) {
            Batch bt;
            while (!pending.empty() && (int)bt.reqs.size() < max_batch) {
                Req r = pending.front();
                if (!cache.admit(r.key, r.n, r.kind, r.abytes, r.fbytes, r.g)) break;
                r.tag = next_tag++;
                put_tag(r.g.A, r.tag);
                if (!r.g.hit) put_tag(r.g.F, r.tag);
                bt.reqs.push_back(r);
                pending.erase(pending.begin());
                ++admitted;
            }
            // nothing in flight and nothing admitted would be a stall
            CHECK(!bt.reqs.empty() || !inflight.empty());
            if (!bt.reqs.empty()) inflight.push_back(bt);
        } else if (!inflight.empty()) {
            const size_t k = rng() % inflight.size();
            Batch bt = inflight[k];
            inflight.erase(inflight.begin() + (long)k);
            for (const Req &r : bt.reqs) {
                const bool ok = rng() % 8 != 0;
                const bool kept = !r.g.hit && r.key >= 0 && ok;
                size_t before = cache.resident();
                cache.retire(r.key, r.n, r.kind, r.abytes, r.fbytes, r.g, ok);
                if (kept && cache.resident() == before + 1) resident_tag[r.g.F] = r.tag;
                ++completed;
            }
        }
        check();
    }
    // drain
    while (!inflight.empty() || !pending.empty()) {
        if (!inflight.empty()) {
            for (const Req &r : inflight.back().reqs) {
                size_t before = cache.resident();
                cache.retire(r.key, r.n, r.kind, r.abytes, r.fbytes, r.g, true);
                if (!r.g.hit && r.key >= 0 && cache.resident() == before + 1) resident_tag[r.g.F] = r.tag;
                ++completed;
            }
            inflight.pop_back();
        } else {
            Batch bt;
            while (!pending.empty() && (int)bt.reqs.size() < max_batch) {
                Req r = pending.front();
                if (!cache.admit(r.key, r.n, r.kind, r.abytes, r.fbytes, r.g)) break;
                r.tag = next_tag++;
                put_tag(r.g.A, r.tag);
                if (!r.g.hit) put_tag(r.g.F, r.tag);
                bt.reqs.push_back(r);
                pending.erase(pending.begin());
                ++admitted;
            }
            CHECK(!bt.reqs.empty());
            inflight.push_back(bt);
        }
        check();
    }
    CHECK(admitted == completed);
    CHECK(arena.used() == cache.resident_bytes());
    std::printf("seed %u budget %zu MiB: %lld requests, %lld evictions, %zu resident\n", seed, budget >> 20,
                completed, cache.evictions(), cache.resident());
    return 0;
}

}  // namespace

// Regression: a repeat request whose own resident factorization fills the
// budget so that its matrix copy does not fit beside it.  Before the fix the
// request was refused while the resident stayed pinned by it, and with
// nothing else in flight the server waited for ever.
void self_pinned_hit()
{
    const size_t budget = (size_t)4 << 20;
    std::vector<char> mem(budget);
    Arena arena;
    arena.init(mem.data(), budget);
    FactorCache cache(arena);
    const int n = 385;                        // A: 2 MiB block, F: 1 MiB block
    Grant g1;
    CHECK(cache.admit(9, n, 0, abytes_of(n), fbytes_of(n), g1) && !g1.hit);
    cache.retire(9, n, 0, abytes_of(n), fbytes_of(n), g1, true);
    CHECK(cache.resident() == 1 && arena.used() == ((size_t)3 << 20));
    Grant g2;
    CHECK(cache.admit(9, n, 0, abytes_of(n), fbytes_of(n), g2));    // served as a miss
    CHECK(!g2.hit && cache.resident() == 0 && cache.evictions() == 1);
    cache.retire(9, n, 0, abytes_of(n), fbytes_of(n), g2, true);
    CHECK(cache.resident() == 1 && arena.used() == ((size_t)3 << 20));
    // the same with room to spare is a hit
    std::vector<char> mem2((size_t)8 << 20);
    Arena a2;
    a2.init(mem2.data(), mem2.size());
    FactorCache c2(a2);
    CHECK(c2.admit(9, n, 0, abytes_of(n), fbytes_of(n), g1) && !g1.hit);
    c2.retire(9, n, 0, abytes_of(n), fbytes_of(n), g1, true);
    CHECK(c2.admit(9, n, 0, abytes_of(n), fbytes_of(n), g2) && g2.hit && g2.F == g1.F && g2.Asrc == g1.A);
    c2.retire(9, n, 0, abytes_of(n), fbytes_of(n), g2, true);
    std::printf("self-pinned hit: ok\n");
}

// Regression: blocks of two sizes that each fit the bump space but not both.
void two_sizes_fit_together()
{
    const size_t budget = (size_t)4 << 20;
    std::vector<char> mem(budget);
    Arena arena;
    arena.init(mem.data(), budget);
    FactorCache cache(arena);
    // a resident of 1 MiB + 1 MiB (n = 300) leaves 2 MiB of bump space
    Grant g0;
    CHECK(cache.admit(1, 300, 0, abytes_of(300), fbytes_of(300), g0));
    cache.retire(1, 300, 0, abytes_of(300), fbytes_of(300), g0, true);
    // n = 385 needs 2 MiB + 1 MiB: each fits alone, not both; the resident must go
    Grant g1;
    CHECK(cache.admit(-1, 385, 0, abytes_of(385), fbytes_of(385), g1));
    CHECK(cache.evictions() == 1 && cache.resident() == 0);
    cache.retire(-1, 385, 0, abytes_of(385), fbytes_of(385), g1, true);
    CHECK(arena.used() == 0);
    std::printf("two sizes: ok\n");
}

int main()
{
    self_pinned_hit();
    two_sizes_fit_together();
    // budgets from tight (a 512 x 512 system barely fits) to roomy
    const size_t budgets[] = {(size_t)4 << 20, (size_t)6 << 20, (size_t)16 << 20, (size_t)64 << 20};
    for (unsigned seed = 1; seed <= 6; ++seed)
        for (size_t b : budgets) run(seed, b, 4000, (int)(1 + seed % 8));
    std::printf("ok\n");
    return 0;
}
