// serve_cache.h -- host-side memory bookkeeping of the request server
// (serve.hip): a power-of-two block arena over one device allocation and the
// cache of resident factorizations (least recently used first out, pinned
// while a request that reuses them is in flight).
//
// Nothing here calls HIP: the arena only hands out addresses inside a range
// the caller allocated.  The same code is built for the CPU by
// tests/serve_cache/stress.cpp and driven through random admission, eviction
// and re-carving sequences under AddressSanitizer and UBSan.
#pragma once

#include <cstddef>
#include <list>
#include <unordered_map>
#include <vector>

namespace magma_amd {
namespace serve {

constexpr int kMinClass = 12;          // 4 KiB
constexpr int kMaxClass = 24;          // 16 MiB

// Power-of-two blocks carved from [base, base + cap) with one free list per
// block size.  Freed blocks are reused for their own size only; once nothing
// at all is allocated the whole range can be re-carved (reset).
class Arena {
public:
    void init(char *base, size_t cap)
    {
        base_ = base;
        cap_ = cap;
        reset();
    }
    static int cls(size_t bytes)
    {
        int c = kMinClass;
        while (c < kMaxClass && ((size_t)1 << c) < bytes) ++c;
        return ((size_t)1 << c) >= bytes ? c : -1;
    }
    static size_t block(size_t bytes)
    {
        const int c = cls(bytes);
        return c < 0 ? 0 : (size_t)1 << c;
    }
    void *alloc(size_t bytes)
    {
        const int c = cls(bytes);
        if (c < 0) return nullptr;
        auto &fl = free_[c - kMinClass];
        if (!fl.empty()) {
            void *p = fl.back();
            fl.pop_back();
            used_ += (size_t)1 << c;
            return p;
        }
        const size_t b = (size_t)1 << c;
        if (top_ + b > cap_) return nullptr;
        void *p = base_ + top_;
        top_ += b;
        used_ += b;
        return p;
    }
    void free(void *p, size_t bytes)
    {
        if (!p) return;
        const int c = cls(bytes);
        free_[c - kMinClass].push_back(p);
        used_ -= (size_t)1 << c;
    }
    // would alloc(bytes) succeed?
    bool can(size_t bytes) const
    {
        const int c = cls(bytes);
        if (c < 0) return false;
        return !free_[c - kMinClass].empty() || top_ + ((size_t)1 << c) <= cap_;
    }
    size_t used() const { return used_; }
    size_t cap() const { return cap_; }
    char *base() const { return base_; }
    // forgets every block; only valid when nothing is allocated
    void reset()
    {
        top_ = 0;
        used_ = 0;
        for (auto &fl : free_) fl.clear();
    }

private:
    char *base_ = nullptr;
    size_t cap_ = 0, top_ = 0, used_ = 0;
    std::vector<void *> free_[kMaxClass - kMinClass + 1];
};

// One request's blocks: its matrix copy A and its factor storage F, which is
// a resident factorization's when hit is set.
struct Grant {
    void *A = nullptr;
    void *F = nullptr;
    const void *Asrc = nullptr;    // a hit's resident matrix copy (the request's source), else null
    bool hit = false;
};

// Resident factorizations by key, with the arena they live in.
class FactorCache {
public:
    explicit FactorCache(Arena &arena) : arena_(arena) {}

    // Blocks of abytes (A) and fbytes (F) for a request of order n and kind,
    // with factor key `key` (< 0: none).  A resident factorization under key
    // with the same order and kind is reused (hit: only A is allocated, the
    // resident is pinned until retire).  Unpinned residents are evicted least
    // recently used first to make room.  Returns false, holding nothing, when
    // the request does not fit even then.
    bool admit(long long key, int n, int kind, size_t abytes, size_t fbytes, Grant &g)
    {
        g = Grant();
        auto it = key >= 0 ? res_.find(key) : res_.end();
        if (it != res_.end() && it->second.n == n && it->second.kind == kind) {
            // pinned before making room, so that the eviction cannot pick it
            it->second.pins++;
            void *A = make_room(abytes, 0) ? arena_.alloc(abytes) : nullptr;
            if (A) {
                g.A = A;
                g.F = it->second.F;
                g.Asrc = it->second.A;
                g.hit = true;
                lru_.splice(lru_.end(), lru_, it->second.lru);
                return true;
            }
            it->second.pins--;
            // no room for the matrix copy while the resident itself is held
            // (with nothing else in flight this would never change): serve
            // the request as a miss, which may evict that resident too; its
            // own factorization replaces the resident when it completes
        }
        if (!make_room(abytes, fbytes)) return false;
        void *A = arena_.alloc(abytes);
        void *F = A ? arena_.alloc(fbytes) : nullptr;
        if (!A || !F) {
            arena_.free(A, abytes);
            return false;
        }
        g.A = A;
        g.F = F;
        return true;
    }

    // The request's batch has completed.  A hit unpins its resident and frees
    // its matrix copy.  A miss with a key whose solve succeeded keeps both
    // blocks as the resident factorization under that key (unless another
    // request made one first); otherwise both blocks are freed.
    void retire(long long key, int n, int kind, size_t abytes, size_t fbytes, const Grant &g, bool success)
    {
        if (g.hit) {
            auto it = res_.find(key);
            if (it != res_.end()) it->second.pins--;
            arena_.free(g.A, abytes);
            return;
        }
        if (key >= 0 && success && !res_.count(key)) {
            Resident e;
            e.n = n;
            e.kind = kind;
            e.A = g.A;
            e.F = g.F;
            e.abytes = abytes;
            e.fbytes = fbytes;
            e.lru = lru_.insert(lru_.end(), key);
            res_.emplace(key, e);
            return;
        }
        arena_.free(g.A, abytes);
        arena_.free(g.F, fbytes);
    }

    size_t resident() const { return res_.size(); }
    long long evictions() const { return evictions_; }
    // bytes held by residents (in block sizes)
    size_t resident_bytes() const
    {
        size_t b = 0;
        for (const auto &kv : res_) b += Arena::block(kv.second.abytes) + Arena::block(kv.second.fbytes);
        return b;
    }
    // fn(key, A, abytes, F, fbytes, pins) for every resident (consistency checks)
    template <class Fn>
    void visit(Fn fn) const
    {
        for (const auto &kv : res_)
            fn(kv.first, kv.second.A, kv.second.abytes, kv.second.F, kv.second.fbytes, kv.second.pins);
    }

private:
    struct Resident {
        int n = 0, kind = 0;
        void *A = nullptr;
        void *F = nullptr;
        size_t abytes = 0, fbytes = 0;
        int pins = 0;
        std::list<long long>::iterator lru;
    };

    // would blocks of a and f bytes (f = 0: none) fit now?  Both blocks may
    // come from the bump space, so the second is tried with the first taken.
    bool fits(size_t a, size_t f)
    {
        if (f == 0) return arena_.can(a);
        void *p = arena_.alloc(a);
        if (!p) return false;
        const bool ok = arena_.can(f);
        arena_.free(p, a);
        return ok;
    }

    bool make_room(size_t a, size_t f)
    {
        while (!fits(a, f)) {
            auto it = lru_.begin();
            while (it != lru_.end() && res_.at(*it).pins > 0) ++it;
            if (it == lru_.end()) {
                // blocks of other sizes may hold the free space: once nothing
                // at all is allocated the arena is re-carved from scratch
                if (arena_.used() == 0) {
                    arena_.reset();
                    return fits(a, f);
                }
                return false;
            }
            const long long key = *it;
            auto r = res_.find(key);
            arena_.free(r->second.A, r->second.abytes);
            arena_.free(r->second.F, r->second.fbytes);
            lru_.erase(it);
            res_.erase(r);
            ++evictions_;
        }
        return true;
    }

    Arena &arena_;
    std::unordered_map<long long, Resident> res_;
    std::list<long long> lru_;        // front = least recently used
    long long evictions_ = 0;
};

}  // namespace serve
}  // namespace magma_amd
