// Admission bookkeeping of the solve server: the HBM budget, the resident
// factorizations kept under a key for later right-hand sides, and their eviction.
//
// Pure host code with no HIP calls (device memory is released through a callback), so
// the same logic is compiled into the library and into a CPU-only test.
//
// Rules:
//   * a request needs its right-hand-side buffers, plus a factorization unless a resident
//     one of the same kind and order is kept under its key (a hit);
//   * a hit is pinned (users > 0) before anything is evicted for it, so the eviction
//     that makes room for its right-hand side can never take its own factorization;
//   * eviction takes idle (users == 0) resident factorizations, least recently used
//     first, and only as many as the request needs;
//   * a pinned factorization is unpinned by the worker that used it, after that
//     worker's device work has finished.
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>
#include <unordered_map>

namespace bola {

struct ResidentFactor {
    int kind = 0;
    size_t n = 0;
    void* f = nullptr;      // the factorization (the library's factorization object)
    void* ipiv = nullptr;   // a second buffer released with it (may be null)
    size_t bytes = 0;       // what it counts against the budget
    int users = 0;          // requests being served from it right now
    uint64_t last_use = 0;  // for least-recently-used eviction
};

inline size_t resident_factor_bytes(size_t n) { return n * n * 8 + ((n * 4 + 255) & ~(size_t)255); }

class ServeCache {
  public:
    // release(f, ipiv): frees an evicted factorization's device memory
    ServeCache(size_t budget, std::function<void(void*, void*)> release) : budget_(budget), release_(release) {}

    size_t budget() const { return budget_; }
    size_t in_use() const { return in_use_; }
    size_t peak() const { return peak_; }
    uint64_t evictions() const { return evictions_; }
    void reset_peak() { peak_ = in_use_; }
    size_t cached_bytes() const {
        size_t b = 0;
        for (auto& kv : map_) b += kv.second.bytes;
        return b;
    }
    size_t size() const { return map_.size(); }

    // A resident factorization under key that can serve (kind, n), or null.
    ResidentFactor* lookup(uint64_t key, int kind, size_t n) {
        if (!key) return nullptr;
        auto it = map_.find(key);
        if (it == map_.end() || it->second.kind != kind || it->second.n != n) return nullptr;
        return &it->second;
    }

    // Bytes a request needs beyond what is resident for it: rhs (its right-hand-side
    // buffers) and, unless a factorization is resident for it, fbytes.
    size_t need(uint64_t key, int kind, size_t n, size_t rhs, size_t fbytes, bool& hit) {
        hit = lookup(key, kind, n) != nullptr;
        return hit ? rhs : fbytes + rhs;
    }

    // Reserves a request's bytes, evicting idle factorizations if needed.  On success
    // the bytes count against the budget until release(), and on a hit the resident
    // factorization is pinned (unpin() after use).  Returns false, with nothing
    // changed except evictions already made, when the request cannot fit now.
    bool reserve(uint64_t key, int kind, size_t n, size_t rhs, size_t fbytes, size_t& reserved, bool& hit) {
        const size_t nb = need(key, kind, n, rhs, fbytes, hit);
        ResidentFactor* own = hit ? lookup(key, kind, n) : nullptr;
        if (own) own->users++;  // pinned before any eviction
        if (in_use_ + nb > budget_ && !make_room(nb)) {
            if (own) own->users--;
            hit = false;
            return false;
        }
        in_use_ += nb;
        if (in_use_ > peak_) peak_ = in_use_;
        reserved = nb;
        return true;
    }

    // The pinned factorization of an admitted hit (null if the bookkeeping is broken).
    ResidentFactor* pinned(uint64_t key) {
        auto it = map_.find(key);
        if (it == map_.end() || it->second.users <= 0) return nullptr;
        it->second.last_use = ++clock_;
        return &it->second;
    }

    void unpin(uint64_t key) {
        auto it = map_.find(key);
        if (it != map_.end() && it->second.users > 0) it->second.users--;
    }

    // A request finished: its reserved bytes are returned, except a factorization it
    // leaves resident under its key.  Returns false (and keeps nothing) when the key
    // is busy with another resident factorization; the caller then frees its own.
    bool finish(size_t reserved, bool keep, uint64_t key, int kind, size_t n, size_t fbytes, void* f, void* ipiv) {
        size_t kept = 0;
        bool stored = false;
        if (keep && key) {
            auto it = map_.find(key);
            if (it != map_.end() && it->second.users == 0) {  // a stale entry under this key
                release_(it->second.f, it->second.ipiv);
                in_use_ -= it->second.bytes;
                map_.erase(it);
                it = map_.end();
            }
            if (it == map_.end()) {
                ResidentFactor r;
                r.kind = kind;
                r.n = n;
                r.f = f;
                r.ipiv = ipiv;
                r.bytes = fbytes;
                r.last_use = ++clock_;
                map_[key] = r;
                kept = r.bytes;
                stored = true;
            }
        }
        in_use_ -= reserved - kept;
        return stored;
    }

    // Frees every resident factorization (the server is closing).
    void clear() {
        for (auto& kv : map_) release_(kv.second.f, kv.second.ipiv);
        for (auto& kv : map_) in_use_ -= kv.second.bytes;
        map_.clear();
    }

  private:
    bool make_room(size_t bytes) {
        while (in_use_ + bytes > budget_) {
            auto victim = map_.end();
            for (auto it = map_.begin(); it != map_.end(); ++it)
                if (it->second.users == 0 &&
                    (victim == map_.end() || it->second.last_use < victim->second.last_use))
                    victim = it;
            if (victim == map_.end()) return false;
            release_(victim->second.f, victim->second.ipiv);
            in_use_ -= victim->second.bytes;
            evictions_++;
            map_.erase(victim);
        }
        return true;
    }

    size_t budget_;
    size_t in_use_ = 0, peak_ = 0;
    uint64_t clock_ = 0, evictions_ = 0;
    std::function<void(void*, void*)> release_;
    std::unordered_map<uint64_t, ResidentFactor> map_;
};

}  // namespace bola
