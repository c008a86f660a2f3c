// A solve server (no reference counterpart: BOLA is a library of one-shot calls).
//
// Requests (LU or Cholesky solves of mixed size, FP64, bit-identical to the one-shot
// entry points) are submitted from any host thread and return a ticket.  A fixed number
// of slots, each a worker thread with its own handle and HIP stream, take requests as
// soon as they are free: a request is admitted when a slot is idle and its device
// buffers fit in the HBM budget (factorizations kept for later right-hand sides count
// against the budget too, and the least recently used idle ones are evicted to make
// room).  Admission scans the queue in arrival order and may pass over a request that
// does not fit yet, but only a bounded number of times, so large requests are not
// starved.  The slots' work overlaps on the device; each request's result is the same
// as when it runs alone, since every slot runs the same deterministic kernels.
#include "common.hpp"
#include "serve_cache.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

namespace bola {
namespace {

double now_us() {
    using namespace std::chrono;
    return (double)duration_cast<nanoseconds>(steady_clock::now().time_since_epoch()).count() * 1e-3;
}

struct Ticket {
    uint64_t id = 0;
    bola_hip_request req{};
    bola_hip_result res{};
    double t_submit = 0.0;
    bool done = false;
    int passed = 0;  // times admission passed over it while it was first in line
};

// The stream of the worker thread that releases device memory (null on other threads).
thread_local hipStream_t tl_stream = nullptr;

void release_device(void* f, void*) { factor_release((bola_hip_factor_s*)f, tl_stream); }

// the factorization object a request needs, its identity in the cache and its size
struct FactorKind {
    int kind, precision, cache_kind;
    size_t bytes;
};
FactorKind factor_kind(const bola_hip_request& r) {
    FactorKind k;
    if (r.kind == BOLA_HIP_SERVE_CHOL) {
        k.kind = BOLA_HIP_FACTOR_CHOL_L;
        k.precision = 0;
    } else if (r.precision == BOLA_HIP_SERVE_FP64) {
        k.kind = BOLA_HIP_FACTOR_LU;
        k.precision = 0;
    } else {
        k.kind = BOLA_HIP_FACTOR_MIXED;
        k.precision = r.precision - 1;
    }
    k.cache_kind = k.kind * 4096 + k.precision;
    k.bytes = bola_hip_factor_bytes(k.kind, k.precision, r.n);
    // the mixed factorization's FP64 fallback is made on first need: count it from the start
    if (k.kind == BOLA_HIP_FACTOR_MIXED) k.bytes += r.n * r.n * 8 + r.n * 4 + 512;
    return k;
}
// right-hand-side buffers on the device: b and x (one buffer when the solve is in place)
size_t rhs_bytes(const bola_hip_request& r, const FactorKind& k) {
    return (k.kind == BOLA_HIP_FACTOR_MIXED ? 2 : 1) * r.n * r.nrhs * 8;
}

}  // namespace
}  // namespace bola

using namespace bola;

struct bola_hip_server_s {
    int device = 0;
    std::mutex mu;
    // One LU at a time above kBigLu: its panel kernel's workgroups (up to 128 of 35 KB of
    // LDS each, or 512 of 70 KB above 8192 rows) must all be resident at once, which
    // several such panels on the slots' streams, beside their GEMMs, cannot be sure of.
    // One large factorization fills the GPU anyway.
    std::mutex big_lu;
    static constexpr size_t kBigLu = 2048;
    std::condition_variable cv_work, cv_done;
    std::deque<Ticket*> queue;
    std::unordered_map<uint64_t, Ticket*> tickets;
    ServeCache cache;
    std::vector<std::thread> workers;
    bool stop = false;
    uint64_t next_id = 1;
    int running = 0;
    // statistics
    uint64_t submitted = 0, completed = 0, rejected = 0, hits = 0, evict_base = 0;
    std::vector<double> lat;
    double t_first = -1.0, t_last = 0.0;

    explicit bola_hip_server_s(size_t budget) : cache(budget, release_device) {}

    // the next admissible request (the lock is held)
    Ticket* admit(size_t& reserved, bool& hit) {
        for (size_t qi = 0; qi < queue.size(); qi++) {
            Ticket* t = queue[qi];
            const bola_hip_request& r = t->req;
            const FactorKind k = factor_kind(r);
            if (cache.reserve(r.factor_key, k.cache_kind, r.n, rhs_bytes(r, k), k.bytes, reserved, hit)) {
                queue.erase(queue.begin() + (long)qi);
                return t;
            }
            // the first in line does not fit yet: others may pass it, a bounded number of times
            if (qi == 0 && ++t->passed > 4 * (int)workers.size()) return nullptr;
        }
        return nullptr;
    }
    void work(int slot);
};

void bola_hip_server_s::work(int slot) {
    (void)hipSetDevice(device);
    bola_hip_handle_t h = nullptr;
    if (bola_hip_create(&h) != 0) return;
    // the slots' work shares the device: no solve whose workgroups wait for one another
    // beyond the factorization's own panel, no CU-partitioned LU streams
    bola_hip_set_exclusive(h, 0);
    hipStream_t st = h->stream;
    tl_stream = st;  // evictions made by this thread are ordered on its stream
    for (;;) {
        Ticket* t = nullptr;
        size_t reserved = 0;
        bool hit = false;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] {
                if (stop) return true;
                t = admit(reserved, hit);
                return t != nullptr;
            });
            if (!t) break;  // stopping
            running++;
        }
        bola_hip_request& r = t->req;
        bola_hip_result& res = t->res;
        res.slot = slot;
        res.cache_hit = hit ? 1 : 0;
        res.queue_us = now_us() - t->t_submit;
        const FactorKind k = factor_kind(r);
        bola_hip_factor_s* fo = nullptr;
        double* xb = nullptr;
        double* bb = nullptr;
        bool keep = false;
        int status = 0, info = 0;
        const size_t n = r.n;
        try {
            if (hit) {
                std::lock_guard<std::mutex> lk(mu);
                ResidentFactor* f = cache.pinned(r.factor_key);
                if (!f) throw HipFailure{hipErrorInvalidValue, "resident factorization missing", __LINE__};
                fo = (bola_hip_factor_s*)f->f;
            } else {
                std::unique_lock<std::mutex> big(big_lu, std::defer_lock);
                if (r.kind == BOLA_HIP_SERVE_LU && n > kBigLu) big.lock();  // released once it is factored
                fo = factor_build(h, k.kind, k.precision, n, r.a, n, &info, true);
            }
            if (info == 0) {
                const size_t rb = std::max<size_t>(n * r.nrhs, 1) * 8;
                BOLA_CHECK(hipMallocAsync((void**)&xb, rb, st));
                BOLA_CHECK(hipMemcpyAsync(xb, r.b, n * r.nrhs * 8, hipMemcpyDefault, st));
                const double* bsrc = xb;
                if (k.kind == BOLA_HIP_FACTOR_MIXED) {  // the refinement reads b while it writes x
                    BOLA_CHECK(hipMallocAsync((void**)&bb, rb, st));
                    BOLA_CHECK(hipMemcpyAsync(bb, xb, n * r.nrhs * 8, hipMemcpyDeviceToDevice, st));
                    bsrc = bb;
                }
                bola_hip_mixed_stats ms{};
                int rc = bola_hip_factor_solve_dev(h, fo, r.nrhs, bsrc, n, xb, n, &ms);
                if (rc < 0) throw HipFailure{hipErrorUnknown, "solve", __LINE__};
                if (rc > 0) info = rc;  // the mixed solver's FP64 fallback met a zero pivot
                if (k.kind == BOLA_HIP_FACTOR_MIXED && rc == 0 && !ms.converged) status = BOLA_HIP_ERR_DEVICE;
                if (rc == 0) BOLA_CHECK(hipMemcpyAsync(r.x, xb, n * r.nrhs * 8, hipMemcpyDefault, st));
                BOLA_CHECK(hipFreeAsync(xb, st));
                xb = nullptr;
                if (bb) BOLA_CHECK(hipFreeAsync(bb, st));
                bb = nullptr;
                BOLA_CHECK(hipStreamSynchronize(st));
                keep = !hit && r.factor_key != 0 && rc == 0;
            }
        } catch (const HipFailure& f) {
            status = BOLA_HIP_ERR_DEVICE;
            (void)hipStreamSynchronize(st);
            if (xb) (void)hipFreeAsync(xb, st);
            if (bb) (void)hipFreeAsync(bb, st);
            keep = false;
            (void)f;
        }
        if (hit) fo = nullptr;  // resident: owned by the cache
        (void)hipStreamSynchronize(st);
        const double t_done = now_us();
        {
            std::lock_guard<std::mutex> lk(mu);
            const bool stored = cache.finish(reserved, keep, r.factor_key, k.cache_kind, n, k.bytes, fo, nullptr);
            if (!stored && fo) factor_release(fo, st);  // not kept (or the key is busy with another)
            if (hit) {
                cache.unpin(r.factor_key);
                hits++;
            }
            res.status = status;
            res.info = info;
            res.latency_us = t_done - t->t_submit;
            t->done = true;
            completed++;
            lat.push_back(res.latency_us);
            t_last = std::max(t_last, t_done);
            running--;
        }
        cv_done.notify_all();
        cv_work.notify_all();
    }
    (void)hipStreamSynchronize(st);
    bola_hip_destroy(h);
}

extern "C" {

int bola_hip_server_create(size_t hbm_budget, int slots, bola_hip_server_t* out) {
    if (!out) return -3;
    *out = nullptr;
    if (slots < 1 || slots > 16) {
        set_error(nullptr, "Argument 2 to server_create had an illegal value (1 <= slots <= 16)");
        return -2;
    }
    if (hbm_budget < ((size_t)1 << 20)) {
        set_error(nullptr, "Argument 1 to server_create had an illegal value (at least 1 MiB)");
        return -1;
    }
    auto* s = new bola_hip_server_s(hbm_budget);
    hipGetDevice(&s->device);
    for (int i = 0; i < slots; i++) s->workers.emplace_back([s, i] { s->work(i); });
    *out = s;
    return 0;
}

int bola_hip_server_submit(bola_hip_server_t s, const bola_hip_request* r, uint64_t* ticket) {
    if (!s || !r || !ticket) return -1;
    if ((r->kind != BOLA_HIP_SERVE_LU && r->kind != BOLA_HIP_SERVE_CHOL) || r->n == 0 || r->nrhs == 0 || !r->b ||
        !r->x || (!r->a && !r->factor_key)) {
        set_error(nullptr, "server_submit: invalid request (kind, n > 0, nrhs > 0, a, b and x)");
        return -2;
    }
    auto* t = new Ticket;
    t->req = *r;
    t->t_submit = now_us();
    std::lock_guard<std::mutex> lk(s->mu);
    bool hit = false;
    if (r->precision != BOLA_HIP_SERVE_FP64 &&
        (r->kind != BOLA_HIP_SERVE_LU || r->precision < 1 || ((r->precision - 1) & 0xFF) > BOLA_HIP_PREC_FP16X3 ||
         ((r->precision - 1) & ~(0xFF | BOLA_HIP_MIX_RBT)) != 0 || r->n > 16384)) {
        set_error(nullptr, "server_submit: invalid precision (mixed precision for LU requests, n <= 16384)");
        return -2;
    }
    const FactorKind k = factor_kind(*r);
    const size_t nb = s->cache.need(r->factor_key, k.cache_kind, r->n, rhs_bytes(*r, k), k.bytes, hit);
    if (nb > s->cache.budget() || (!hit && !r->a)) {
        delete t;
        s->rejected++;
        set_error(nullptr, !r->a ? "server_submit: no matrix and no cached factorization under this key"
                                 : "server_submit: the request alone exceeds the HBM budget");
        return -3;
    }
    t->id = s->next_id++;
    *ticket = t->id;
    s->tickets[t->id] = t;
    s->queue.push_back(t);
    s->submitted++;
    if (s->t_first < 0.0) s->t_first = t->t_submit;
    s->cv_work.notify_all();
    return 0;
}

static int finish(bola_hip_server_t s, Ticket* t, bola_hip_result* res) {
    if (res) *res = t->res;
    s->tickets.erase(t->id);
    delete t;
    return 0;
}

int bola_hip_server_wait(bola_hip_server_t s, uint64_t ticket, bola_hip_result* res) {
    if (!s) return -1;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return -2;
    Ticket* t = it->second;
    s->cv_done.wait(lk, [&] { return t->done; });
    return finish(s, t, res);
}

int bola_hip_server_poll(bola_hip_server_t s, uint64_t ticket, bola_hip_result* res) {
    if (!s) return -1;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return -2;
    if (!it->second->done) return 1;
    return finish(s, it->second, res);
}

int bola_hip_server_drain(bola_hip_server_t s) {
    if (!s) return -1;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->queue.empty() && s->running == 0; });
    return 0;
}

int bola_hip_server_get_stats(bola_hip_server_t s, bola_hip_server_stats* st) {
    if (!s || !st) return -1;
    std::lock_guard<std::mutex> lk(s->mu);
    *st = bola_hip_server_stats{};
    st->submitted = s->submitted;
    st->completed = s->completed;
    st->rejected = s->rejected;
    st->cache_hits = s->hits;
    st->evictions = s->cache.evictions() - s->evict_base;
    st->hbm_in_use = s->cache.in_use();
    st->hbm_peak = s->cache.peak();
    st->hbm_budget = s->cache.budget();
    st->hbm_cached = s->cache.cached_bytes();
    if (!s->lat.empty()) {
        std::vector<double> v = s->lat;
        std::sort(v.begin(), v.end());
        auto pct = [&](double p) {
            const size_t i = (size_t)std::min<double>((double)v.size() - 1, std::ceil(p * (double)v.size()) - 1);
            return v[i];
        };
        st->p50_us = pct(0.50);
        st->p99_us = pct(0.99);
        double sum = 0.0;
        for (double x : v) sum += x;
        st->mean_us = sum / (double)v.size();
        const double span = s->t_last - s->t_first;
        st->solves_per_s = span > 0.0 ? (double)s->completed / (span * 1e-6) : 0.0;
    }
    return 0;
}

int bola_hip_server_reset_stats(bola_hip_server_t s) {
    if (!s) return -1;
    std::lock_guard<std::mutex> lk(s->mu);
    s->submitted = s->completed = s->rejected = s->hits = 0;
    s->evict_base = s->cache.evictions();
    s->lat.clear();
    s->t_first = -1.0;
    s->t_last = 0.0;
    s->cache.reset_peak();
    return 0;
}

int bola_hip_server_destroy(bola_hip_server_t s) {
    if (!s) return 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        // requests not yet admitted are cancelled
        for (Ticket* t : s->queue) {
            t->res.status = BOLA_HIP_ERR_CANCELLED;
            t->done = true;
        }
        s->queue.clear();
        s->stop = true;
    }
    s->cv_work.notify_all();
    s->cv_done.notify_all();
    for (auto& w : s->workers) w.join();
    hipSetDevice(s->device);
    s->cache.clear();  // the workers have joined: no factorization is in use
    for (auto& kv : s->tickets) delete kv.second;
    delete s;
    return 0;
}

}  // extern "C"
