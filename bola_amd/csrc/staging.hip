// Host <-> device staging for the host-pointer entry points (the reference's functions
// take host slices: getrf.rs:37, gemm.rs:88-102).
//
// Operands travel through a pinned host buffer (hipHostMalloc, grown on demand and kept
// by the handle) on a copy stream of the handle's own:
//   * upload: host threads copy a chunk of columns from the caller's matrix into the
//     pinned buffer, the chunk's DMA is queued at once and the next chunk is copied while
//     it runs; a stream that needs columns [0, c) waits for the chunk that completes them,
//     so a factorization starts on its first columns while the rest is still arriving;
//   * download: row strips (or the whole matrix) are copied device -> pinned on the copy
//     stream as soon as the work that finalises them is queued, and host threads copy each
//     strip into the caller's matrix as soon as its DMA has completed.
// Small operands (below 1 MiB) skip the pinned path.
#include "common.hpp"

#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>

namespace bola {

namespace {

constexpr size_t kDirectBytes = (size_t)1 << 20;  // below this, a plain asynchronous copy
constexpr size_t kChunkBytes = (size_t)16 << 20;  // upload chunk

}  // namespace

// A fixed set of host threads that run one parallel loop at a time.
struct CopyPool {
    std::vector<std::thread> th;
    std::mutex mu;
    std::condition_variable cv_go, cv_done;
    std::function<void(size_t, size_t)> job;
    size_t items = 0;
    uint64_t gen = 0;
    int pending = 0;
    bool quit = false;

    explicit CopyPool(int n) {
        for (int t = 0; t < n; t++)
            th.emplace_back([this, t, n] {
                uint64_t seen = 0;
                for (;;) {
                    std::function<void(size_t, size_t)> f;
                    size_t cnt;
                    {
                        std::unique_lock<std::mutex> lk(mu);
                        cv_go.wait(lk, [&] { return quit || gen != seen; });
                        if (quit) return;
                        seen = gen;
                        f = job;
                        cnt = items;
                    }
                    const size_t lo = cnt * t / n, hi = cnt * (t + 1) / n;
                    if (lo < hi) f(lo, hi);
                    std::lock_guard<std::mutex> lk(mu);
                    if (--pending == 0) cv_done.notify_all();
                }
            });
    }
    ~CopyPool() {
        {
            std::lock_guard<std::mutex> lk(mu);
            quit = true;
        }
        cv_go.notify_all();
        for (auto& t : th) t.join();
    }
    // f(lo, hi) over [0, n) split across the threads; returns when all are done
    void run(size_t n, std::function<void(size_t, size_t)> f) {
        if (n == 0) return;
        std::unique_lock<std::mutex> lk(mu);
        job = std::move(f);
        items = n;
        pending = (int)th.size();
        gen++;
        cv_go.notify_all();
        cv_done.wait(lk, [&] { return pending == 0; });
    }
};

struct PinnedStage {
    char* buf = nullptr;
    size_t bytes = 0;
    hipStream_t copy = nullptr;
    std::vector<hipEvent_t> evs;
    size_t nev = 0;
    hipEvent_t last = nullptr;  // the newest DMA that touches buf
    CopyPool* pool = nullptr;

    hipEvent_t event() {
        if (nev == evs.size()) {
            hipEvent_t e;
            BOLA_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            evs.push_back(e);
        }
        return evs[nev++];
    }
    ~PinnedStage() {
        if (copy) (void)hipStreamSynchronize(copy);
        for (auto e : evs) (void)hipEventDestroy(e);
        if (copy) (void)hipStreamDestroy(copy);
        if (buf) (void)hipHostFree(buf);
        delete pool;
    }
};

PinnedStage* pinned_stage(bola_hip_handle_s* h, size_t need) {
    if (!h->pin) h->pin = new PinnedStage;
    PinnedStage* p = h->pin;
    if (!p->copy) BOLA_CHECK(hipStreamCreateWithFlags(&p->copy, hipStreamNonBlocking));
    if (!p->pool) {
        const unsigned hw = std::thread::hardware_concurrency();
        p->pool = new CopyPool((int)std::max(1u, std::min(16u, hw ? hw : 4u)));
    }
    // the previous transfer's DMAs are done with the buffer before it is reused
    if (p->last) BOLA_CHECK(hipEventSynchronize(p->last));
    p->last = nullptr;
    p->nev = 0;
    if (need > p->bytes) {
        if (p->buf) BOLA_CHECK(hipHostFree(p->buf));
        p->buf = nullptr;
        p->bytes = 0;
        const size_t nb = need + (need >> 3);
        BOLA_CHECK(hipHostMalloc((void**)&p->buf, nb, hipHostMallocDefault));
        p->bytes = nb;
    }
    return p;
}

void release_pinned_stage(bola_hip_handle_s* h) {
    delete h->pin;
    h->pin = nullptr;
}

// ---- upload

HostUpload::HostUpload(bola_hip_handle_s* h_, double* dst_, size_t dld_, const double* src_, size_t ld_,
                       size_t rows_, size_t cols_)
    : h(h_), dst(dst_), dld(dld_), src(src_), ld(ld_), rows(rows_), cols(cols_) {
    if (rows == 0 || cols == 0) {
        done = cols;
        return;
    }
    direct = rows * cols * 8 < kDirectBytes;
    if (direct) {
        BOLA_CHECK(hipMemcpy2DAsync(dst, dld * 8, src, ld * 8, rows * 8, cols, hipMemcpyHostToDevice, h->stream));
        done = cols;
        return;
    }
    p = pinned_stage(h, rows * cols * 8);
    chunk = std::max<size_t>(1, kChunkBytes / (rows * 8));
    // the device buffer may still be read by earlier work on the handle's stream
    hipEvent_t e = p->event();
    BOLA_CHECK(hipEventRecord(e, h->stream));
    BOLA_CHECK(hipStreamWaitEvent(p->copy, e, 0));
}

void HostUpload::need(hipStream_t s, size_t c_end) {
    c_end = std::min(c_end, cols);
    if (cols == 0) return;
    if (direct) {  // copied on the handle's stream
        if (s && s != h->stream) {
            hipEvent_t e;
            BOLA_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            BOLA_CHECK(hipEventRecord(e, h->stream));
            BOLA_CHECK(hipStreamWaitEvent(s, e, 0));
            BOLA_CHECK(hipEventDestroy(e));
        }
        return;
    }
    while (done < c_end) {
        const size_t c0 = done, cc = std::min(chunk, cols - done);
        double* pb = (double*)p->buf;
        const double* sp = src;
        const size_t r = rows, l = ld;
        p->pool->run(cc, [=](size_t lo, size_t hi) {
            for (size_t c = c0 + lo; c < c0 + hi; c++) std::memcpy(pb + c * r, sp + c * l, r * 8);
        });
        BOLA_CHECK(hipMemcpy2DAsync(dst + c0 * dld, dld * 8, pb + c0 * rows, rows * 8, rows * 8, cc,
                                    hipMemcpyHostToDevice, p->copy));
        ready = p->event();
        BOLA_CHECK(hipEventRecord(ready, p->copy));
        p->last = ready;
        done += cc;
    }
    if (ready && s) BOLA_CHECK(hipStreamWaitEvent(s, ready, 0));
}

// ---- download

HostDownload::HostDownload(bola_hip_handle_s* h_, double* dst_, size_t ld_, const double* src_, size_t dld_,
                           size_t rows_, size_t cols_, bool shares_upload_buffer)
    : h(h_), dst(dst_), ld(ld_), src(src_), dld(dld_), rows(rows_), cols(cols_) {
    if (rows == 0 || cols == 0) return;
    direct = rows * cols * 8 < kDirectBytes;
    // the strips land in the pinned buffer at the matrix's own positions (leading dim rows)
    if (!direct) p = shares_upload_buffer ? h->pin : pinned_stage(h, rows * cols * 8);
}

void HostDownload::strip(size_t r0, size_t r1, const hipStream_t* after, int nafter) {
    if (rows == 0 || cols == 0 || r1 <= r0) return;
    if (direct) {
        for (int i = 1; i < nafter; i++) {
            hipEvent_t e;
            BOLA_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            BOLA_CHECK(hipEventRecord(e, after[i]));
            BOLA_CHECK(hipStreamWaitEvent(after[0], e, 0));
            BOLA_CHECK(hipEventDestroy(e));
        }
        BOLA_CHECK(hipMemcpy2DAsync(dst + r0, ld * 8, src + r0, dld * 8, (r1 - r0) * 8, cols, hipMemcpyDeviceToHost,
                                    after[0]));
        strips.push_back(Strip{r0, r1, nullptr});
        return;
    }
    for (int i = 0; i < nafter; i++) {
        hipEvent_t e = p->event();
        BOLA_CHECK(hipEventRecord(e, after[i]));
        BOLA_CHECK(hipStreamWaitEvent(p->copy, e, 0));
    }
    double* pb = (double*)p->buf;
    BOLA_CHECK(hipMemcpy2DAsync(pb + r0, rows * 8, src + r0, dld * 8, (r1 - r0) * 8, cols, hipMemcpyDeviceToHost,
                                p->copy));
    hipEvent_t e = p->event();
    BOLA_CHECK(hipEventRecord(e, p->copy));
    p->last = e;
    strips.push_back(Strip{r0, r1, e});
}

void HostDownload::finish() {
    if (rows == 0 || cols == 0) return;
    if (direct) {
        BOLA_CHECK(hipStreamSynchronize(h->stream));
        return;
    }
    double* pb = (double*)p->buf;
    for (const Strip& s : strips) {
        BOLA_CHECK(hipEventSynchronize(s.ev));
        const size_t r0 = s.r0, nr = s.r1 - s.r0, r = rows, l = ld;
        double* d = dst;
        p->pool->run(cols, [=](size_t lo, size_t hi) {
            for (size_t c = lo; c < hi; c++) std::memcpy(d + c * l + r0, pb + c * r + r0, nr * 8);
        });
    }
}

void h2d_mat(bola_hip_handle_s* h, double* d, size_t dld, const double* hsrc, size_t ld, size_t rows, size_t cols) {
    HostUpload u(h, d, dld, hsrc, ld, rows, cols);
    u.need(h->stream, cols);
}

void d2h_mat(bola_hip_handle_s* h, double* hdst, size_t ld, const double* d, size_t dld, size_t rows, size_t cols) {
    HostDownload dl(h, hdst, ld, d, dld, rows, cols, false);
    hipStream_t st = h->stream;
    dl.strip(0, rows, &st, 1);
    dl.finish();
}

}  // namespace bola
