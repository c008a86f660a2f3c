// C ABI of libbola_hip.so: handles, errors, profiling, and every entry point of
// include/bola_hip.h.  Host-pointer entry points stage operands through the calling
// thread's default handle; device-pointer entry points enqueue on the handle's stream.
#include <algorithm>
#include <cstdio>
#include <mutex>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "common.hpp"

namespace bola {

// defined in potrf.hip
void chol_factor(bola_hip_handle_s* h, char uplo, size_t n, double* a, size_t lda, int32_t* info,
                 bool recursive_entry);
void launch_dsyrk(hipStream_t st, char uplo, char trans, size_t n, size_t k, double alpha, const double* a, size_t lda,
                  double beta, double* c, size_t ldc, Abort abort);

namespace {

thread_local std::string t_last_error;
thread_local bola_hip_handle_s* t_default = nullptr;

struct DefaultHandleReaper {
    ~DefaultHandleReaper() {}
};

hipEvent_t take_event(bola_hip_handle_s* h) {
    if (!h->event_pool.empty()) {
        hipEvent_t e = h->event_pool.back();
        h->event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    BOLA_CHECK(hipEventCreate(&e));
    return e;
}

}  // namespace

// Every phase is also a roctx range under its name (the reference's ScopedTimer names,
// profiling.rs:30-57), so a rocprofv3 --marker-trace shows the host side of each phase
// beside the kernels it enqueued.
Phase::Phase(bola_hip_handle_s* h_, const char* n, double fl, hipStream_t s) : h(h_), name(n), flops(fl) {
    roctxRangePushA(n);
    st = s ? s : (h ? h->stream : nullptr);
    if (h && h->profile) {
        e0 = take_event(h);
        BOLA_CHECK(hipEventRecord(e0, st));
    }
}

Phase::~Phase() {
    roctxRangePop();
    if (h && h->profile && e0) {
        hipEvent_t e1 = nullptr;
        try {
            e1 = take_event(h);
            (void)hipEventRecord(e1, st);
            h->pending.push_back(PendingEvent{name, e0, e1, flops});
        } catch (...) {
        }
    }
}

void collect_profile(bola_hip_handle_s* h) {
    if (h->pending.empty()) return;
    BOLA_CHECK(hipStreamSynchronize(h->stream));
    if (h->stream2) BOLA_CHECK(hipStreamSynchronize(h->stream2));
    if (h->stream3) BOLA_CHECK(hipStreamSynchronize(h->stream3));
    for (hipStream_t ls : {h->lu_crit, h->lu_bulk, h->lu_left})
        if (ls) BOLA_CHECK(hipStreamSynchronize(ls));
    for (auto& p : h->pending) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, p.start, p.stop);
        PhaseStat& s = h->stats[p.name];
        s.calls += 1;
        s.ms += ms;
        s.flops += p.flops;
        h->event_pool.push_back(p.start);
        h->event_pool.push_back(p.stop);
    }
    h->pending.clear();
}

void* Workspace::get(size_t need) {
    if (need <= bytes) return ptr;
    bola_hip_handle_s* h = owner;
    hipStream_t st = h->stream;
    // the free waits for everything queued on the handle's side streams ...
    hipStream_t side[5] = {h->stream2, h->stream3, h->lu_crit, h->lu_bulk, h->lu_left};
    if (!h->ws_event) BOLA_CHECK(hipEventCreateWithFlags(&h->ws_event, hipEventDisableTiming));
    for (hipStream_t s2 : side) {
        if (!s2) continue;
        BOLA_CHECK(hipEventRecord(h->ws_event, s2));
        BOLA_CHECK(hipStreamWaitEvent(st, h->ws_event, 0));
    }
    if (ptr) BOLA_CHECK(hipFreeAsync(ptr, st));
    ptr = nullptr;
    bytes = 0;
    const size_t nb = need + (need >> 2);
    BOLA_CHECK(hipMallocAsync(&ptr, nb, st));
    bytes = nb;
    // ... and the side streams' later work waits for the new buffer
    BOLA_CHECK(hipEventRecord(h->ws_event, st));
    for (hipStream_t s2 : side)
        if (s2) BOLA_CHECK(hipStreamWaitEvent(s2, h->ws_event, 0));
    return ptr;
}

void set_error(bola_hip_handle_s* h, const std::string& msg) {
    if (h) h->last_error = msg;
    t_last_error = msg;
}

bola_hip_handle_s* default_handle() {
    if (!t_default) {
        bola_hip_handle_t h = nullptr;
        if (bola_hip_create(&h) != 0) throw HipFailure{hipErrorInitializationError, "bola_hip_create", __LINE__};
        t_default = h;
    }
    return t_default;
}

int device_failure(bola_hip_handle_s* h, const HipFailure& f) {
    char buf[256];
    snprintf(buf, sizeof buf, "HIP error %d (%s) in %s at line %d", (int)f.err, hipGetErrorString(f.err), f.what,
             f.line);
    set_error(h, buf);
    return BOLA_HIP_ERR_DEVICE;
}

}  // namespace bola

using namespace bola;

namespace {

#define API_BEGIN(h) \
    try {
#define API_END(h)                                   \
    }                                                \
    catch (const HipFailure& f) {                    \
        return device_failure(h, f);                 \
    }                                                \
    catch (...) {                                    \
        set_error(h, "unexpected host exception");   \
        return BOLA_HIP_ERR_DEVICE;                  \
    }

// staging of host operands: one device region per operand, compact leading dimension
struct Stager {
    bola_hip_handle_s* h;
    std::vector<size_t> offs;
    size_t total = 0;
    size_t add(size_t bytes) {
        size_t o = total;
        total += (bytes + 255) & ~(size_t)255;
        offs.push_back(o);
        return o;
    }
    char* base = nullptr;
    void commit() { base = (char*)h->stage.get(std::max<size_t>(total, 256)); }
    template <class T>
    T* at(size_t off) { return (T*)(base + off); }
};

// A host-pointer LU: columns uploaded as the factorization reaches them, row strips
// downloaded as they become final.
struct HostLuHooks : LuHooks {
    HostUpload& up;
    HostDownload& dn;
    HostLuHooks(HostUpload& u, HostDownload& d) : up(u), dn(d) {}
    void need_cols(hipStream_t s, int c_end) override { up.need(s, (size_t)c_end); }
    void rows_done(int r0, int r1, const hipStream_t* after, int nafter) override {
        dn.strip((size_t)r0, (size_t)r1, after, nafter);
    }
};

bool is_ch(char c, char w) { return c == w || c == (char)(w + 32); }

int32_t* info_buf(bola_hip_handle_s* h) { return h->d_info; }

// Reads the two status words after a factorization and renders the reference's text.
int finish_lu(bola_hip_handle_s* h, int32_t* info_dev) {
    int32_t inf[2] = {0, 0};
    BOLA_CHECK(hipMemcpyAsync(inf, info_dev, sizeof inf, hipMemcpyDeviceToHost, h->stream));
    BOLA_CHECK(hipStreamSynchronize(h->stream));
    if (h->profile) collect_profile(h);
    if (inf[0] == 0) return 0;
    if (inf[0] == BOLA_HIP_ERR_DEVICE) {
        set_error(h, "device synchronisation timed out inside the panel factorization");
        return BOLA_HIP_ERR_DEVICE;
    }
    char buf[128];
    snprintf(buf, sizeof buf, "Matrix is singular. The pivot in column %d is zero.", inf[1]);
    set_error(h, buf);
    return inf[0];
}

int finish_chol(bola_hip_handle_s* h, int32_t* info_dev, bool print_global) {
    int32_t inf[2] = {0, 0};
    BOLA_CHECK(hipMemcpyAsync(inf, info_dev, sizeof inf, hipMemcpyDeviceToHost, h->stream));
    BOLA_CHECK(hipStreamSynchronize(h->stream));
    if (h->profile) collect_profile(h);
    if (inf[0] == 0) return 0;
    char buf[128];
    snprintf(buf, sizeof buf, "Matrix is not positive-definite. Failure at column %d.", print_global ? inf[0] : inf[1]);
    set_error(h, buf);
    return inf[0];
}

void gemm_dev(bola_hip_handle_s* h, char transa, char transb, size_t m, size_t n, size_t k, double alpha,
              const double* a, size_t lda, const double* b, size_t ldb, double beta, double* c, size_t ldc) {
    if (m == 0 || n == 0) return;
    Phase ph(h, "GEMM", 2.0 * (double)m * n * k);
    if (alpha == 0.0) {
        // gemm.rs:110-126: only the beta pass
        if (beta == 0.0) launch_fill(h->stream, c, ldc, m, n, 0.0);
        else if (beta != 1.0) launch_scale(h->stream, c, ldc, m, n, beta);
        return;
    }
    launch_dgemm(h->stream, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, nullptr, true);
}

}  // namespace

extern "C" {

// ------------------------------------------------------------------ handle

int bola_hip_create(bola_hip_handle_t* out) {
    bola_hip_handle_s* h = new bola_hip_handle_s();
    try {
        BOLA_CHECK(hipGetDevice(&h->device));
        hipDeviceProp_t prop;
        BOLA_CHECK(hipGetDeviceProperties(&prop, h->device));
        h->num_cus = prop.multiProcessorCount;
        set_num_cus(h->num_cus);
        BOLA_CHECK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        h->own_stream = true;
        BOLA_CHECK(hipMalloc(&h->d_info, 4 * sizeof(int32_t)));
    } catch (const HipFailure& f) {
        int r = device_failure(nullptr, f);
        delete h;
        *out = nullptr;
        return r;
    }
    *out = h;
    return 0;
}

int bola_hip_destroy(bola_hip_handle_t h) {
    if (!h) return 0;
    (void)hipStreamSynchronize(h->stream);
    if (h->stream2) (void)hipStreamSynchronize(h->stream2);
    if (h->stream3) (void)hipStreamSynchronize(h->stream3);
    for (hipStream_t ls : {h->lu_crit, h->lu_bulk, h->lu_left})
        if (ls) (void)hipStreamSynchronize(ls);
    for (auto& p : h->pending) {
        (void)hipEventDestroy(p.start);
        (void)hipEventDestroy(p.stop);
    }
    for (auto e : h->event_pool) (void)hipEventDestroy(e);
    h->ws.release();
    h->ws_aux.release();
    h->stage.release();
    h->mws.release();
    h->mws2.release();
    h->bws.release();
    h->cws.release();
    release_pinned_stage(h);
    if (h->d_info) (void)hipFree(h->d_info);
    for (auto e : h->sync_events) (void)hipEventDestroy(e);
    if (h->ws_event) (void)hipEventDestroy(h->ws_event);
    if (h->stream2) (void)hipStreamDestroy(h->stream2);
    if (h->stream3) (void)hipStreamDestroy(h->stream3);
    for (hipStream_t ls : {h->lu_crit, h->lu_bulk, h->lu_left})
        if (ls) (void)hipStreamDestroy(ls);
    if (h->own_stream) (void)hipStreamDestroy(h->stream);
    if (t_default == h) t_default = nullptr;
    delete h;
    return 0;
}

int bola_hip_set_stream(bola_hip_handle_t h, void* stream) {
    if (!h) return -1;
    API_BEGIN(h)
    BOLA_CHECK(hipStreamSynchronize(h->stream));
    if (h->own_stream) BOLA_CHECK(hipStreamDestroy(h->stream));
    h->own_stream = false;
    h->stream = (hipStream_t)stream;
    return 0;
    API_END(h)
}

void* bola_hip_get_stream(bola_hip_handle_t h) { return h ? (void*)h->stream : nullptr; }

int bola_hip_set_exclusive(bola_hip_handle_t h, int exclusive) {
    if (!h) return -1;
    h->exclusive_device = exclusive != 0;
    return 0;
}

int bola_hip_synchronize(bola_hip_handle_t h) {
    API_BEGIN(h)
    BOLA_CHECK(hipStreamSynchronize(h->stream));
    if (h->profile) collect_profile(h);
    return 0;
    API_END(h)
}

const char* bola_hip_error_string(bola_hip_handle_t h) { return h ? h->last_error.c_str() : ""; }
const char* bola_hip_last_error(void) { return t_last_error.c_str(); }

// ------------------------------------------------------------------ profiling

int bola_hip_profile_enable(bola_hip_handle_t h, int on) {
    if (!h) return -1;
    h->profile = on != 0;
    return 0;
}

int bola_hip_profile_reset(bola_hip_handle_t h) {
    if (!h) return -1;
    API_BEGIN(h)
    collect_profile(h);
    h->stats.clear();
    return 0;
    API_END(h)
}

int bola_hip_profile_get(bola_hip_handle_t h, const char* name, double* ms, long* calls, double* flops) {
    if (!h || !name) return -1;
    API_BEGIN(h)
    collect_profile(h);
    auto it = h->stats.find(name);
    if (ms) *ms = it == h->stats.end() ? 0.0 : it->second.ms;
    if (calls) *calls = it == h->stats.end() ? 0 : it->second.calls;
    if (flops) *flops = it == h->stats.end() ? 0.0 : it->second.flops;
    return 0;
    API_END(h)
}

int bola_hip_profile_report(bola_hip_handle_t h, double iterations, char* buf, size_t len) {
    if (!h || !buf || len == 0) return -1;
    API_BEGIN(h)
    collect_profile(h);
    std::string out;
    char line[256];
    snprintf(line, sizeof line, "\n--- Function Call Profiling (Average over %g iterations) ---\n", iterations);
    out += line;
    double total = 0.0;
    for (auto& kv : h->stats) total = std::max(total, kv.second.ms);
    if (total <= 0.0) {
        out += "No profiling data recorded.\n";
    } else {
        std::vector<std::pair<std::string, PhaseStat>> v(h->stats.begin(), h->stats.end());
        std::sort(v.begin(), v.end(), [](auto& a, auto& b) { return a.second.ms > b.second.ms; });
        snprintf(line, sizeof line, "%-10s | %-12s | %-15s | %-15s | %-10s\n", "Function", "Calls", "Avg Time/Iter",
                 "Avg Time/Call", "Percentage");
        out += line;
        out += std::string(78, '-') + "\n";
        for (auto& kv : v) {
            if (kv.second.calls == 0) continue;
            snprintf(line, sizeof line, "%-10s | %-12ld | %12.3fms | %12.3fus | %9.2f%%\n", kv.first.c_str(),
                     kv.second.calls, kv.second.ms / iterations, 1000.0 * kv.second.ms / kv.second.calls,
                     100.0 * kv.second.ms / total);
            out += line;
        }
    }
    snprintf(buf, len, "%s", out.c_str());
    return 0;
    API_END(h)
}

// ------------------------------------------------------------------ device flavour

int bola_hip_idamax_dev(bola_hip_handle_t h, size_t n, const double* x, int64_t* result_dev) {
    API_BEGIN(h)
    Phase ph(h, "IDAMAX");
    void* ws = h->ws_aux.get(idamax_ws_bytes(n));
    launch_idamax(h->stream, n, x, result_dev, ws);
    return 0;
    API_END(h)
}

int bola_hip_dlaswp_dev(bola_hip_handle_t h, size_t n, double* a, size_t lda, int32_t k1, int32_t k2,
                        const int32_t* ipiv, int32_t incx) {
    API_BEGIN(h)
    if (n == 0 || incx == 0 || k1 > k2) return 0;
    Phase ph(h, "LASWP");
    void* ws = h->ws_aux.get(laswp_ws_bytes((size_t)(k2 - k1) + 1));
    launch_laswp(h->stream, n, a, lda, k1, k2, ipiv, incx, ws, nullptr);
    return 0;
    API_END(h)
}

int bola_hip_dger_dev(bola_hip_handle_t h, size_t m, size_t n, double alpha, const double* x, const double* y,
                      size_t ldy, double* a, size_t lda) {
    API_BEGIN(h)
    Phase ph(h, "GER", 2.0 * (double)m * n);
    launch_dger(h->stream, m, n, alpha, x, y, ldy, a, lda);
    return 0;
    API_END(h)
}

int bola_hip_dscal_dev(bola_hip_handle_t h, size_t n, double alpha, double* x, size_t incx) {
    API_BEGIN(h)
    Phase ph(h, "SCAL", (double)n);
    launch_dscal(h->stream, n, alpha, x, incx);
    return 0;
    API_END(h)
}

int bola_hip_daxpy_dev(bola_hip_handle_t h, size_t n, double alpha, const double* x, size_t incx, double* y,
                       size_t incy) {
    API_BEGIN(h)
    Phase ph(h, "DAXPY", 2.0 * (double)n);
    launch_daxpy(h->stream, n, alpha, x, incx, y, incy);
    return 0;
    API_END(h)
}

int bola_hip_ddot_dev(bola_hip_handle_t h, size_t n, const double* x, size_t incx, const double* y, size_t incy,
                      double* result_dev) {
    API_BEGIN(h)
    Phase ph(h, "DOT", 2.0 * (double)n);
    void* ws = h->ws_aux.get(ddot_ws_bytes(n));
    launch_ddot(h->stream, n, x, incx, y, incy, result_dev, ws);
    return 0;
    API_END(h)
}

int bola_hip_dgemm_dev(bola_hip_handle_t h, char transa, char transb, size_t m, size_t n, size_t k, double alpha,
                       const double* a, size_t lda, const double* b, size_t ldb, double beta, double* c, size_t ldc) {
    API_BEGIN(h)
    gemm_dev(h, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
    return 0;
    API_END(h)
}

int bola_hip_dtrsm_dev(bola_hip_handle_t h, char side, char uplo, char transa, char diag, size_t m, size_t n,
                       double alpha, const double* a, size_t lda, double* b, size_t ldb) {
    API_BEGIN(h)
    Phase ph(h, "TRSM", (is_ch(side, 'L') ? (double)m * m * n : (double)n * n * m));
    launch_dtrsm(h->stream, side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb, nullptr);
    return 0;
    API_END(h)
}

int bola_hip_dsyrk_dev(bola_hip_handle_t h, char uplo, char trans, size_t n, size_t k, double alpha,
                       const double* a, size_t lda, double beta, double* c, size_t ldc) {
    API_BEGIN(h)
    Phase ph(h, "SYRK", (double)n * (n + 1) * k);
    launch_dsyrk(h->stream, uplo, trans, n, k, alpha, a, lda, beta, c, ldc, nullptr);
    return 0;
    API_END(h)
}

int bola_hip_dgetrf_dev(bola_hip_handle_t h, size_t m, size_t n, double* a, size_t lda, int32_t* ipiv,
                        int32_t* info_dev) {
    API_BEGIN(h)
    if (lda < std::max<size_t>(1, m)) {
        char buf[96];
        snprintf(buf, sizeof buf, "Argument 4 to getrf had an illegal value of %zu", lda);
        set_error(h, buf);
        return -4;
    }
    int32_t* info = info_dev ? info_dev : h->d_info;
    lu_factor(h, m, n, a, lda, ipiv, info, false);
    if (info_dev) return 0;
    return finish_lu(h, info);
    API_END(h)
}

int bola_hip_dgetrf_noalloc_dev(bola_hip_handle_t h, size_t m, size_t n, double* a, size_t lda, int32_t* ipiv,
                                int32_t* info_dev) {
    if (!h) return -1;
    API_BEGIN(h)
    if (lda < std::max<size_t>(1, m)) {
        char buf[96];
        snprintf(buf, sizeof buf, "Argument 4 to getrf had an illegal value of %zu", lda);
        set_error(h, buf);
        return -4;
    }
    size_t have = 0;
    FixedScratch fx;
    fx.ptr = noalloc_scratch(h, &have);
    fx.bytes = have;
    const size_t need = lu_scratch_bytes(m, n, false);
    if (need > have) {
        char buf[160];
        snprintf(buf, sizeof buf,
                 "getrf_noalloc: a %zu x %zu factorization needs %zu bytes of scratch, more than the %zu available",
                 m, n, need, have);
        set_error(h, buf);
        return -2;
    }
    int32_t* info = info_dev ? info_dev : h->d_info;
    lu_factor(h, m, n, a, lda, ipiv, info, false, &fx);
    if (info_dev) return 0;
    return finish_lu(h, info);
    API_END(h)
}

int bola_hip_dgetrf2_dev(bola_hip_handle_t h, size_t m, size_t n, double* a, size_t lda, int32_t* ipiv,
                         int32_t* info_dev) {
    API_BEGIN(h)
    if (lda < std::max<size_t>(1, m)) {
        char buf[96];
        snprintf(buf, sizeof buf, "Argument 4 to getrf2 had an illegal value of %zu", lda);
        set_error(h, buf);
        return -4;
    }
    int32_t* info = info_dev ? info_dev : h->d_info;
    lu_factor(h, m, n, a, lda, ipiv, info, true);
    if (info_dev) return 0;
    return finish_lu(h, info);
    API_END(h)
}

int bola_hip_dpotrf_dev(bola_hip_handle_t h, char uplo, size_t n, double* a, size_t lda, int32_t* info_dev) {
    API_BEGIN(h)
    if (uplo != 'U' && uplo != 'L') {
        set_error(h, "Argument 1 to potrf had an illegal value");
        return -1;
    }
    if (lda < std::max<size_t>(1, n)) {
        set_error(h, "Argument 4 to potrf had an illegal value");
        return -4;
    }
    int32_t* info = info_dev ? info_dev : h->d_info;
    chol_factor(h, uplo, n, a, lda, info, false);
    if (info_dev) return 0;
    return finish_chol(h, info, n <= 32);
    API_END(h)
}

int bola_hip_dpotrf2_dev(bola_hip_handle_t h, char uplo, size_t n, double* a, size_t lda, int32_t* info_dev) {
    API_BEGIN(h)
    int32_t* info = info_dev ? info_dev : h->d_info;
    chol_factor(h, uplo, n, a, lda, info, true);
    if (info_dev) return 0;
    return finish_chol(h, info, true);
    API_END(h)
}

int bola_hip_dgetrs_dev(bola_hip_handle_t h, size_t n, size_t nrhs, const double* lu, size_t lda, const int32_t* ipiv,
                        double* b, size_t ldb) {
    API_BEGIN(h)
    if (n == 0 || nrhs == 0) return 0;
    Phase ph(h, "GETRS", 2.0 * (double)n * n * nrhs);
    const bool pers = h->exclusive_device && ln_persistent_fits(n, nrhs);
    const size_t fo = (laswp_ws_bytes(n) + 255) & ~(size_t)255;
    char* ws = (char*)h->ws_aux.get(fo + (pers ? ln_persistent_ws_bytes(n, nrhs) : 0));
    {
        Phase pl(h, "LASWP");
        launch_laswp(h->stream, nrhs, b, ldb, 1, (int32_t)n, ipiv, 1, ws, nullptr);
    }
    {
        Phase pt(h, "TRSM", (double)n * n * nrhs);
        if (pers) {
            launch_dtrsm_ln_persistent(h->stream, false, 1, n, nrhs, lu, lda, b, ldb, ws + fo);
            launch_dtrsm_ln_persistent(h->stream, true, 0, n, nrhs, lu, lda, b, ldb, ws + fo);
        } else {
            launch_dtrsm(h->stream, 'L', 'L', 'N', 'U', n, nrhs, 1.0, lu, lda, b, ldb, nullptr);
            launch_dtrsm(h->stream, 'L', 'U', 'N', 'N', n, nrhs, 1.0, lu, lda, b, ldb, nullptr);
        }
    }
    return 0;
    API_END(h)
}

int bola_hip_dpotrs_dev(bola_hip_handle_t h, char uplo, size_t n, size_t nrhs, const double* f, size_t lda, double* b,
                        size_t ldb) {
    API_BEGIN(h)
    if (n == 0 || nrhs == 0) return 0;
    Phase ph(h, "POTRS", 2.0 * (double)n * n * nrhs);
    if (uplo == 'U') {
        launch_dtrsm(h->stream, 'L', 'U', 'T', 'N', n, nrhs, 1.0, f, lda, b, ldb, nullptr);
        launch_dtrsm(h->stream, 'L', 'U', 'N', 'N', n, nrhs, 1.0, f, lda, b, ldb, nullptr);
    } else {
        launch_dtrsm(h->stream, 'L', 'L', 'N', 'N', n, nrhs, 1.0, f, lda, b, ldb, nullptr);
        launch_dtrsm(h->stream, 'L', 'L', 'T', 'N', n, nrhs, 1.0, f, lda, b, ldb, nullptr);
    }
    return 0;
    API_END(h)
}

// ------------------------------------------------------------------ host flavour

size_t bola_hip_idamax(size_t n, const double* x) {
    bola_hip_handle_s* h = nullptr;
    try {
        h = default_handle();
        Stager s{h};
        size_t ox = s.add(std::max<size_t>(n, 1) * 8), oo = s.add(8);
        s.commit();
        if (n) BOLA_CHECK(hipMemcpyAsync(s.at<double>(ox), x, n * 8, hipMemcpyHostToDevice, h->stream));
        void* ws = h->ws_aux.get(idamax_ws_bytes(n));
        launch_idamax(h->stream, n, s.at<double>(ox), s.at<int64_t>(oo), ws);
        int64_t r = 0;
        BOLA_CHECK(hipMemcpyAsync(&r, s.at<int64_t>(oo), 8, hipMemcpyDeviceToHost, h->stream));
        BOLA_CHECK(hipStreamSynchronize(h->stream));
        return (size_t)r;
    } catch (const HipFailure& f) {
        device_failure(h, f);
        return (size_t)-1;
    }
}

int bola_hip_dlaswp(size_t n, double* a, size_t lda, int32_t k1, int32_t k2, const int32_t* ipiv, int32_t incx) {
    bola_hip_handle_s* h = nullptr;
    API_BEGIN(h)
    h = default_handle();
    if (n == 0 || incx == 0 || k1 > k2) return 0;
    const size_t absinc = (size_t)(incx < 0 ? -(int64_t)incx : incx);
    const size_t npiv_span = (size_t)(k1 - 1) + (size_t)(k2 - k1) * absinc + 1;
    int64_t maxrow = k2;
    for (int32_t i = k1; i <= k2; i++) {
        // both directions pair row i with ipiv[(k1-1) + (i-k1)*|incx|] (laswp.rs:30-37)
        size_t ix = (size_t)(k1 - 1) + (size_t)(i - k1) * absinc;
        if (ipiv[ix] < 1) {
            set_error(h, "laswp: pivot index below 1");
            return -6;
        }
        maxrow = std::max<int64_t>(maxrow, ipiv[ix]);
    }
    if ((size_t)maxrow > lda) {
        set_error(h, "laswp: pivot index beyond the leading dimension");
        return -6;
    }
    const size_t rows = (size_t)maxrow;
    Stager s{h};
    size_t oa = s.add(rows * n * 8), op = s.add(npiv_span * 4);
    s.commit();
    h2d_mat(h, s.at<double>(oa), rows, a, lda, rows, n);
    BOLA_CHECK(hipMemcpyAsync(s.at<int32_t>(op), ipiv, npiv_span * 4, hipMemcpyHostToDevice, h->stream));
    void* ws = h->ws_aux.get(laswp_ws_bytes((size_t)(k2 - k1) + 1));
    launch_laswp(h->stream, n, s.at<double>(oa), rows, k1, k2, s.at<int32_t>(op), incx, ws, nullptr);
    d2h_mat(h, a, lda, s.at<double>(oa), rows, rows, n);
    BOLA_CHECK(hipStreamSynchronize(h->stream));
    return 0;
    API_END(h)
}

static size_t span(size_t n, size_t inc) { return n ? (n - 1) * inc + 1 : 0; }

int bola_hip_dger(size_t m, size_t n, double alpha, const double* x, const double* y, size_t ldy, double* a,
                  size_t lda) {
    bola_hip_handle_s* h = nullptr;
    API_BEGIN(h)
    h = default_handle();
    if (m == 0 || n == 0) return 0;
    Stager s{h};
    size_t ox = s.add(m * 8), oy = s.add(span(n, ldy) * 8), oa = s.add(m * n * 8);
    s.commit();
    BOLA_CHECK(hipMemcpyAsync(s.at<double>(ox), x, m * 8, hipMemcpyHostToDevice, h->stream));
    BOLA_CHECK(hipMemcpyAsync(s.at<double>(oy), y, span(n, ldy) * 8, hipMemcpyHostToDevice, h->stream));
    h2d_mat(h, s.at<double>(oa), m, a, lda, m, n);
    launch_dger(h->stream, m, n, alpha, s.at<double>(ox), s.at<double>(oy), ldy, s.at<double>(oa), m);
    d2h_mat(h, a, lda, s.at<double>(oa), m, m, n);
    BOLA_CHECK(hipStreamSynchronize(h->stream));
    return 0;
    API_END(h)
}

int bola_hip_dscal(size_t n, double alpha, double* x, size_t incx) {
    bola_hip_handle_s* h = nullptr;
    API_BEGIN(h)
    h = default_handle();
    if (n == 0) return 0;
    Stager s{h};
    size_t ox = s.add(span(n, incx) * 8);
    s.commit();
    BOLA_CHECK(hipMemcpyAsync(s.at<double>(ox), x, span(n, incx) * 8, hipMemcpyHostToDevice, h->stream));
    launch_dscal(h->stream, n, alpha, s.at<double>(ox), incx);
    BOLA_CHECK(hipMemcpyAsync(x, s.at<double>(ox), span(n, incx) * 8, hipMemcpyDeviceToHost, h->stream));
    BOLA_CHECK(hipStreamSynchronize(h->stream));
    return 0;
    API_END(h)
}

int bola_hip_daxpy(size_t n, double alpha, const double* x, size_t incx, double* y, size_t incy) {
    bola_hip_handle_s* h = nullptr;
    API_BEGIN(h)
    h = default_handle();
    if (n == 0 || alpha == 0.0) return 0;
    Stager s{h};
    size_t ox = s.add(span(n, incx) * 8), oy = s.add(span(n, incy) * 8);
    s.commit();
    BOLA_CHECK(hipMemcpyAsync(s.at<double>(ox), x, span(n, incx) * 8, hipMemcpyHostToDevice, h->stream));
    BOLA_CHECK(hipMemcpyAsync(s.at<double>(oy), y, span(n, incy) * 8, hipMemcpyHostToDevice, h->stream));
    launch_daxpy(h->stream, n, alpha, s.at<double>(ox), incx, s.at<double>(oy), incy);
    BOLA_CHECK(hipMemcpyAsync(y, s.at<double>(oy), span(n, incy) * 8, hipMemcpyDeviceToHost, h->stream));
    BOLA_CHECK(hipStreamSynchronize(h->stream));
    return 0;
    API_END(h)
}

double bola_hip_ddot(size_t n, const double* x, size_t incx, const double* y, size_t incy) {
    bola_hip_handle_s* h = nullptr;
    try {
        h = default_handle();
        if (n == 0) return 0.0;
        Stager s{h};
        size_t ox = s.add(span(n, incx) * 8), oy = s.add(span(n, incy) * 8), oo = s.add(8);
        s.commit();
        BOLA_CHECK(hipMemcpyAsync(s.at<double>(ox), x, span(n, incx) * 8, hipMemcpyHostToDevice, h->stream));
        BOLA_CHECK(hipMemcpyAsync(s.at<double>(oy), y, span(n, incy) * 8, hipMemcpyHostToDevice, h->stream));
        void* ws = h->ws_aux.get(ddot_ws_bytes(n));
        launch_ddot(h->stream, n, s.at<double>(ox), incx, s.at<double>(oy), incy, s.at<double>(oo), ws);
        double r = 0.0;
        BOLA_CHECK(hipMemcpyAsync(&r, s.at<double>(oo), 8, hipMemcpyDeviceToHost, h->stream));
        BOLA_CHECK(hipStreamSynchronize(h->stream));
        return r;
    } catch (const HipFailure& f) {
        device_failure(h, f);
        return 0.0 / 0.0;
    }
}

int bola_hip_dgemm(char transa, char transb, size_t m, size_t n, size_t k, double alpha, const double* a, size_t lda,
                   const double* b, size_t ldb, double beta, double* c, size_t ldc) {
    bola_hip_handle_s* h = nullptr;
    API_BEGIN(h)
    h = default_handle();
    if (m == 0 || n == 0) return 0;
    const bool na = is_ch(transa, 'N'), nb = is_ch(transb, 'N');
    const size_t ar = na ? m : k, ac = na ? k : m, br = nb ? k : n, bc = nb ? n : k;
    Stager s{h};
    size_t oa = s.add(ar * ac * 8), ob = s.add(br * bc * 8), oc = s.add(m * n * 8);
    s.commit();
    const bool use_ab = alpha != 0.0 && k > 0;
    if (use_ab) {
        h2d_mat(h, s.at<double>(oa), ar, a, lda, ar, ac);
        h2d_mat(h, s.at<double>(ob), br, b, ldb, br, bc);
    }
    h2d_mat(h, s.at<double>(oc), m, c, ldc, m, n);
    gemm_dev(h, transa, transb, m, n, k, alpha, s.at<double>(oa), std::max<size_t>(ar, 1), s.at<double>(ob),
             std::max<size_t>(br, 1), beta, s.at<double>(oc), m);
    d2h_mat(h, c, ldc, s.at<double>(oc), m, m, n);
    BOLA_CHECK(hipStreamSynchronize(h->stream));
    return 0;
    API_END(h)
}

int bola_hip_dtrsm(char side, char uplo, char transa, char diag, size_t m, size_t n, double alpha, const double* a,
                   size_t lda, double* b, size_t ldb) {
    bola_hip_handle_s* h = nullptr;
    API_BEGIN(h)
    h = default_handle();
    if (m == 0 || n == 0) return 0;
    const size_t na = is_ch(side, 'L') ? m : n;
    Stager s{h};
    size_t oa = s.add(na * na * 8), ob = s.add(m * n * 8);
    s.commit();
    h2d_mat(h, s.at<double>(oa), na, a, lda, na, na);
    h2d_mat(h, s.at<double>(ob), m, b, ldb, m, n);
    launch_dtrsm(h->stream, side, uplo, transa, diag, m, n, alpha, s.at<double>(oa), na, s.at<double>(ob), m, nullptr);
    d2h_mat(h, b, ldb, s.at<double>(ob), m, m, n);
    BOLA_CHECK(hipStreamSynchronize(h->stream));
    return 0;
    API_END(h)
}

int bola_hip_dsyrk(char uplo, char trans, size_t n, size_t k, double alpha, const double* a, size_t lda, double beta,
                   double* c, size_t ldc) {
    bola_hip_handle_s* h = nullptr;
    API_BEGIN(h)
    h = default_handle();
    if (n == 0) return 0;
    const bool tr = !(trans == 'N' || trans == 'n');
    const size_t ar = tr ? k : n, ac = tr ? n : k;
    Stager s{h};
    size_t oa = s.add(std::max<size_t>(ar * ac, 1) * 8), oc = s.add(n * n * 8);
    s.commit();
    h2d_mat(h, s.at<double>(oa), ar, a, lda, ar, ac);
    h2d_mat(h, s.at<double>(oc), n, c, ldc, n, n);
    launch_dsyrk(h->stream, uplo, trans, n, k, alpha, s.at<double>(oa), std::max<size_t>(ar, 1), beta,
                 s.at<double>(oc), n, nullptr);
    d2h_mat(h, c, ldc, s.at<double>(oc), n, n, n);
    BOLA_CHECK(hipStreamSynchronize(h->stream));
    return 0;
    API_END(h)
}

static int getrf_host(size_t m, size_t n, double* a, size_t lda, int32_t* ipiv, bool rec) {
    bola_hip_handle_s* h = nullptr;
    API_BEGIN(h)
    h = default_handle();
    if (lda < std::max<size_t>(1, m)) {
        char buf[96];
        snprintf(buf, sizeof buf, "Argument 4 to %s had an illegal value of %zu", rec ? "getrf2" : "getrf", lda);
        set_error(h, buf);
        return -4;
    }
    if (m == 0 || n == 0) return 0;
    const size_t mn = std::min(m, n);
    Stager s{h};
    size_t oa = s.add(m * n * 8), op = s.add(mn * 4);
    s.commit();
    BOLA_CHECK(hipMemcpyAsync(s.at<int32_t>(op), ipiv, mn * 4, hipMemcpyHostToDevice, h->stream));
    HostUpload up(h, s.at<double>(oa), m, a, lda, m, n);
    HostDownload dn(h, a, lda, s.at<double>(oa), m, m, n, true);
    HostLuHooks hooks(up, dn);
    lu_factor(h, m, n, s.at<double>(oa), m, s.at<int32_t>(op), h->d_info, rec, nullptr, &hooks);
    dn.finish();
    BOLA_CHECK(hipMemcpyAsync(ipiv, s.at<int32_t>(op), mn * 4, hipMemcpyDeviceToHost, h->stream));
    return finish_lu(h, h->d_info);
    API_END(h)
}

int bola_hip_dgetrf(size_t m, size_t n, double* a, size_t lda, int32_t* ipiv) {
    return getrf_host(m, n, a, lda, ipiv, false);
}

int bola_hip_dgetrf2(size_t m, size_t n, double* a, size_t lda, int32_t* ipiv) {
    return getrf_host(m, n, a, lda, ipiv, true);
}

static int potrf_host(char uplo, size_t n, double* a, size_t lda, bool rec) {
    bola_hip_handle_s* h = nullptr;
    API_BEGIN(h)
    h = default_handle();
    if (!rec) {
        if (uplo != 'U' && uplo != 'L') {
            set_error(h, "Argument 1 to potrf had an illegal value");
            return -1;
        }
        if (lda < std::max<size_t>(1, n)) {
            set_error(h, "Argument 4 to potrf had an illegal value");
            return -4;
        }
    }
    if (n == 0) return 0;
    Stager s{h};
    size_t oa = s.add(n * n * 8);
    s.commit();
    h2d_mat(h, s.at<double>(oa), n, a, lda, n, n);
    chol_factor(h, uplo, n, s.at<double>(oa), n, h->d_info, rec);
    d2h_mat(h, a, lda, s.at<double>(oa), n, n, n);
    return finish_chol(h, h->d_info, rec || n <= 32);
    API_END(h)
}

int bola_hip_dpotrf(char uplo, size_t n, double* a, size_t lda) { return potrf_host(uplo, n, a, lda, false); }
int bola_hip_dpotrf2(char uplo, size_t n, double* a, size_t lda) { return potrf_host(uplo, n, a, lda, true); }

int bola_hip_dgetrs(size_t n, size_t nrhs, const double* lu, size_t lda, const int32_t* ipiv, double* b, size_t ldb) {
    bola_hip_handle_s* h = nullptr;
    API_BEGIN(h)
    h = default_handle();
    if (n == 0 || nrhs == 0) return 0;
    for (size_t i = 0; i < n; i++)
        if (ipiv[i] < 1 || (size_t)ipiv[i] > n) {
            set_error(h, "getrs: pivot index out of range");
            return -5;
        }
    Stager s{h};
    size_t oa = s.add(n * n * 8), op = s.add(n * 4), ob = s.add(n * nrhs * 8);
    s.commit();
    h2d_mat(h, s.at<double>(oa), n, lu, lda, n, n);
    BOLA_CHECK(hipMemcpyAsync(s.at<int32_t>(op), ipiv, n * 4, hipMemcpyHostToDevice, h->stream));
    h2d_mat(h, s.at<double>(ob), n, b, ldb, n, nrhs);
    int r = bola_hip_dgetrs_dev(h, n, nrhs, s.at<double>(oa), n, s.at<int32_t>(op), s.at<double>(ob), n);
    if (r) return r;
    d2h_mat(h, b, ldb, s.at<double>(ob), n, n, nrhs);
    BOLA_CHECK(hipStreamSynchronize(h->stream));
    return 0;
    API_END(h)
}

int bola_hip_dpotrs(char uplo, size_t n, size_t nrhs, const double* f, size_t lda, double* b, size_t ldb) {
    bola_hip_handle_s* h = nullptr;
    API_BEGIN(h)
    h = default_handle();
    if (n == 0 || nrhs == 0) return 0;
    Stager s{h};
    size_t oa = s.add(n * n * 8), ob = s.add(n * nrhs * 8);
    s.commit();
    h2d_mat(h, s.at<double>(oa), n, f, lda, n, n);
    h2d_mat(h, s.at<double>(ob), n, b, ldb, n, nrhs);
    int r = bola_hip_dpotrs_dev(h, uplo, n, nrhs, s.at<double>(oa), n, s.at<double>(ob), n);
    if (r) return r;
    d2h_mat(h, b, ldb, s.at<double>(ob), n, n, nrhs);
    BOLA_CHECK(hipStreamSynchronize(h->stream));
    return 0;
    API_END(h)
}

int bola_hip_dgesv_mixed(int precision, size_t n, size_t nrhs, const double* a, size_t lda, const double* b,
                         size_t ldb, double* x, size_t ldx, bola_hip_mixed_stats* stats) {
    bola_hip_handle_s* h = nullptr;
    API_BEGIN(h)
    h = default_handle();
    if (lda < std::max<size_t>(1, n)) {
        set_error(h, "Argument 5 to dgesv_mixed had an illegal value");
        return -5;
    }
    if (n == 0 || nrhs == 0) return bola_hip_dgesv_mixed_dev(h, precision, n, nrhs, nullptr, 1, nullptr, 1, nullptr, 1,
                                                             stats);
    Stager s{h};
    size_t oa = s.add(n * n * 8), ob = s.add(n * nrhs * 8), ox = s.add(n * nrhs * 8);
    s.commit();
    h2d_mat(h, s.at<double>(oa), n, a, lda, n, n);
    h2d_mat(h, s.at<double>(ob), n, b, ldb, n, nrhs);
    int r = bola_hip_dgesv_mixed_dev(h, precision, n, nrhs, s.at<double>(oa), n, s.at<double>(ob), n, s.at<double>(ox),
                                     n, stats);
    if (r < 0) return r;
    d2h_mat(h, x, ldx, s.at<double>(ox), n, n, nrhs);
    BOLA_CHECK(hipStreamSynchronize(h->stream));
    return r;
    API_END(h)
}

}  // extern "C"
