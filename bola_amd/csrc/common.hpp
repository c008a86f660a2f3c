// Shared definitions for the gfx950 dense-solver library: the handle, error plumbing,
// profiling phases and small device helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "bola_hip.h"

struct bola_hip_handle_s;

namespace bola {
struct PinnedStage;  // staging.hip

// Thrown inside the library only; every C-ABI entry point catches it.
struct HipFailure {
    hipError_t err;
    const char* what;
    int line;
};

#define BOLA_CHECK(expr)                                           \
    do {                                                           \
        hipError_t e_ = (expr);                                    \
        if (e_ != hipSuccess) throw ::bola::HipFailure{e_, #expr, __LINE__}; \
    } while (0)

#define BOLA_LAUNCH_CHECK() BOLA_CHECK(hipGetLastError())

// Growable device workspace owned by a handle.  Growth is stream-ordered on the handle's
// stream: the old buffer is freed after all work queued on the handle's streams, the new
// one is visible to all of them, and no other handle (or server slot) is stalled.
struct Workspace {
    void* ptr = nullptr;
    size_t bytes = 0;
    bola_hip_handle_s* owner = nullptr;
    void* get(size_t need);  // api.hip
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

// Where work queued behind a factorization stops when a pivot fails.
//
// `word` is the factorization's status word: 0, a global 1-based failing column (a zero
// pivot found by the panel of 32 columns starting at 32*floor((v-1)/32)), or negative
// (a device error: everything stops).  An operation names the column `col` of the first
// panel whose results it applies.  It runs only for panels before the failing one, so
// that a factorization stopped by a zero pivot leaves the matrix in the reference's
// state however its work was scheduled (getrf.rs:53-98 returns at the failing panel).
//   kcols == 0: all or nothing -- the operation applies the panel at `col` only;
//   kcols == 1: it applies consecutive panels from `col` on (k of a trailing GEMM counts
//               their columns, a list of row maps holds one per panel) and keeps those
//               before the failing panel.
// Converting a plain pointer gives "stop on any failure" (col beyond every matrix),
// which is what sequential factorizations (Cholesky, the recursive LU) need.
constexpr int kAbortAny = 1 << 30;
struct Abort {
    const int32_t* word = nullptr;
    int col = kAbortAny;
    int kcols = 0;
    Abort() = default;
    __host__ __device__ Abort(const int32_t* w) : word(w) {}  // NOLINT: implicit on purpose
    __host__ __device__ static Abort panel(const int32_t* w, int c) {
        Abort a(w);
        a.col = c;
        return a;
    }
    __host__ __device__ static Abort span(const int32_t* w, int c) {
        Abort a(w);
        a.col = c;
        a.kcols = 1;
        return a;
    }
};
// columns completed before the failure (everything when there is none)
__device__ __forceinline__ long abort_done(const Abort& a) {
    if (!a.word) return 1L << 62;
    const int v = __hip_atomic_load(a.word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (v == 0) return 1L << 62;
    if (v < 0) return -(1L << 62);
    return 32L * ((v - 1) / 32);
}
__device__ __forceinline__ bool abort_stop(const Abort& a) { return abort_done(a) <= (long)a.col; }
// how many of `count` units (k columns, or 32-column panels when `per` = 32) may run
__device__ __forceinline__ int abort_count(const Abort& a, int count, int per) {
    const long done = abort_done(a);
    if (!a.kcols) return done > (long)a.col ? count : 0;
    const long c = (done - (long)a.col) / per;
    return (int)(c < 0 ? 0 : (c > count ? count : c));
}

// Device scratch handed to a factorization instead of the handle's workspaces.
struct FixedScratch {
    void* ptr = nullptr;
    size_t bytes = 0;
};

// A composed row permutation: row `row[s]` receives the old contents of row `src[s]`.
// Produced by the LASWP compose step (one map per group of up to 64 interchanges) and by
// the LU panel kernel (one map per panel).
constexpr int ROWMAP_MAX = 128;
struct RowMap {
    int32_t count;
    int32_t pad;
    int64_t row[ROWMAP_MAX];
    int64_t src[ROWMAP_MAX];
};

struct PhaseStat {
    long calls = 0;
    double ms = 0.0;
    double flops = 0.0;
};

struct PendingEvent {
    std::string name;
    hipEvent_t start, stop;
    double flops;
};

}  // namespace bola

struct bola_hip_handle_s {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int num_cus = 0;
    bola::Workspace ws;         // factorization scratch (panel flags/slots, row maps)
    bola::Workspace ws_aux;     // LASWP group maps, IDAMAX / DOT partials
    bola::Workspace stage;      // operands of host-pointer calls
    bola::Workspace mws;        // mixed-precision solver: fp32 factors, vectors, GMRES basis
    bola::Workspace mws2;       // its FP64 fallback copy of A
    bola::Workspace bws;        // batched solves: staged descriptors, per-workgroup scratch
    bola::Workspace cws;        // Cholesky: the saved region of the matrix (restored when a pivot fails)
    // the handle's work has the device to itself (off for the server's slots): one-launch
    // triangular solves, the LU's CU-partitioned streams
    bool exclusive_device = true;
    int32_t* d_info = nullptr;  // device status word for factorizations
    std::string last_error;
    // profiling: per-phase GPU time measured with events on the handle's stream
    bool profile = false;
    std::vector<bola::PendingEvent> pending;
    std::map<std::string, bola::PhaseStat> stats;
    std::vector<hipEvent_t> event_pool;
    // second stream for work that overlaps the critical path, and ordering events
    hipStream_t stream2 = nullptr;
    hipStream_t stream3 = nullptr;  // a third one: side work that must not wait for the second
    // the LU's streams on disjoint sets of CUs: its critical path, and its bulk update and
    // left interchanges (lu_split: CUs i with (i / 8) % lu_split == 0 serve the critical
    // path; 0: no split, -1: not yet created)
    hipStream_t lu_crit = nullptr, lu_bulk = nullptr, lu_left = nullptr;
    int lu_split = -1;
    std::vector<hipEvent_t> sync_events;
    hipEvent_t ws_event = nullptr;  // orders workspace growth against the side streams
    bola::PinnedStage* pin = nullptr;  // host-pointer calls: pinned staging, copy stream, copy threads
    hipStream_t aux_stream() {
        if (!stream2) {
            if (hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking) != hipSuccess) stream2 = nullptr;
        }
        return stream2 ? stream2 : stream;
    }
    bola_hip_handle_s() {
        for (bola::Workspace* w : {&ws, &ws_aux, &stage, &mws, &mws2, &bws, &cws}) w->owner = this;
    }
    hipStream_t aux_stream3() {
        if (!stream3) {
            if (hipStreamCreateWithFlags(&stream3, hipStreamNonBlocking) != hipSuccess) stream3 = nullptr;
        }
        return stream3 ? stream3 : stream;
    }
};

namespace bola {

// RAII phase marker: records events around a phase when profiling is on.
struct Phase {
    bola_hip_handle_s* h;
    const char* name;
    double flops;
    hipEvent_t e0 = nullptr;
    hipStream_t st = nullptr;
    Phase(bola_hip_handle_s* h_, const char* n, double fl = 0.0, hipStream_t s = nullptr);
    ~Phase();
};

void set_error(bola_hip_handle_s* h, const std::string& msg);
// Records a HIP failure on h (or the thread's host-flavour error) and returns BOLA_HIP_ERR_DEVICE.
int device_failure(bola_hip_handle_s* h, const HipFailure& f);

// Every C-ABI entry point body sits between these: HIP failures become return codes.
#define BOLA_API_BEGIN try {
#define BOLA_API_END(h)                                              \
    }                                                                \
    catch (const ::bola::HipFailure& f_) {                           \
        return ::bola::device_failure(h, f_);                        \
    }                                                                \
    catch (...) {                                                    \
        ::bola::set_error(h, "unexpected host exception");           \
        return BOLA_HIP_ERR_DEVICE;                                  \
    }
bola_hip_handle_s* default_handle();
void collect_profile(bola_hip_handle_s* h);

inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// v when keep, else +0.  Written as a bitwise AND with a mask the compiler cannot see
// through, so that the load feeding v is not moved into a branch on keep (where it would
// be waited for on its own): loads from clamped addresses then all stay in flight at once.
__device__ __forceinline__ float keep_or_zero(float v, bool keep) {
    uint32_t m = keep ? 0xFFFFFFFFu : 0u;
    asm volatile("" : "+v"(m));
    return __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, v) & m);
}
__device__ __forceinline__ double keep_or_zero(double v, bool keep) {
    uint32_t m = keep ? 0xFFFFFFFFu : 0u;
    asm volatile("" : "+v"(m));
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint64_t mm = ((uint64_t)m << 32) | m;
    return __builtin_bit_cast(double, b & mm);
}

// The i-th ordering event of a handle (created on first use, timing disabled).
inline hipEvent_t handle_event(bola_hip_handle_s* h, int i) {
    while ((int)h->sync_events.size() <= i) {
        hipEvent_t e;
        BOLA_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        h->sync_events.push_back(e);
    }
    return h->sync_events[i];
}

// ---------------------------------------------------------------- launchers
// (defined in the .hip translation units; all enqueue on `st`)

// idamax.hip
void launch_idamax(hipStream_t st, size_t n, const double* x, int64_t* out, void* ws);
size_t idamax_ws_bytes(size_t n);

// blas12.hip
void launch_dscal(hipStream_t st, size_t n, double alpha, double* x, size_t incx);
void launch_daxpy(hipStream_t st, size_t n, double alpha, const double* x, size_t incx, double* y, size_t incy);
void launch_ddot(hipStream_t st, size_t n, const double* x, size_t incx, const double* y, size_t incy, double* out,
                 void* ws);
size_t ddot_ws_bytes(size_t n);
void launch_dger(hipStream_t st, size_t m, size_t n, double alpha, const double* x, const double* y, size_t ldy,
                 double* a, size_t lda);
void launch_fill(hipStream_t st, double* a, size_t lda, size_t m, size_t n, double v);
void launch_scale(hipStream_t st, double* a, size_t lda, size_t m, size_t n, double beta);
void launch_scale_tri(hipStream_t st, char uplo, double* a, size_t lda, size_t n, double beta);

// laswp.hip
size_t laswp_ws_bytes(size_t npiv);
void launch_laswp(hipStream_t st, size_t n, double* a, size_t lda, int32_t k1, int32_t k2, const int32_t* ipiv,
                  int32_t incx, void* ws, Abort abort = Abort());
// Applies `nmaps` row maps in order to the columns [0, left) and [skip_to, skip_to + right)
// of a.  When fix_ipiv is non-null, block 0 also adds fix_off to fix_ipiv[0..fix_cnt).
// Maps stop as `abort` says (one map per 32-column panel when it counts panels); the
// pivot fix only when the first map is applied.
void launch_rowmap_apply(hipStream_t st, double* a, size_t lda, size_t left, size_t skip_to, size_t right,
                         const RowMap* maps, int nmaps, Abort abort, int32_t* fix_ipiv, int fix_cnt, int fix_off);

// gemm.hip
// `abort`: read at kernel start; inside factorizations work queued behind a failed pivot
// is skipped (all of it, or with Abort::span the k columns of panels from the failing one on).
void launch_dgemm(hipStream_t st, char transa, char transb, size_t m, size_t n, size_t k, double alpha,
                  const double* a, size_t lda, const double* b, size_t ldb, double beta, double* c, size_t ldc,
                  Abort abort = Abort(), bool fast = false);
// slab: depth of one accumulate-then-fold step (64 = gemm.rs KC; 32 = the syrk.rs block
// loop, or one LU panel per step); mask: 0 all of C, 1 only rows in a later 32-block than
// the column, 2 earlier; edge_m / edge_n: first row / column (relative to C) that takes
// the reference's unfused edge path, -1 for this call's own m - m%8 / n - n%4.  fast: the
// register budget of two workgroups per CU (stand-alone products); otherwise one, which
// leaves more of the machine to a factorization's panel running beside it.
void launch_dgemm_ex(hipStream_t st, char transa, char transb, size_t m, size_t n, size_t k, double alpha,
                     const double* a, size_t lda, const double* b, size_t ldb, double beta, double* c, size_t ldc,
                     Abort abort, int slab, int mask, long edge_m = -1, long edge_n = -1, bool fast = false,
                     int ncu = 0);  // ncu: CUs the stream may use (0: the device's)
void set_num_cus(int n);

// staging.hip: host <-> device transfers of the host-pointer entry points through pinned
// memory.  An upload makes a stream wait for the columns it needs (need), so work on the
// first columns can start while later ones are still in transit; a download copies row
// strips back once the streams that finalise them have queued their work (strip), and
// finish() returns when the caller's matrix holds all of them.
void h2d_mat(bola_hip_handle_s* h, double* d, size_t dld, const double* hsrc, size_t ld, size_t rows, size_t cols);
void d2h_mat(bola_hip_handle_s* h, double* hdst, size_t ld, const double* d, size_t dld, size_t rows, size_t cols);
void release_pinned_stage(bola_hip_handle_s* h);
struct HostUpload {
    HostUpload(bola_hip_handle_s* h, double* dst, size_t dld, const double* src, size_t ld, size_t rows, size_t cols);
    void need(hipStream_t s, size_t c_end);  // s waits until columns [0, c_end) are on the device
    bola_hip_handle_s* h;
    double* dst;
    size_t dld;
    const double* src;
    size_t ld, rows, cols;
    bool direct = false;
    PinnedStage* p = nullptr;
    size_t chunk = 0, done = 0;
    hipEvent_t ready = nullptr;
};
struct HostDownload {
    // shares_upload_buffer: the matrix came in through a HostUpload of the same shape
    HostDownload(bola_hip_handle_s* h, double* dst, size_t ld, const double* src, size_t dld, size_t rows, size_t cols,
                 bool shares_upload_buffer);
    void strip(size_t r0, size_t r1, const hipStream_t* after, int nafter);  // rows [r0, r1) are final
    void finish();
    struct Strip {
        size_t r0, r1;
        hipEvent_t ev;
    };
    bola_hip_handle_s* h;
    double* dst;
    size_t ld;
    const double* src;
    size_t dld, rows, cols;
    bool direct = false;
    PinnedStage* p = nullptr;
    std::vector<Strip> strips;
};
// What a factorization tells its caller while it queues work: columns it is about to
// need, rows that no later work changes.
struct LuHooks {
    virtual void need_cols(hipStream_t s, int c_end) = 0;
    virtual void rows_done(int r0, int r1, const hipStream_t* after, int nafter) = 0;
    virtual ~LuHooks() = default;
};

// getrf.hip: the blocked LU (getrf.rs:37-101) or, recursive_entry, getrf2 (getrf2.rs:39-97)
// on the handle's streams; `fixed`: scratch to use instead of the handle's workspaces (at
// least lu_scratch_bytes(m, n, recursive_entry) bytes).
void lu_factor(bola_hip_handle_s* h, size_t m, size_t n, double* a, size_t lda, int32_t* ipiv, int32_t* info,
               bool recursive_entry, const FixedScratch* fixed = nullptr, LuHooks* hooks = nullptr);
size_t lu_scratch_bytes(size_t m, size_t n, bool recursive_entry);
// The no-allocation entry point's device scratch on the current device (bytes: its size).
void* noalloc_scratch(bola_hip_handle_s* h, size_t* bytes);

// mixed.hip: factorization objects (the C ABI's bola_hip_factor_*); memory from the
// stream-ordered allocator on the handle's stream when stream_ordered.
bola_hip_factor_s* factor_build(bola_hip_handle_s* h, int kind, int precision, size_t n, const double* a, size_t lda,
                                int* info, bool stream_ordered);
void factor_release(bola_hip_factor_s* f, hipStream_t st);

// trsm.hip
void launch_dtrsm(hipStream_t st, char side, char uplo, char transa, char diag, size_t m, size_t n, double alpha,
                  const double* a, size_t lda, double* b, size_t ldb, Abort abort = Abort());
// One-launch left/no-transpose solve, alpha = 1 (trsm.rs:95-170, same results as
// launch_dtrsm): m = 64 * 2^k with 2 to 256 leaves, n % 16 == 0, scratch at most 1 GiB.
// Its workgroups wait for one another, so it runs only where nothing else fills the
// device; ws: ln_persistent_ws_bytes(m, n) of scratch, 256-byte aligned.
bool ln_persistent_fits(size_t m, size_t n);
size_t ln_persistent_ws_bytes(size_t m, size_t n);
void launch_dtrsm_ln_persistent(hipStream_t st, bool upper, int unit, size_t m, size_t n, const double* a, size_t lda,
                                double* b, size_t ldb, void* ws);

}  // namespace bola
