// upload.cpp -- copies between the device and large pageable host buffers (scene views, refit
// geometry, build inputs and outputs, ray batches, rendered frames).
//
// hipMemcpyAsync to or from pageable memory stages the data through the runtime's own buffers
// with one host thread.  Here a large copy is cut into chunks; host threads copy a chunk into
// (or out of) one of two pinned buffers owned by the context while the DMA engine moves the
// other, so the host copy and the PCIe transfer overlap.  An event per buffer keeps a buffer from
// being reused before the transfer that reads or fills it has finished.  Small copies (up to 64
// KB) go straight to hipMemcpyAsync.
#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "api_internal.hpp"

// Host threads that copy slices of a staged chunk, started once per context (with its pinned
// buffers) and kept until bc_destroy, instead of threads started for every chunk.  A copy hands
// out slices under the pool's mutex; the calling thread takes slices too, so a copy finishes
// whatever the workers do.
struct CopyPool {
    std::mutex m;
    std::condition_variable work, done;
    std::vector<std::thread> th;
    char* dst = nullptr;
    const char* src = nullptr;
    size_t n = 0, per = 0;
    unsigned parts = 0, next = 0, pending = 0;
    bool stop = false;

    // Takes the next slice of the current copy, if any (the mutex held).
    bool take(size_t& a, size_t& b) {
        if (next >= parts) return false;
        a = (size_t)next++ * per;
        b = std::min(n, a + per);
        return true;
    }
    void finish_one() {
        if (--pending == 0) done.notify_all();
    }
    void worker() {
        std::unique_lock<std::mutex> lk(m);
        for (;;) {
            work.wait(lk, [&] { return stop || next < parts; });
            if (stop) return;
            size_t a, b;
            while (take(a, b)) {
                lk.unlock();
                if (b > a) std::memcpy(dst + a, src + a, b - a);
                lk.lock();
                finish_one();
            }
        }
    }
};

namespace {

// Up to this, one plain hipMemcpyAsync.  Pageable copies of a few hundred KB cost up to ~80 us each
// in the runtime's own staging (a 64-row chunk's image and guides: 1.56 ms per closed-box frame of
// four chunks, against 0.88 ms through the pinned buffers), so only small ones go direct.
constexpr size_t kDirectMax = size_t(64) << 10;
constexpr unsigned kMaxThreads = 8;

// Starts the context's copy threads (up to kMaxThreads - 1 besides the caller).  It never
// throws: a thread that cannot be started leaves the pool smaller, an empty pool copies on the
// calling thread alone.
CopyPool* copy_pool(bc_ctx* c) {
    if (c->copy_pool) return c->copy_pool;
    CopyPool* p = nullptr;
    try {
        p = new CopyPool();
        const unsigned hw = std::thread::hardware_concurrency();
        const unsigned nth = std::max(1u, std::min(kMaxThreads, hw));
        for (unsigned i = 1; i < nth; i++) p->th.emplace_back([p] { p->worker(); });
    } catch (...) {
    }
    c->copy_pool = p;
    return p;
}

// Copies n bytes with the pool's threads and the caller's.  It never throws (it runs inside the
// C entry points): without a pool, or if its mutex fails, the caller copies alone.
void par_copy(CopyPool* p, char* dst, const char* src, size_t n) noexcept {
    if (!p || p->th.empty() || n < (size_t(1) << 20)) {
        std::memcpy(dst, src, n);
        return;
    }
    try {
        std::unique_lock<std::mutex> lk(p->m);
        const unsigned parts = (unsigned)p->th.size() + 1;
        p->dst = dst;
        p->src = src;
        p->n = n;
        p->per = ((n + parts - 1) / parts + 4095) & ~size_t(4095);
        p->parts = parts;
        p->next = 0;
        p->pending = parts;
        p->work.notify_all();
        size_t a, b;
        while (p->take(a, b)) {
            lk.unlock();
            if (b > a) std::memcpy(dst + a, src + a, b - a);
            lk.lock();
            p->finish_one();
        }
        p->done.wait(lk, [&] { return p->pending == 0; });
        p->parts = 0;
    } catch (...) {
        std::memcpy(dst, src, n);  // the mutex failed before any slice was handed out
    }
}

// Allocates the two pinned buffers on first use.  Returns false, leaving no error behind, when the
// runtime cannot provide them (pinned memory is a limited resource): the copies then go straight
// to hipMemcpyAsync as small ones do.
// The pinned buffers are about to be reused: they no longer hold rows rendered ahead.
void drop_ahead(bc_ctx* c) {
    c->pin_ahead = false;
    c->pin_pre_y0 = -1;
}

bool ensure_pinned(bc_ctx* c) {
    if (c->h_pin[0]) return true;
    void* p[2] = {nullptr, nullptr};
    hipEvent_t e[2] = {nullptr, nullptr};
    bool good = true;
    for (int i = 0; i < 2 && good; i++)
        good = hipHostMalloc(&p[i], bc_ctx::kPinChunk, hipHostMallocDefault) == hipSuccess &&
               hipEventCreateWithFlags(&e[i], hipEventDisableTiming) == hipSuccess;
    if (!good) {
        for (int i = 0; i < 2; i++) {
            if (p[i]) (void)hipHostFree(p[i]);
            if (e[i]) (void)hipEventDestroy(e[i]);
        }
        (void)hipGetLastError();  // a later launch check must not report this failure
        return false;
    }
    for (int i = 0; i < 2; i++) {
        c->h_pin[i] = p[i];
        c->pin_ev[i] = e[i];
        c->pin_busy[i] = false;
    }
    return true;
}

}  // namespace

bool bc_h2d(bc_ctx* c, void* dst, const void* src, size_t bytes, hipStream_t st, const char* what) {
    if (bytes == 0) return true;
    if (bytes <= kDirectMax || !ensure_pinned(c))
        return bc_check(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st), what);
    drop_ahead(c);
    CopyPool* pool = copy_pool(c);
    const char* s = static_cast<const char*>(src);
    char* d = static_cast<char*>(dst);
    int b = c->pin_next;
    for (size_t off = 0; off < bytes; off += bc_ctx::kPinChunk) {
        const size_t n = std::min(bc_ctx::kPinChunk, bytes - off);
        if (c->pin_busy[b] && !bc_check(c, hipEventSynchronize(c->pin_ev[b]), what)) return false;
        c->pin_busy[b] = false;
        par_copy(pool, static_cast<char*>(c->h_pin[b]), s + off, n);
        if (!bc_check(c, hipMemcpyAsync(d + off, c->h_pin[b], n, hipMemcpyHostToDevice, st), what) ||
            !bc_check(c, hipEventRecord(c->pin_ev[b], st), what))
            return false;
        c->pin_busy[b] = true;
        b ^= 1;
    }
    c->pin_next = b;
    return true;
}

bool bc_d2h(bc_ctx* c, void* dst, const void* src, size_t bytes, hipStream_t st, const char* what) {
    if (bytes == 0) return true;
    if (bytes <= kDirectMax || !ensure_pinned(c))
        return bc_check(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st), what) &&
               bc_check(c, hipStreamSynchronize(st), what);
    drop_ahead(c);
    for (int i = 0; i < 2; i++) {  // uploads still reading the buffers finish first
        if (c->pin_busy[i] && !bc_check(c, hipEventSynchronize(c->pin_ev[i]), what)) return false;
        c->pin_busy[i] = false;
    }
    CopyPool* pool = copy_pool(c);
    const char* s = static_cast<const char*>(src);
    char* d = static_cast<char*>(dst);
    const size_t K = bc_ctx::kPinChunk, nchunk = (bytes + K - 1) / K;
    // chunk k lands in buffer k & 1; chunk k + 1 is queued before chunk k is copied out, into
    // the buffer whose chunk (k - 1) the host has already copied out
    auto queue = [&](size_t k) {
        const size_t n = std::min(K, bytes - k * K);
        return bc_check(c, hipMemcpyAsync(c->h_pin[k & 1], s + k * K, n, hipMemcpyDeviceToHost, st), what) &&
               bc_check(c, hipEventRecord(c->pin_ev[k & 1], st), what);
    };
    if (!queue(0)) return false;
    for (size_t k = 0; k < nchunk; k++) {
        if (k + 1 < nchunk && !queue(k + 1)) return false;
        if (!bc_check(c, hipEventSynchronize(c->pin_ev[k & 1]), what)) return false;
        par_copy(pool, d + k * K, static_cast<const char*>(c->h_pin[k & 1]), std::min(K, bytes - k * K));
    }
    return true;
}

// Two device-to-host copies that each fit one pinned buffer (a rendered chunk's image and guides):
// both transfers are queued into the pinned buffers and waited for once, then copied out, instead
// of two pageable copies with a wait each.  Larger ones take bc_d2h.
bool bc_d2h_pair(bc_ctx* c, void* dst0, const void* src0, size_t n0, void* dst1, const void* src1, size_t n1,
                 hipStream_t st, const char* what) {
    if (n0 > bc_ctx::kPinChunk || n1 > bc_ctx::kPinChunk || n0 + n1 <= kDirectMax || !ensure_pinned(c))
        return bc_d2h(c, dst0, src0, n0, st, what) && bc_d2h(c, dst1, src1, n1, st, what);
    drop_ahead(c);
    for (int i = 0; i < 2; i++) {  // uploads still reading the buffers finish first
        if (c->pin_busy[i] && !bc_check(c, hipEventSynchronize(c->pin_ev[i]), what)) return false;
        c->pin_busy[i] = false;
    }
    if ((n0 && !bc_check(c, hipMemcpyAsync(c->h_pin[0], src0, n0, hipMemcpyDeviceToHost, st), what)) ||
        (n1 && !bc_check(c, hipMemcpyAsync(c->h_pin[1], src1, n1, hipMemcpyDeviceToHost, st), what)) ||
        !bc_check(c, hipStreamSynchronize(st), what))
        return false;
    CopyPool* pool = copy_pool(c);
    if (n0) par_copy(pool, static_cast<char*>(dst0), static_cast<const char*>(c->h_pin[0]), n0);
    if (n1) par_copy(pool, static_cast<char*>(dst1), static_cast<const char*>(c->h_pin[1]), n1);
    return true;
}

int bc_d2h_hold(bc_ctx* c, const void* src0, size_t n0, const void* src1, size_t n1, hipStream_t st,
                const char* what) {
    if (n0 > bc_ctx::kPinChunk || n1 > bc_ctx::kPinChunk || !ensure_pinned(c)) return 0;
    drop_ahead(c);
    for (int i = 0; i < 2; i++) {  // uploads still reading the buffers finish first
        if (c->pin_busy[i] && !bc_check(c, hipEventSynchronize(c->pin_ev[i]), what)) return -1;
        c->pin_busy[i] = false;
    }
    if ((n0 && !bc_check(c, hipMemcpyAsync(c->h_pin[0], src0, n0, hipMemcpyDeviceToHost, st), what)) ||
        (n1 && !bc_check(c, hipMemcpyAsync(c->h_pin[1], src1, n1, hipMemcpyDeviceToHost, st), what)) ||
        !bc_check(c, hipStreamSynchronize(st), what))
        return -1;
    c->pin_ahead = true;
    return 1;
}

int bc_d2h_prefetch(bc_ctx* c, const void* src0, size_t n0, const void* src1, size_t n1, hipStream_t st,
                    const char* what) {
    if (n0 + n1 > bc_ctx::kPinChunk || !ensure_pinned(c)) return -1;
    drop_ahead(c);
    const int b = c->pin_next;
    if (c->pin_busy[b] && !bc_check(c, hipEventSynchronize(c->pin_ev[b]), what)) return -2;
    c->pin_busy[b] = false;
    char* h = static_cast<char*>(c->h_pin[b]);
    if ((n0 && !bc_check(c, hipMemcpyAsync(h, src0, n0, hipMemcpyDeviceToHost, st), what)) ||
        (n1 && !bc_check(c, hipMemcpyAsync(h + n0, src1, n1, hipMemcpyDeviceToHost, st), what)) ||
        !bc_check(c, hipEventRecord(c->pin_ev[b], st), what))
        return -2;
    c->pin_busy[b] = true;
    c->pin_next = b ^ 1;
    return b;
}

void bc_host_copy(bc_ctx* c, void* dst, const void* src, size_t n) {
    if (n <= kDirectMax) {
        std::memcpy(dst, src, n);
        return;
    }
    par_copy(copy_pool(c), static_cast<char*>(dst), static_cast<const char*>(src), n);
}

void bc_pinned_release(bc_ctx* c) {
    if (CopyPool* p = c->copy_pool) {
        try {
            std::lock_guard<std::mutex> lk(p->m);
            p->stop = true;
        } catch (...) {
            p->stop = true;
        }
        p->work.notify_all();
        for (std::thread& t : p->th)
            if (t.joinable()) t.join();
        delete p;
        c->copy_pool = nullptr;
    }
    for (int i = 0; i < 2; i++) {
        if (c->pin_ev[i]) {
            if (c->pin_busy[i]) (void)hipEventSynchronize(c->pin_ev[i]);
            (void)hipEventDestroy(c->pin_ev[i]);
        }
        if (c->h_pin[i]) (void)hipHostFree(c->h_pin[i]);
        c->h_pin[i] = nullptr;
        c->pin_ev[i] = nullptr;
        c->pin_busy[i] = false;
    }
    drop_ahead(c);
}
