// upload.cpp -- copies between the device and large pageable host buffers (scene views, refit
// geometry, build inputs and outputs, ray batches, rendered frames).
//
// hipMemcpyAsync to or from pageable memory stages the data through the runtime's own buffers
// with one host thread.  Here a large copy is cut into chunks; host threads copy a chunk into
// (or out of) one of two pinned buffers owned by the context while the DMA engine moves the
// other, so the host copy and the PCIe transfer overlap.  An event per buffer keeps a buffer from
// being reused before the transfer that reads or fills it has finished.  Small copies go
// straight to hipMemcpyAsync.
#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

#include "api_internal.hpp"

namespace {

constexpr size_t kDirectMax = size_t(4) << 20;  // up to this, one plain hipMemcpyAsync
constexpr unsigned kMaxThreads = 8;

// Copies n bytes with up to nth threads (the caller's among them).  It never throws: it is called
// inside the C entry points, so a thread that cannot be started (std::system_error) or a vector
// that cannot grow leaves its part of the copy to the calling thread, and the threads already
// started are joined before returning.
void par_copy(char* dst, const char* src, size_t n, unsigned nth) noexcept {
    if (nth <= 1 || n < (size_t(1) << 20)) {
        std::memcpy(dst, src, n);
        return;
    }
    const size_t per = ((n + nth - 1) / nth + 4095) & ~size_t(4095);
    std::thread pool[kMaxThreads];
    unsigned started = 0;
    size_t a = per;  // [per, a) is covered by started threads
    for (; a < n && started < kMaxThreads; a += per) {
        const size_t b = std::min(n, a + per);
        try {
            pool[started] = std::thread([=] { std::memcpy(dst + a, src + a, b - a); });
        } catch (...) {
            break;
        }
        started++;
    }
    std::memcpy(dst, src, std::min(n, per));
    if (a < n) std::memcpy(dst + a, src + a, n - a);
    for (unsigned i = 0; i < started; i++) pool[i].join();
}

bool ensure_pinned(bc_ctx* c) {
    if (c->h_pin[0]) return true;
    void* p[2] = {nullptr, nullptr};
    hipEvent_t e[2] = {nullptr, nullptr};
    bool good = true;
    for (int i = 0; i < 2 && good; i++)
        good = bc_check(c, hipHostMalloc(&p[i], bc_ctx::kPinChunk, hipHostMallocDefault), "pinned upload buffer") &&
               bc_check(c, hipEventCreateWithFlags(&e[i], hipEventDisableTiming), "upload event");
    if (!good) {
        for (int i = 0; i < 2; i++) {
            if (p[i]) (void)hipHostFree(p[i]);
            if (e[i]) (void)hipEventDestroy(e[i]);
        }
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
    if (bytes <= kDirectMax) return bc_check(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st), what);
    if (!ensure_pinned(c)) return false;
    const unsigned hw = std::thread::hardware_concurrency();
    const unsigned nth = std::max(1u, std::min(kMaxThreads, hw));
    const char* s = static_cast<const char*>(src);
    char* d = static_cast<char*>(dst);
    int b = c->pin_next;
    for (size_t off = 0; off < bytes; off += bc_ctx::kPinChunk) {
        const size_t n = std::min(bc_ctx::kPinChunk, bytes - off);
        if (c->pin_busy[b] && !bc_check(c, hipEventSynchronize(c->pin_ev[b]), what)) return false;
        c->pin_busy[b] = false;
        par_copy(static_cast<char*>(c->h_pin[b]), s + off, n, nth);
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
    if (bytes <= kDirectMax)
        return bc_check(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st), what) &&
               bc_check(c, hipStreamSynchronize(st), what);
    if (!ensure_pinned(c)) return false;
    for (int i = 0; i < 2; i++) {  // uploads still reading the buffers finish first
        if (c->pin_busy[i] && !bc_check(c, hipEventSynchronize(c->pin_ev[i]), what)) return false;
        c->pin_busy[i] = false;
    }
    const unsigned hw = std::thread::hardware_concurrency();
    const unsigned nth = std::max(1u, std::min(kMaxThreads, hw));
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
        par_copy(d + k * K, static_cast<const char*>(c->h_pin[k & 1]), std::min(K, bytes - k * K), nth);
    }
    return true;
}

void bc_pinned_release(bc_ctx* c) {
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
}
