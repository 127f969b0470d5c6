// Pinned staging for the host-pointer entry points (the sgemm_ocl-style use
// of the library, sgemm_ocl2.h:166-219: copy in, compute, copy out).
//
// Copies from or to pageable host memory go through two pinned slots
// (hipHostMalloc) in chunks: while the DMA engine moves chunk c between a
// slot and the device, host threads copy chunk c+1 between the caller's
// memory and the other slot.  The host-side copy is split over a small pool
// of threads, because one thread's memcpy is slower than the PCIe link.
// A slot is reused only after the event recorded behind its last DMA has
// completed, so calls may follow each other without waiting.
#include "internal.h"

#include <atomic>
#include <condition_variable>
#include <mutex>
#include <string.h>
#include <thread>
#include <vector>

namespace ugemm {

void release_staging();

namespace {

// Splits a memcpy over the calling thread and the pool's workers.
class CopyPool {
  public:
    explicit CopyPool(int workers) {
        for (int i = 0; i < workers; ++i) th_.emplace_back([this] { loop(); });
    }
    ~CopyPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    void copy(void *dst, const void *src, size_t n) {
        const size_t min_part = 1 << 20;
        const int parts = (int)std::min<size_t>(th_.size() + 1, std::max<size_t>(1, n / min_part));
        if (parts <= 1) {
            memcpy(dst, src, n);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            dst_ = (char *)dst;
            src_ = (const char *)src;
            n_ = n;
            parts_ = parts;
            next_.store(1);          // part 0 is the caller's
            left_ = parts;
            ++gen_;
        }
        cv_.notify_all();
        do_part(0);
        finish_one();
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [this] { return left_ == 0; });
    }

  private:
    void do_part(int p) {
        const size_t per = (n_ + parts_ - 1) / parts_;
        const size_t lo = (size_t)p * per, hi = std::min(n_, lo + per);
        if (lo < hi) memcpy(dst_ + lo, src_ + lo, hi - lo);
    }
    void finish_one() {
        std::lock_guard<std::mutex> lk(mu_);
        if (--left_ == 0) done_.notify_all();
    }
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            for (int p; (p = next_.fetch_add(1)) < parts_;) {
                do_part(p);
                finish_one();
            }
        }
    }
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    bool stop_ = false;
    uint64_t gen_ = 0;
    char *dst_ = nullptr;
    const char *src_ = nullptr;
    size_t n_ = 0;
    int parts_ = 0, left_ = 0;
    std::atomic<int> next_{0};
};

constexpr size_t kSlotBytes = 32u << 20;

struct Stager {
    void *slot[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    CopyPool *pool = nullptr;
    int device = -1;
    // Only the host threads are stopped at exit; the pinned slots belong to
    // the HIP runtime, which may already be shut down by then.
    ~Stager() { delete pool; }
};

Stager g_stage;   // used under the device mutex (one host-pointer call at a time)

int ensure_stager() {
    Stager &S = g_stage;
    int dev = 0;
    UGEMM_HIP_TRY(hipGetDevice(&dev));
    if (S.slot[0] && S.device == dev) return UGEMM_OK;
    if (S.slot[0]) release_staging();
    for (int i = 0; i < 2; ++i) {
        if (hipHostMalloc(&S.slot[i], kSlotBytes, hipHostMallocDefault) != hipSuccess) {
            S.slot[i] = nullptr;
            release_staging();
            return fail(UGEMM_ERR_NOMEM, "pinned staging: hipHostMalloc failed");
        }
        UGEMM_HIP_TRY(hipEventCreateWithFlags(&S.ev[i], hipEventDisableTiming));
    }
    if (!S.pool) {
        const unsigned hw = std::thread::hardware_concurrency();
        S.pool = new CopyPool((int)std::max(1u, std::min(7u, hw > 2 ? hw / 2 - 1 : 1u)));
    }
    S.device = dev;
    return UGEMM_OK;
}

size_t chunk_for(size_t bytes) {
    // At least four chunks for overlap, at most a slot, at least 1 MiB.
    size_t c = (bytes + 3) / 4;
    c = std::max<size_t>(c, 1u << 20);
    c = std::min(c, kSlotBytes);
    return (c + 255) & ~(size_t)255;
}

}  // namespace

void release_staging() {
    Stager &S = g_stage;
    for (int i = 0; i < 2; ++i) {
        if (S.ev[i]) {
            (void)hipEventSynchronize(S.ev[i]);
            (void)hipEventDestroy(S.ev[i]);
            S.ev[i] = nullptr;
        }
        if (S.slot[i]) {
            (void)hipHostFree(S.slot[i]);
            S.slot[i] = nullptr;
        }
    }
    S.device = -1;
}

int staged_h2d(void *dst, const void *src, size_t bytes, hipStream_t s) {
    if (bytes == 0) return UGEMM_OK;
    int rc = ensure_stager();
    if (rc) return rc;
    Stager &S = g_stage;
    const size_t chunk = chunk_for(bytes);
    int k = 0;
    for (size_t off = 0; off < bytes; off += chunk, k ^= 1) {
        const size_t n = std::min(chunk, bytes - off);
        UGEMM_HIP_TRY(hipEventSynchronize(S.ev[k]));   // the slot's previous DMA is done
        S.pool->copy(S.slot[k], (const char *)src + off, n);
        UGEMM_HIP_TRY(hipMemcpyAsync((char *)dst + off, S.slot[k], n, hipMemcpyHostToDevice, s));
        UGEMM_HIP_TRY(hipEventRecord(S.ev[k], s));
    }
    return UGEMM_OK;
}

int staged_d2h(void *dst, const void *src, size_t bytes, hipStream_t s) {
    if (bytes == 0) return UGEMM_OK;
    int rc = ensure_stager();
    if (rc) return rc;
    Stager &S = g_stage;
    const size_t chunk = chunk_for(bytes);
    const size_t nchunks = (bytes + chunk - 1) / chunk;
    auto issue = [&](size_t c) -> int {
        const size_t off = c * chunk, n = std::min(chunk, bytes - off);
        const int k = (int)(c & 1);
        UGEMM_HIP_TRY(hipEventSynchronize(S.ev[k]));   // an earlier use of the slot is finished
        UGEMM_HIP_TRY(hipMemcpyAsync(S.slot[k], (const char *)src + off, n, hipMemcpyDeviceToHost, s));
        UGEMM_HIP_TRY(hipEventRecord(S.ev[k], s));
        return UGEMM_OK;
    };
    if ((rc = issue(0))) return rc;
    for (size_t c = 0; c < nchunks; ++c) {
        if (c + 1 < nchunks && (rc = issue(c + 1))) return rc;   // next DMA runs while chunk c is copied out
        const int k = (int)(c & 1);
        UGEMM_HIP_TRY(hipEventSynchronize(S.ev[k]));
        const size_t off = c * chunk, n = std::min(chunk, bytes - off);
        S.pool->copy((char *)dst + off, S.slot[k], n);
    }
    return UGEMM_OK;
}

}  // namespace ugemm
