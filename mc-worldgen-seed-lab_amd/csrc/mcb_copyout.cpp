// mcb_copyout.cpp - device results to a caller's host buffer (usually pageable memory the caller
// allocated) faster than one pageable hipMemcpy.
//
// * While the device still computes, host threads write one byte into every 4 KiB page of the destination,
//   so its first-touch page faults (and the zeroing of fresh pages, huge pages included) run in parallel
//   and overlap the kernels instead of stalling the copy.
// * The bytes then travel in blocks: block k is copied device -> pinned staging on the caller's stream
//   while the threads move block k - 1 from its staging buffer into the destination.
// Waits are short polls (a block's copy takes tens of microseconds, and sleeping on a futex or on the
// device's completion signal costs about as much).  Staging (kStage pinned blocks of kBlock bytes, from
// hipHostMalloc, per device) and the threads are kept for the life of the process; calls are serialised.
// Destinations the device reaches directly (pinned host memory, device memory) take one plain copy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>

#include "mcb_internal.h"

namespace {

constexpr size_t kBlock = (size_t)4 << 20;     // bytes per staged block
constexpr int kStage = 3;                      // staged blocks in flight
constexpr size_t kDirect = (size_t)1 << 20;    // below this, one plain copy
constexpr int kThreads = 8;                    // host copy threads
constexpr size_t kPage = 4096;

// A fixed pool of threads that run one job (a memcpy or a page touch, split in kThreads parts) at a time.
// Workers poll for the next job for ~200 us before they sleep, so the jobs of one copy start quickly.
class CopyPool {
public:
    CopyPool() {
        for (int i = 0; i < kThreads; i++) std::thread([this, i] { run(i); }).detach();
    }
    // dst[0..n) = src[0..n) (src null: one byte written into each 4 KiB page of dst instead); returns
    // when every part is done
    void run_job(char *dst, const char *src, size_t n) {
        dst_ = dst;
        src_ = src;
        n_ = n;
        left_.store(kThreads, std::memory_order_relaxed);
        {
            std::lock_guard<std::mutex> lk(mu_);
            gen_.fetch_add(1, std::memory_order_release);
        }
        if (sleepers_.load(std::memory_order_acquire)) cv_.notify_all();
        while (left_.load(std::memory_order_acquire)) std::this_thread::yield();
    }

private:
    void run(int id) {
        uint64_t seen = 0;
        for (;;) {
            uint64_t g = gen_.load(std::memory_order_acquire);
            if (g == seen) {
                const auto t0 = std::chrono::steady_clock::now();
                while ((g = gen_.load(std::memory_order_acquire)) == seen &&
                       std::chrono::steady_clock::now() - t0 < std::chrono::microseconds(200))
                    std::this_thread::yield();
                if (g == seen) {
                    std::unique_lock<std::mutex> lk(mu_);
                    sleepers_.fetch_add(1, std::memory_order_acq_rel);
                    cv_.wait(lk, [&] { return (g = gen_.load(std::memory_order_acquire)) != seen; });
                    sleepers_.fetch_sub(1, std::memory_order_acq_rel);
                }
            }
            seen = g;
            char *d = dst_;
            const char *s = src_;
            const size_t n = n_;
            // part id of n, in whole pages
            const size_t part = ((n + kThreads - 1) / kThreads + kPage - 1) & ~(kPage - 1);
            const size_t a = std::min(n, part * (size_t)id), b = std::min(n, a + part);
            if (s) {
                if (b > a) memcpy(d + a, s + a, b - a);
            } else {
                for (size_t i = a; i < b; i += kPage) ((volatile char *)d)[i] = 0;
            }
            left_.fetch_sub(1, std::memory_order_acq_rel);
        }
    }
    std::mutex mu_;
    std::condition_variable cv_;
    std::atomic<uint64_t> gen_{0};
    std::atomic<int> left_{0}, sleepers_{0};
    char *dst_ = nullptr;
    const char *src_ = nullptr;
    size_t n_ = 0;
};

struct Staging {
    void *buf[kStage] = {nullptr, nullptr, nullptr};
    hipEvent_t ev[kStage] = {nullptr, nullptr, nullptr};
    bool ok = false;
};

std::mutex g_mu;

Staging *staging() {
    static std::map<int, Staging> per_device;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    Staging &S = per_device[dev];
    if (!S.ok) {
        for (int i = 0; i < kStage; i++) {
            if (!S.buf[i] && hipHostMalloc(&S.buf[i], kBlock, hipHostMallocDefault) != hipSuccess) return nullptr;
            if (!S.ev[i] && hipEventCreateWithFlags(&S.ev[i], hipEventDisableTiming) != hipSuccess) return nullptr;
        }
        S.ok = true;
    }
    return &S;
}

CopyPool &pool() {
    static CopyPool *p = new CopyPool();        // never destroyed: its threads outlive static destruction
    return *p;
}

hipError_t wait_event(hipEvent_t ev) {
    for (;;) {
        hipError_t e = hipEventQuery(ev);
        if (e != hipErrorNotReady) return e;
        std::this_thread::yield();
    }
}

// dst is memory the device reaches directly (pinned host memory, such as mcgpu_host_alloc's slots, or
// device memory): one DMA copies there at full rate, so it is not staged
bool direct_target(const void *dst) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, dst) != hipSuccess) {
        (void)hipGetLastError();                 // pageable memory the runtime does not know
        return false;
    }
    return a.type == hipMemoryTypeHost || a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

}  // namespace

hipError_t mcb::copy_to_host(void *dst, const void *src, size_t bytes, hipStream_t st) {
    if (bytes == 0) return hipSuccess;
    std::lock_guard<std::mutex> lk(g_mu);
    Staging *S = bytes < kDirect || direct_target(dst) ? nullptr : staging();
    if (!S) {
        hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, st);
        return e == hipSuccess ? hipStreamSynchronize(st) : e;
    }
    const size_t nblk = (bytes + kBlock - 1) / kBlock;
    auto issue = [&](size_t k) -> hipError_t {
        const size_t off = k * kBlock, len = std::min(kBlock, bytes - off);
        hipError_t e = hipMemcpyAsync(S->buf[k % kStage], (const char *)src + off, len, hipMemcpyDeviceToHost, st);
        return e == hipSuccess ? hipEventRecord(S->ev[k % kStage], st) : e;
    };
    // on an error, copies already queued into the staging blocks finish before the next call reuses them
    auto fail = [&](hipError_t e) {
        (void)hipStreamSynchronize(st);
        return e;
    };
    for (size_t k = 0; k < std::min(nblk, (size_t)kStage); k++)
        if (hipError_t e = issue(k)) return fail(e);
    pool().run_job((char *)dst, nullptr, bytes);
    for (size_t k = 0; k < nblk; k++) {
        if (hipError_t e = wait_event(S->ev[k % kStage])) return fail(e);
        const size_t off = k * kBlock, len = std::min(kBlock, bytes - off);
        pool().run_job((char *)dst + off, (const char *)S->buf[k % kStage], len);
        if (k + kStage < nblk)
            if (hipError_t e = issue(k + kStage)) return fail(e);
    }
    return hipSuccess;
}
