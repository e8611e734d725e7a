// Host memory copies of lattice-sized buffers split over a few threads.
//
// One thread's memcpy into freshly mapped pages moves 2-4 GB/s on the host: a
// 64 MB result takes 15-30 ms, longer than the device needs to compute it.  A
// copy of at least PAR_COPY_MIN bytes is cut into parts that the calling thread
// and a small pool of helper threads claim one at a time, so the copy and the
// destination's page faults run in parallel.  Several callers may copy at once;
// the pool serves them in turn and each caller also copies its own parts, so a
// busy pool never stalls a copy.
#pragma once

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>

namespace hostcopy {

constexpr size_t PAR_COPY_MIN = (size_t)4 << 20;
constexpr size_t PART_MIN = (size_t)1 << 20;
constexpr int HELPERS_MAX = 7;

struct Batch {
    char* dst;
    const char* src;
    size_t n, part;
    int parts;
    bool scan = false;  // also find whether a byte copied is above 1
    std::atomic<int> next{0};
    std::atomic<bool> above{false};
    int users = 0;     // helpers copying parts of it (under the pool's lock)
    int finished = 0;  // parts copied (under the pool's lock)
};

// Whether a byte of the n at p is above 1 (read as unsigned), 64 bits a step.
inline bool any_above_one(const char* p, size_t n) {
    uint64_t hi = 0;
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
        uint64_t w;
        std::memcpy(&w, p + i, 8);
        hi |= w;
    }
    for (; i < n; ++i) hi |= (uint8_t)p[i];
    return (hi & 0xFEFEFEFEFEFEFEFEull) != 0;
}

inline int copy_parts(Batch* b) {
    int k = 0;
    for (int i; (i = b->next.fetch_add(1, std::memory_order_relaxed)) < b->parts; ++k) {
        const size_t o = (size_t)i * b->part, len = std::min(b->part, b->n - o);
        std::memcpy(b->dst + o, b->src + o, len);
        if (b->scan && any_above_one(b->dst + o, len)) b->above.store(true, std::memory_order_relaxed);
    }
    return k;
}

struct Pool {
    std::mutex mu;
    std::condition_variable work, done;
    std::deque<Batch*> queue;
    int helpers = 0;

    void run() {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            work.wait(lk, [&] { return !queue.empty(); });
            Batch* b = queue.front();
            queue.pop_front();
            ++b->users;
            lk.unlock();
            const int k = copy_parts(b);
            lk.lock();
            b->finished += k;
            --b->users;
            done.notify_all();
        }
    }
};

// The pool lives until the process ends (its threads are detached and wait on
// the queue), so no copy can outlive it.
inline Pool& pool() {
    static Pool* p = [] {
        Pool* q = new Pool();
        const unsigned hw = std::thread::hardware_concurrency();
        q->helpers = (int)std::min<unsigned>(HELPERS_MAX, hw > 1 ? hw - 1 : 0);
        for (int i = 0; i < q->helpers; ++i) std::thread([q] { q->run(); }).detach();
        return q;
    }();
    return *p;
}

// Copies n bytes; with above, also stores whether a byte copied is above 1
// (read as unsigned), found while each part is still in cache.
inline void copy(void* dst, const void* src, size_t n, bool* above = nullptr) {
    Pool& P = pool();
    if (n < PAR_COPY_MIN || P.helpers == 0) {
        std::memcpy(dst, src, n);
        if (above) *above = any_above_one((const char*)dst, n);
        return;
    }
    Batch b;
    b.dst = (char*)dst;
    b.src = (const char*)src;
    b.n = n;
    b.scan = above != nullptr;
    // about four parts per thread, whole cache lines each
    b.part = std::max(PART_MIN, (n / (4 * (size_t)(P.helpers + 1)) + 63) / 64 * 64);
    b.parts = (int)((n + b.part - 1) / b.part);
    const int asks = std::min(P.helpers, b.parts - 1);
    {
        std::lock_guard<std::mutex> lk(P.mu);
        for (int i = 0; i < asks; ++i) P.queue.push_back(&b);
    }
    P.work.notify_all();
    const int k = copy_parts(&b);
    std::unique_lock<std::mutex> lk(P.mu);
    b.finished += k;
    // every part is claimed now: drop the requests no helper has taken
    P.queue.erase(std::remove(P.queue.begin(), P.queue.end(), &b), P.queue.end());
    P.done.wait(lk, [&] { return b.users == 0 && b.finished == b.parts; });
    if (above) *above = b.above.load();
}

}  // namespace hostcopy
