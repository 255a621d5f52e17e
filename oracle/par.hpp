// ORACLE — test infrastructure only (see bb.hpp).
//
// Fork-join data parallelism, restated from src/par.rs. Chunk boundaries
// depend only on lengths, so results are identical at any thread count.
#pragma once
#include <algorithm>
#include <atomic>
#include <cstddef>
#include <thread>
#include <vector>

namespace oracle {

// par.rs:17-24 — cached worker budget. The oracle lets the caller pin it
// (thread-count-independence tests and the 1/16/nproc CPU timings).
inline std::atomic<size_t>& max_workers_slot() {
    static std::atomic<size_t> slot{0};
    return slot;
}
inline size_t max_workers() {
    size_t v = max_workers_slot().load();
    if (v == 0) {
        unsigned hc = std::thread::hardware_concurrency();
        v = hc == 0 ? 1 : hc;
        max_workers_slot().store(v);
    }
    return v;
}
inline void set_max_workers(size_t n) { max_workers_slot().store(n == 0 ? 0 : n); }

// par.rs:30-34
inline size_t worker_count(size_t units, size_t min_per_thread) {
    size_t floor = std::max<size_t>(min_per_thread, 1);
    size_t by_work = units / floor;
    return std::min(std::max<size_t>(std::min(by_work, max_workers()), 1), std::max<size_t>(units, 1));
}

// par.rs:38-56 — f(base, ptr, len) on disjoint chunks.
template <typename T, typename F>
inline void for_each_indexed(T* items, size_t len, size_t min_chunk, F f) {
    size_t workers = worker_count(len, min_chunk);
    if (workers <= 1) {
        f(size_t(0), items, len);
        return;
    }
    size_t chunk = (len + workers - 1) / workers;
    std::vector<std::thread> threads;
    for (size_t base = 0; base < len; base += chunk) {
        size_t n = std::min(chunk, len - base);
        threads.emplace_back([&f, items, base, n]() { f(base, items + base, n); });
    }
    for (auto& t : threads) t.join();
}

}  // namespace oracle
