// Device context: one HIP stream, a device memory pool for per-proof
// buffers, a descriptor ring for batched kernel tables, a staging area for
// device-to-host collection, and cached per-geometry constant tables.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <utility>
#include <vector>

#include "kernels.hpp"
#include "profiler.hpp"
#include "status.hpp"

namespace z0p {

inline void hip_check(hipError_t e) {
    if (e != hipSuccess) {
        if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) throw Status(Z0P_ERR_OUT_OF_MEMORY);
        throw Status(Z0P_ERR_DEVICE);
    }
}

// Device-resident constant tables for one LDE geometry (log n, log B).
struct GeometryTables {
    int logn = 0, logb = 0;
    std::vector<void*> owned;
    PowTable w{}, winv{};            // omega_m^{+e}, omega_m^{-e}
    CosetTables coset{};
    const uint32_t* zinv = nullptr;  // [B] 1/(s_j^n - 1), Montgomery
    const uint32_t* zvan = nullptr;  // [B] s_j^n - 1, Montgomery
    const uint32_t* xn = nullptr;    // [n] omega_n^brev(q), Montgomery
    const uint32_t* finv = nullptr;  // [m/4] Omega_m^-brev(i), Montgomery
    PointInverses pinv;              // built on first use by an AIR proof (Ctx::point_inverses)
    ~GeometryTables() {
        for (void* p : owned) (void)hipFree(p);
    }
};

// First-fit allocator over one device allocation (host-side bookkeeping):
// proofs of different sizes come and go without device allocation calls.
class DevicePool {
   public:
    ~DevicePool() {
        if (base_) (void)hipFree(base_);
    }
    size_t capacity() const { return cap_; }
    size_t in_use() const { return in_use_; }
    // Resize; only valid while nothing is allocated.
    void reset_capacity(size_t bytes) {
        require(in_use_ == 0, Z0P_ERR_INVALID_ARGUMENT);
        if (base_) hip_check(hipFree(base_));
        base_ = nullptr;
        cap_ = 0;
        free_.clear();
        if (bytes == 0) return;
        hip_check(hipMalloc(&base_, bytes));
        cap_ = bytes;
        free_[0] = bytes;
    }
    bool alloc(size_t bytes, uint8_t*& out) {
        bytes = (bytes + 255) & ~size_t(255);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            size_t off = it->first, sz = it->second;
            free_.erase(it);
            if (sz > bytes) free_[off + bytes] = sz - bytes;
            in_use_ += bytes;
            out = base_ + off;
            return true;
        }
        return false;
    }
    void release(uint8_t* p, size_t bytes) {
        bytes = (bytes + 255) & ~size_t(255);
        size_t off = size_t(p - base_);
        in_use_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    uint8_t* base() const { return base_; }

   private:
    uint8_t* base_ = nullptr;
    size_t cap_ = 0, in_use_ = 0;
    std::map<size_t, size_t> free_;
};

// Pinned host ring mirrored in device memory: the descriptor tables and
// small uploads of one scheduler iteration. Pushes fill the host side and
// return the matching device address; flush() uploads everything pushed so
// far with one copy. In replay mode pushes only recompute addresses, which
// match the recorded ones when the same pushes repeat in the same order.
// Reset after the stream has been synchronised, or, while other work is in
// flight, restart in a region that work does not use (reset_to).
class DescRing final : public TableSink {
   public:
    ~DescRing() override {
        if (host_) (void)hipHostFree(host_);
        if (dev_) (void)hipFree(dev_);
    }
    void init(size_t bytes) {
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&host_), bytes, hipHostMallocDefault));
        hip_check(hipMalloc(reinterpret_cast<void**>(&dev_), bytes));
        cap_ = limit_ = bytes;
    }
    size_t capacity() const { return cap_; }
    const void* push(const void* p, size_t bytes) override {
        size_t at = (used_ + 255) & ~size_t(255);
        if (replay_) {
            require(at + bytes <= flushed_, Z0P_ERR_DEVICE);
        } else {
            require(at + bytes <= limit_, Z0P_ERR_OUT_OF_MEMORY);
            std::memcpy(host_ + at, p, bytes);
        }
        used_ = at + bytes;
        return dev_ + at;
    }
    size_t mark() const { return used_; }
    void flush(hipStream_t s) {
        if (used_ > flushed_)
            hip_check(hipMemcpyAsync(dev_ + flushed_, host_ + flushed_, used_ - flushed_, hipMemcpyHostToDevice, s));
        flushed_ = used_;
    }
    void replay_from(size_t mark) {
        used_ = mark;
        replay_ = true;
    }
    void end_replay() { replay_ = false; }
    void reset() { reset_to(0, cap_); }
    // Pushes fill [base, limit) from now on.
    void reset_to(size_t base, size_t limit) {
        used_ = flushed_ = base;
        limit_ = std::min(limit, cap_);
        replay_ = false;
    }

   private:
    uint8_t* host_ = nullptr;
    uint8_t* dev_ = nullptr;
    size_t cap_ = 0, limit_ = 0, used_ = 0, flushed_ = 0;
    bool replay_ = false;
};

// A pinned host buffer and a device buffer of the same size, grown on
// demand and kept (the batched verification's staging).
struct PinnedPair {
    uint8_t* host = nullptr;
    uint8_t* dev = nullptr;
    size_t cap = 0;
    void reserve(size_t bytes) {
        if (bytes <= cap) return;
        release();
        const size_t want = std::max<size_t>(bytes + bytes / 4, size_t(1) << 16);
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&host), want, hipHostMallocDefault));
        const hipError_t e = hipMalloc(reinterpret_cast<void**>(&dev), want);
        if (e != hipSuccess) {
            (void)hipHostFree(host);
            host = nullptr;
            hip_check(e);
        }
        cap = want;
    }
    void release() {
        if (host) (void)hipHostFree(host);
        if (dev) (void)hipFree(dev);
        host = dev = nullptr;
        cap = 0;
    }
    ~PinnedPair() { release(); }
};

// Persistent host worker threads for the per-proof transcript steps of a
// scheduler iteration (each proof owns its transcript, so they are
// independent). run() also works on the calling thread.
class HostPool {
   public:
    explicit HostPool(unsigned workers) {
        for (unsigned i = 0; i < workers; ++i) threads_.emplace_back([this] { loop(); });
    }
    ~HostPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : threads_) t.join();
    }
    size_t workers() const { return threads_.size(); }
    void run(size_t count, const std::function<void(size_t)>& f) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = &f;
            count_ = count;
            next_.store(0);
            busy_ = threads_.size();
            ++gen_;
        }
        cv_.notify_all();
        drain();
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [this] { return busy_ == 0; });
        job_ = nullptr;
    }

   private:
    void drain() {
        for (size_t i = next_.fetch_add(1); i < count_; i = next_.fetch_add(1)) (*job_)(i);
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
            drain();
            std::lock_guard<std::mutex> lk(mu_);
            if (--busy_ == 0) done_.notify_all();
        }
    }
    std::vector<std::thread> threads_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    const std::function<void(size_t)>* job_ = nullptr;
    size_t count_ = 0, busy_ = 0;
    std::atomic<size_t> next_{0};
    uint64_t gen_ = 0;
    bool stop_ = false;
};

// Launch shapes that change speed and never results; set by hand
// (z0p_ctx_set_schedule) or measured by the autotuner (tuner.cpp).
struct Schedule {
    int tail_max_log = 17;       // levels of at most 2^k digests go to the quad-lane tail
    int tail_group_max_log = 7;  // tail workgroup: at most 2^k digests (2^(k+1) threads)
    int tail_spread_log = 0;     // smaller groups so a level spans about 2^k workgroups
    int levels_fuse = 2;         // wide levels per launch (1 or 2)
    int small_max_log = 10;      // fold + commit in one workgroup for layers of at most 2^k values (0: off)
    int grind_blocks = 4096;     // grind workgroups per launch, shared by the proofs grinding
    int fold_fuse_max_log = 0;   // fold + next layer's leaves in one kernel for layers of at most 2^k values (0: off)
    int compose_fuse = 1;        // composition + FRI layer 0's leaves in one kernel (0: separate kernels)
    int trace_fuse = 1;          // linked column + trace leaves in one kernel (0: separate kernels)
    int subtree_levels = 0;      // tree levels above level 1 reduced in LDS per leaf workgroup (0: off)
    int compose_subtree = 0;     // the same inside the composition kernel (its own limit; 0: off)
    int graph = 0;               // each iteration's launches captured into one HIP graph (0: direct launches)
    int lde_split = 1;           // LDE row and column passes in kernels of their own (0: inside k_ntt)
    int grind_pool = 1024;       // workgroups of the pooled search when several proofs grind (0: per-proof workgroups)
    int fold_commit_max_log = 18;  // FRI layers of 2^8..2^k values committed in one launch (fold, leaves, tree; 0: off)
    int tree_top = 0;              // a proof tree's tail levels in one staged launch instead of tail launches (0: off)
    int wave_streams = 0;          // the scheduler's two waves on two streams, so their kernels overlap (0: one stream)

    // Digests per tail workgroup (log2) for a level of `count` digests.
    int tail_group_log(uint64_t count) const {
        int lc = 0;
        while ((uint64_t(1) << lc) < count) ++lc;
        int g = lc - tail_spread_log;
        if (g > tail_group_max_log) g = tail_group_max_log;
        if (g < TAIL_GROUP_MIN_LOG) g = TAIL_GROUP_MIN_LOG;
        return g;
    }
    bool valid() const {
        return tail_max_log >= TAIL_GROUP_MIN_LOG && tail_max_log <= 24 && tail_group_max_log >= TAIL_GROUP_MIN_LOG &&
               tail_group_max_log <= TAIL_GROUP_MAX_LOG && tail_spread_log >= 0 && tail_spread_log <= 16 &&
               (levels_fuse == 1 || levels_fuse == 2) && small_max_log >= 0 && small_max_log <= 10 &&
               (small_max_log == 0 || small_max_log >= 4) && grind_blocks >= 16 && grind_blocks <= (1 << 20) &&
               (fold_fuse_max_log == 0 || (fold_fuse_max_log >= 3 && fold_fuse_max_log <= 27)) &&
               (compose_fuse == 0 || compose_fuse == 1) && (trace_fuse == 0 || trace_fuse == 1) &&
               subtree_levels >= 0 && subtree_levels <= SUBTREE_MAX_LEVELS && compose_subtree >= 0 &&
               compose_subtree <= SUBTREE_MAX_LEVELS && (graph == 0 || graph == 1) &&
               (lde_split == 0 || lde_split == 1) && grind_pool >= 0 && grind_pool <= 8192 &&
               (fold_commit_max_log == 0 || (fold_commit_max_log >= 8 && fold_commit_max_log <= 27)) &&
               (tree_top == 0 || tree_top == 1) && (wave_streams == 0 || wave_streams == 1);
    }
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    Profiler prof;
    DevicePool pool;
    DescRing ring;
    size_t hbm_budget = 0;  // 0: the pool grows to fit (between batches)
    Schedule sched;

    // Staging for the words the host transcript needs: pinned host memory
    // the kernels write directly (stage_dev is its device address).
    uint32_t* stage_dev = nullptr;
    uint32_t* stage_host = nullptr;
    size_t stage_words = 0;
    // A second staging buffer: the scheduler alternates two waves of proofs,
    // and the wave on the device writes one buffer while the host reads the
    // other wave's words from the other (swap_wave exchanges them).
    uint32_t* alt_stage_dev = nullptr;
    uint32_t* alt_stage_host = nullptr;
    size_t alt_stage_words = 0;
    // Two waves once this many proofs are active (0: never).
    size_t waves_min = 8;
    // Completion flag in pinned host memory: the iteration's last kernel
    // stores the iteration number there after a system-scope fence.
    uint64_t* flag_host = nullptr;
    uint64_t* flag_dev = nullptr;
    uint64_t seq = 0;
    // The other wave's stream, completion flag and admission event (schedule
    // `wave_streams`): swap_wave exchanges them with `stream`, `flag_*` and
    // `setup_done` together with the staging buffers, so each wave's
    // launches, flag and staging stay on one side. `setup_done` is recorded
    // on a side's stream after the device work of admitting proofs there
    // (uploads), before that side's next wave; a wave launched on the other
    // side waits for it.
    hipStream_t alt_stream = nullptr;
    uint64_t* alt_flag_host = nullptr;
    uint64_t* alt_flag_dev = nullptr;
    hipEvent_t setup_done = nullptr, alt_setup_done = nullptr;
    // Mailbox for pre-enqueued phases (pinned host memory, device-mapped):
    // word 0 = gate counter the device polls, word 1 = gate timeout flag,
    // then MAILBOX_SLOTS constant slots of MAILBOX_SLOT bytes.
    static constexpr size_t MAILBOX_SLOT = 512, MAILBOX_SLOTS = 32, MAILBOX_HEAD = 256;
    uint8_t* mailbox_host = nullptr;
    uint8_t* mailbox_dev = nullptr;
    uint64_t gate = 0;
    bool pipeline_single = true;  // pre-enqueue the phases of a lone proof
    // Polls of host memory a pre-enqueued phase's gate makes before it
    // gives up (a few seconds); the gate then reports a timeout.
    uint32_t gate_spin = 1u << 22;

    // Device buffer of z0p_transcript_grind: the grind head and the running
    // minimum nonce (allocated on its first call).
    uint8_t* grind_buf = nullptr;

    // Scratch arena for the kernel-level diagnostic hooks.
    uint8_t* scratch = nullptr;
    size_t scratch_cap = 0, scratch_used = 0;

    uint32_t* tw4096 = nullptr;      // [2048] omega_4096^e, Montgomery
    uint32_t* tw4096_inv = nullptr;  // [2048] omega_4096^{-e}, Montgomery
    std::map<std::pair<int, int>, std::unique_ptr<GeometryTables>> tables;
    std::unique_ptr<HostPool> host_pool;  // created on first use by a batch
    HostPool& host_workers() {
        if (!host_pool) {
            const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
            host_pool.reset(new HostPool(std::min(hw, 16u) - 1));
        }
        return *host_pool;
    }
    // Batched verification staging: two slots of (checks, siblings, flags).
    PinnedPair verify_stage[2][3];
    hipEvent_t verify_done[2] = {nullptr, nullptr};

    Launch launch() { return Launch{stream, &prof, &ring, false, sched.lde_split != 0, uint32_t(sched.grind_pool)}; }

    // Runs a launch sequence dry (tables only), uploads the tables with one
    // copy, then replays it for real: as direct launches, or (schedule
    // `graph`, profiling off) captured into a HIP graph that is instantiated
    // and launched once. The sequence holds kernel launches only (no
    // memset or copy nodes).
    template <typename F>
    void enqueue(F&& launches) {
        const size_t mark = ring.mark();
        Launch dry = launch();
        dry.dry = true;
        launches(static_cast<const Launch&>(dry));
        ring.flush(stream);
        ring.replay_from(mark);
        if (sched.graph && !prof.enabled) {
            hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            hipGraph_t g = nullptr;
            try {
                launches(static_cast<const Launch&>(launch()));
            } catch (...) {
                (void)hipStreamEndCapture(stream, &g);
                if (g) (void)hipGraphDestroy(g);
                ring.end_replay();
                throw;
            }
            hip_check(hipStreamEndCapture(stream, &g));
            hipGraphExec_t x = nullptr;
            const hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            hip_check(e);
            graph_execs.push_back(x);
            hip_check(hipGraphLaunch(x, stream));
        } else {
            launches(static_cast<const Launch&>(launch()));
        }
        ring.end_replay();
    }

    // Executable graphs launched since the stream last drained; destroyed
    // once the work they launched has completed.
    std::vector<hipGraphExec_t> graph_execs;
    void release_graphs() {
        for (hipGraphExec_t x : graph_execs) (void)hipGraphExecDestroy(x);
        graph_execs.clear();
    }

    // Switches to the other wave's staging buffer and, with `streams`, also
    // to its stream, flag and admission event.
    void swap_wave(bool streams) {
        std::swap(stage_dev, alt_stage_dev);
        std::swap(stage_host, alt_stage_host);
        std::swap(stage_words, alt_stage_words);
        if (streams) {
            std::swap(stream, alt_stream);
            std::swap(flag_host, alt_flag_host);
            std::swap(flag_dev, alt_flag_dev);
            std::swap(setup_done, alt_setup_done);
        }
    }
    // Waves on two streams: only without per-launch profiling, whose event
    // resolution synchronises one stream at a time.
    bool wave_streams() const { return sched.wave_streams != 0 && !prof.enabled; }

    void reserve_stage(size_t words) {
        if (words <= stage_words) return;
        hip_check(hipStreamSynchronize(stream));
        if (stage_host) hip_check(hipHostFree(stage_host));
        stage_host = stage_dev = nullptr;
        size_t w = words + (words >> 2) + 1024;
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&stage_host), w * 4,
                                hipHostMallocMapped | hipHostMallocCoherent));
        hip_check(hipHostGetDevicePointer(reinterpret_cast<void**>(&stage_dev), stage_host, 0));
        stage_words = w;
    }

    // Waits until the flag reaches `s` (spinning: the kernels are short and
    // the host transcript step is on the critical path). The stream is
    // queried now and then so a device failure surfaces as an error.
    // sync_events: when profiling, synchronise to resolve kernel events
    // (not while pre-enqueued phases still wait on host gates).
    // reset_ring: also restart the descriptor ring (not while another
    // launch that reads it is still queued).
    void wait_flag(uint64_t s, bool sync_events = true, bool reset_ring = true) {
        const volatile uint64_t* f = flag_host;
        for (uint64_t spins = 1;; ++spins) {
            if (__atomic_load_n(const_cast<const uint64_t*>(f), __ATOMIC_ACQUIRE) >= s) break;
            if ((spins & 1023) == 0) {
                hipError_t e = hipStreamQuery(stream);
                if (e == hipSuccess) {
                    if (__atomic_load_n(const_cast<const uint64_t*>(f), __ATOMIC_ACQUIRE) >= s) break;
                    throw Status(Z0P_ERR_DEVICE);  // stream idle, flag never written
                }
                if (e != hipErrorNotReady) hip_check(e);
            }
        }
        if (prof.enabled && sync_events) {
            hip_check(hipStreamSynchronize(stream));
            prof.resolve();
        }
        if (reset_ring) {
            ring.reset();
            release_graphs();  // nothing else is queued: every graph has run
        }
    }

    void reserve_scratch(size_t bytes) {
        if (bytes > scratch_cap) {
            hip_check(hipStreamSynchronize(stream));
            if (scratch) hip_check(hipFree(scratch));
            scratch = nullptr;
            scratch_cap = 0;
            hip_check(hipMalloc(&scratch, bytes));
            scratch_cap = bytes;
        }
        scratch_used = 0;
    }
    template <typename T>
    T* scratch_alloc(size_t count) {
        size_t bytes = (count * sizeof(T) + 255) & ~size_t(255);
        require(scratch_used + bytes <= scratch_cap, Z0P_ERR_OUT_OF_MEMORY);
        T* p = reinterpret_cast<T*>(scratch + scratch_used);
        scratch_used += bytes;
        return p;
    }

    void sync() {
        hip_check(hipStreamSynchronize(alt_stream));
        hip_check(hipStreamSynchronize(stream));
        if (prof.enabled) prof.resolve();
        ring.reset();
        release_graphs();
    }

    // Device buffers released by IVC accumulators, kept by size for the
    // next accumulator on this context (freed with the context).
    std::multimap<size_t, void*> spare;
    void* take_buffer(size_t bytes) {
        auto it = spare.find(bytes);
        if (it != spare.end()) {
            void* p = it->second;
            spare.erase(it);
            return p;
        }
        void* p = nullptr;
        hip_check(hipMalloc(&p, bytes));
        return p;
    }
    // The caller has synchronised every use of `p` on this context's stream.
    void give_buffer(size_t bytes, void* p) {
        if (p) spare.emplace(bytes, p);
    }

    const GeometryTables& geometry(int logn, int logb);
    // The geometry's per-point inverses (PointInverses), built on first use.
    const PointInverses& point_inverses(int logn, int logb);

    ~Ctx();
};

}  // namespace z0p
