// Batched AIR verification with the Merkle path hashing on the device
// (a batched verification service over air::verify, air.rs:607-752 and
// fri.rs:686-859). Host threads parse each proof and replay its transcript
// with every other check (air_verify_deferred); the proofs' path checks are
// staged into the context's pinned buffers and hashed by k_verify_paths, one
// path per lane. Proofs go in chunks: the host verifies chunk k + 1 while
// the device hashes chunk k.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <functional>
#include <thread>

#include "prover.hpp"

namespace z0p {

namespace {

static_assert(sizeof(PathCheck) == PATH_CHECK_WORDS * 4, "path check records are 28 words");

constexpr size_t CHUNK = 64;  // proofs per device launch

// One chunk in flight: its proofs and where each one's checks start.
struct Slot {
    size_t first = 0, count = 0;
    std::vector<size_t> begin;
    bool busy = false;
};

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

int status_of(const std::exception_ptr& e) {
    try {
        std::rethrow_exception(e);
    } catch (const Status& s) {
        return s.code;
    } catch (const std::bad_alloc&) {
        return Z0P_ERR_OUT_OF_MEMORY;
    } catch (...) {
        return Z0P_ERR_DEVICE;
    }
}

}  // namespace

void air_verify_batch_device(Ctx& c, const z0p_air_params* params, Transcript* const* trs,
                             const uint8_t* const* proofs, const size_t* lens, size_t count, unsigned threads,
                             int* statuses) {
    if (count == 0) return;
    for (hipEvent_t& e : c.verify_done)
        if (!e) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    // On an error, the copies and kernels still queued must finish before
    // the staging they use is reused or freed.
    struct Drain {
        hipStream_t s;
        bool armed = true;
        ~Drain() {
            if (armed) (void)hipStreamSynchronize(s);
        }
    } drain{c.stream};
    Slot slots[2];
    std::vector<int> host(count, Z0P_OK);

    // Waits for a slot's device work and turns its flags into statuses.
    auto resolve = [&](int k) {
        Slot& s = slots[k];
        if (!s.busy) return;
        const double t0 = now_ms();
        hip_check(hipEventSynchronize(c.verify_done[k]));
        c.prof.host("~verify_wait_device", now_ms() - t0);
        const uint8_t* ok = c.verify_stage[k][2].host;
        for (size_t j = 0; j < s.count; ++j) {
            const size_t i = s.first + j;
            int st = host[i];
            for (size_t r = s.begin[j]; r < s.begin[j + 1]; ++r)
                if (!ok[r]) {
                    st = Z0P_ERR_BAD_MERKLE_PATH;  // recorded before any host failure
                    break;
                }
            statuses[i] = st;
        }
        s.busy = false;
    };

    std::vector<PathChecks> pcs(std::min(CHUNK, count));
    for (size_t first = 0, round = 0; first < count; first += CHUNK, ++round) {
        const size_t n = std::min(CHUNK, count - first);
        for (size_t k = 0; k < n; ++k) {
            pcs[k].checks.clear();
            pcs[k].siblings.clear();
        }
        // Host half of every proof in the chunk.
        auto one = [&](size_t k) {
            const size_t i = first + k;
            try {
                air_verify_deferred(params[i], *trs[i], proofs[i], lens[i], pcs[k]);
            } catch (...) {
                host[i] = status_of(std::current_exception());
            }
        };
        // Runs f(k) for k < n on the chosen threads.
        auto parallel = [&](const std::function<void(size_t)>& f) {
            if (threads == 1 || n == 1) {
                for (size_t k = 0; k < n; ++k) f(k);
            } else if (threads == 0) {
                c.host_workers().run(n, f);
            } else {
                std::atomic<size_t> next{0};
                auto work = [&] {
                    for (size_t k = next.fetch_add(1); k < n; k = next.fetch_add(1)) f(k);
                };
                std::vector<std::thread> pool;
                for (unsigned w = 1; w < std::min<size_t>(threads, n); ++w) pool.emplace_back(work);
                work();
                for (auto& t : pool) t.join();
            }
        };
        const double t_host = now_ms();
        parallel(one);
        c.prof.host("~verify_host", now_ms() - t_host);
        const double t_stage = now_ms();
        // The slot used two chunks ago is free once its results are read.
        const int k = int(round & 1);
        Slot& s = slots[k];
        resolve(k);
        s.first = first;
        s.count = n;
        s.begin.assign(n + 1, 0);
        size_t words = 0;
        for (size_t j = 0; j < n; ++j) {
            s.begin[j + 1] = s.begin[j] + pcs[j].checks.size();
            words += pcs[j].siblings.size();
        }
        const size_t total = s.begin[n];
        require(total < (size_t(1) << 32) && words / 8 < (size_t(1) << 32), Z0P_ERR_OUT_OF_MEMORY);
        PinnedPair& ckb = c.verify_stage[k][0];
        PinnedPair& sib = c.verify_stage[k][1];
        PinnedPair& okb = c.verify_stage[k][2];
        ckb.reserve(total * sizeof(PathCheck));
        sib.reserve(words * 4);
        okb.reserve(total);
        auto* ck = reinterpret_cast<PathCheck*>(ckb.host);
        auto* sw = reinterpret_cast<uint32_t*>(sib.host);
        std::vector<size_t> sib_base(n + 1, 0);
        for (size_t j = 0; j < n; ++j) sib_base[j + 1] = sib_base[j] + pcs[j].siblings.size() / 8;
        // Each proof's records into the pinned staging, in parallel.
        parallel([&](size_t j) {
            const PathChecks& pc = pcs[j];
            for (size_t r = 0; r < pc.checks.size(); ++r) {
                PathCheck rec = pc.checks[r];
                rec.sib += uint32_t(sib_base[j]);
                ck[s.begin[j] + r] = rec;
            }
            if (!pc.siblings.empty())
                std::memcpy(sw + 8 * sib_base[j], pc.siblings.data(), pc.siblings.size() * 4);
        });
        c.prof.host("~verify_stage", now_ms() - t_stage);
        if (total) {
            hip_check(hipMemcpyAsync(ckb.dev, ckb.host, total * sizeof(PathCheck), hipMemcpyHostToDevice, c.stream));
            if (words) hip_check(hipMemcpyAsync(sib.dev, sib.host, words * 4, hipMemcpyHostToDevice, c.stream));
            launch_verify_paths(c.stream, reinterpret_cast<const uint32_t*>(ckb.dev),
                                reinterpret_cast<const uint32_t*>(sib.dev), uint32_t(total), okb.dev);
            hip_check(hipGetLastError());
            hip_check(hipMemcpyAsync(okb.host, okb.dev, total, hipMemcpyDeviceToHost, c.stream));
        }
        hip_check(hipEventRecord(c.verify_done[k], c.stream));
        s.busy = true;
    }
    resolve(0);
    resolve(1);
    drain.armed = false;
}

}  // namespace z0p
