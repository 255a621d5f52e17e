// Per-kernel timing with HIP events recorded on the launching stream.
// When enabled, every launcher brackets its kernel with an event pair; the
// host resolves elapsed times after the next stream synchronisation, or
// without waiting for the pairs that have completed (resolve_ready).
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

namespace z0p {

struct Profiler {
    bool enabled = false;
    std::string only;  // when not empty: time launches of this name only
    struct Pending {
        const char* name;
        hipEvent_t start, stop;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    struct Stat {
        uint64_t launches = 0;
        double total_ms = 0.0;
    };
    std::map<std::string, Stat> stats;

    hipEvent_t take() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }

    // Host-side intervals of the scheduler (build + launch, device wait,
    // transcript steps), recorded under the given name whether or not
    // kernel events are enabled (two clock reads per iteration).
    void host(const char* name, double ms) {
        Stat& st = stats[name];
        st.launches += 1;
        st.total_ms += ms;
    }

    // Call after the stream has been synchronised.
    void resolve() { resolve_pairs(false); }
    // The pairs whose stop event has completed; the others stay pending.
    void resolve_ready() { resolve_pairs(true); }
    void resolve_pairs(bool ready_only) {
        size_t keep = 0;
        for (size_t i = 0; i < pending.size(); ++i) {
            Pending& p = pending[i];
            if (ready_only && p.stop && hipEventQuery(p.stop) != hipSuccess) {
                pending[keep++] = p;
                continue;
            }
            float ms = 0.f;
            if (p.start && p.stop && hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
                Stat& s = stats[p.name];
                s.launches += 1;
                s.total_ms += ms;
            }
            if (p.start) pool.push_back(p.start);
            if (p.stop) pool.push_back(p.stop);
        }
        pending.resize(keep);
    }

    ~Profiler() {
        for (auto& p : pending) {
            if (p.start) (void)hipEventDestroy(p.start);
            if (p.stop) (void)hipEventDestroy(p.stop);
        }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

struct ProfScope {
    Profiler* p;
    hipStream_t s;
    Profiler::Pending ev{nullptr, nullptr, nullptr};
    ProfScope(Profiler* prof, const char* name, hipStream_t stream) : p(prof && prof->enabled ? prof : nullptr), s(stream) {
        if (p && !p->only.empty() && p->only != name) p = nullptr;
        if (!p) return;
        ev.name = name;
        ev.start = p->take();
        ev.stop = p->take();
        if (ev.start) (void)hipEventRecord(ev.start, s);
    }
    ~ProfScope() {
        if (!p) return;
        if (ev.stop) (void)hipEventRecord(ev.stop, s);
        p->pending.push_back(ev);
    }
};

}  // namespace z0p
