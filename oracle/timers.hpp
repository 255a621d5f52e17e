// ORACLE — test infrastructure only (see bb.hpp).
//
// Per-stage wall-clock timers for the host-CPU baseline. The reference
// publishes no stage breakdown; these give the first one.
#pragma once
#include <chrono>
#include <map>
#include <string>

namespace oracle {

struct StageTimers {
    std::map<std::string, double> ms;
    void add(const std::string& k, double v) { ms[k] += v; }
    void clear() { ms.clear(); }
};

inline StageTimers& timers() {
    static StageTimers t;
    return t;
}

struct ScopedStage {
    std::string name;
    std::chrono::steady_clock::time_point t0;
    explicit ScopedStage(std::string n) : name(std::move(n)), t0(std::chrono::steady_clock::now()) {}
    ~ScopedStage() {
        auto t1 = std::chrono::steady_clock::now();
        timers().add(name, std::chrono::duration<double, std::milli>(t1 - t0).count());
    }
};

}  // namespace oracle
