// Schedule autotuner (include/z0p.h: z0p_ctx_autotune). Proof bytes do not
// depend on the schedule, so the tuner simply proves the caller's workload
// under candidate schedules and keeps the fastest: coordinate descent over
// the knobs, each candidate timed as the median wall time of `reps` runs of
// the whole workload (one pipelined proof, or one batch).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "params.hpp"
#include "prover.hpp"

namespace z0p {
namespace {

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

double run_once(Ctx& c, const std::vector<z0p_air_params>& work) {
    static const uint8_t label[] = "z0p/autotune";
    std::vector<Transcript> trs(work.size(), Transcript(label, sizeof(label) - 1));
    const double t0 = now_ms();
    if (work.size() == 1) {
        (void)air_prove(c, work[0], trs[0]);
        return now_ms() - t0;
    }
    std::vector<std::unique_ptr<Proof, void (*)(Proof*)>> ps;
    size_t want = 0;
    for (size_t i = 0; i < work.size(); ++i) {
        ps.emplace_back(proof_new_air(c, work[i], &trs[i]), proof_free);
        want += proof_mem_bytes(ps.back().get());
    }
    Engine e(c, c.hbm_budget ? c.hbm_budget : auto_budget(c, want), 256);
    for (auto& p : ps) e.submit(p.get());
    e.run();
    for (auto& p : ps) require(proof_status(p.get()) == Z0P_OK, proof_status(p.get()));
    return now_ms() - t0;
}

double median_ms(Ctx& c, const std::vector<z0p_air_params>& work, unsigned reps) {
    std::vector<double> t;
    for (unsigned r = 0; r < reps; ++r) t.push_back(run_once(c, work));
    std::sort(t.begin(), t.end());
    return t[t.size() / 2];
}

struct Knob {
    const char* name;
    int Schedule::*field;
    std::vector<int> values;
};

}  // namespace

TuneResult autotune(Ctx& c, const std::vector<z0p_air_params>& work, unsigned reps) {
    for (const z0p_air_params& p : work) {
        validate_air(p);
        validate_fri(air_fri_params(p));
    }
    require(!work.empty() && reps > 0, Z0P_ERR_INVALID_ARGUMENT);
    const std::vector<Knob> knobs = {
        {"tail_group_max_log", &Schedule::tail_group_max_log, {5, 6, 7, 8, 9}},
        {"tail_spread_log", &Schedule::tail_spread_log, {0, 6, 7, 8}},
        {"tail_max_log", &Schedule::tail_max_log, {10, 12, 14, 15, 16, 17, 18}},
        {"levels_fuse", &Schedule::levels_fuse, {1, 2}},
        {"small_max_log", &Schedule::small_max_log, {0, 8, 10}},
        {"grind_blocks", &Schedule::grind_blocks, {1024, 2048, 4096, 8192}},
        {"fold_fuse_max_log", &Schedule::fold_fuse_max_log, {0, 12, 14, 16, 18, 22}},
        {"compose_fuse", &Schedule::compose_fuse, {0, 1}},
        {"trace_fuse", &Schedule::trace_fuse, {0, 1}},
        {"subtree_levels", &Schedule::subtree_levels, {0, 2, 4, 6, 8}},
        {"compose_subtree", &Schedule::compose_subtree, {0, 1, 2, 4}},
        {"graph", &Schedule::graph, {0, 1}},
        {"lde_split", &Schedule::lde_split, {0, 1}},
        {"grind_pool", &Schedule::grind_pool, {0, 512, 1024, 2048}},
        {"fold_commit_max_log", &Schedule::fold_commit_max_log, {0, 12, 14, 16, 18, 20, 22}},
        {"tree_top", &Schedule::tree_top, {0, 1}},
        {"wave_streams", &Schedule::wave_streams, {0, 1}},
    };
    // A candidate must beat the incumbent by this fraction to replace it, so
    // run-to-run noise does not move the schedule.
    constexpr double MARGIN = 0.015;
    TuneResult res;
    const Schedule start = c.sched;
    try {
        (void)run_once(c, work);  // warm: geometry tables, pool, code objects
        Schedule best = start;
        double best_ms = median_ms(c, work, reps);
        res.start_ms = best_ms;
        char line[160];
        for (int sweep = 0; sweep < 2; ++sweep) {
            bool moved = false;
            for (const Knob& k : knobs) {
                c.sched = best;
                best_ms = median_ms(c, work, reps);  // incumbent, re-measured against drift
                int keep = best.*k.field;
                for (int v : k.values) {
                    if (v == best.*k.field) continue;
                    Schedule s = best;
                    s.*k.field = v;
                    if (!s.valid()) continue;
                    c.sched = s;
                    const double ms = median_ms(c, work, reps);
                    ++res.trials;
                    std::snprintf(line, sizeof line, "%s=%d:%.4f;", k.name, v, ms);
                    res.log += line;
                    if (ms < best_ms * (1.0 - MARGIN)) {
                        best_ms = ms;
                        keep = v;
                    }
                }
                if (keep != best.*k.field) moved = true;
                best.*k.field = keep;
            }
            if (!moved) break;
        }
        c.sched = best;
        res.best = best;
        res.best_ms = median_ms(c, work, reps);
    } catch (...) {
        c.sched = start;
        throw;
    }
    return res;
}

}  // namespace z0p
