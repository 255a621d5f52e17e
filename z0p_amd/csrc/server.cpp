// Batched and served proving on one device (include/z0p.h: z0p_air_prove_batch,
// z0p_server_*). Both feed the proof scheduler (prover.cpp: Engine): proofs
// share every kernel launch and are admitted as pool memory frees up. The
// server runs two schedulers (small and large requests), each on its own host
// thread and stream, and admits requests that arrive while others are on the
// device into the running batch of their lane (continuous batching); a
// request's proof bytes do not depend on what it ran beside.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <thread>

#include "params.hpp"
#include "prover.hpp"

namespace {

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return Z0P_OK;
    } catch (const z0p::Status& s) {
        return s.code;
    } catch (const std::bad_alloc&) {
        return Z0P_ERR_OUT_OF_MEMORY;
    } catch (...) {
        return Z0P_ERR_DEVICE;
    }
}

// The proof's buffer, handed to the caller (freed with z0p_free).
uint8_t* hand_over(z0p::Bytes& v, size_t* len) {
    *len = v.size();
    return v.release();
}

constexpr size_t DEFAULT_MAX_ACTIVE = 256;

// Same clock as the scheduler's proof timestamps.
double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace

using z0p::Proof;

struct ServerRequest {
    uint64_t id = 0;
    z0p::Transcript tr;
    z0p_air_params params{};
    Proof* proof = nullptr;
    int status = Z0P_OK;
    double submit_ms = 0, admit_ms = 0, done_ms = 0;
    ~ServerRequest() {
        if (proof) z0p::proof_free(proof);
    }
};

// Requests of at least 2^LARGE_LOG steps go to the large-proof lane, the rest
// to the small-proof lane. Each lane is one scheduler on its own stream (and
// its own context), run by its own host thread: short proofs keep cycling
// through their phases while long ones occupy the other lane, and the two
// streams' kernels share the device.
constexpr uint64_t LARGE_LOG = 16;
constexpr int LANES = 2;

struct z0p_server;

struct Lane {
    z0p_server* s = nullptr;
    z0p::Ctx* c = nullptr;
    bool owned = false;  // created by the server (the large lane)
    size_t budget = 0;
    std::unique_ptr<z0p::Engine> engine;
    std::thread worker;
    std::condition_variable cv_work;
    std::deque<ServerRequest*> incoming;
    std::map<uint64_t, ServerRequest*> running;
    void loop();
};

struct z0p_server {
    size_t max_active = DEFAULT_MAX_ACTIVE;
    Lane lanes[LANES];
    std::mutex mu;
    std::condition_variable cv_done;
    bool stop = false;
    uint64_t next_id = 1;
    size_t pending = 0;  // submitted and not yet polled
    std::deque<ServerRequest*> finished;

    void complete(ServerRequest* r) {
        if (r->done_ms == 0) r->done_ms = r->admit_ms = now_ms();
        std::lock_guard<std::mutex> lk(mu);
        finished.push_back(r);
        cv_done.notify_all();
    }
    ~z0p_server() {
        for (Lane& l : lanes) {
            l.engine.reset();
            if (l.owned) delete l.c;
        }
    }
};

void Lane::loop() {
    (void)hipSetDevice(c->device);
    z0p::Engine* e = engine.get();
    auto on_done = [this](Proof* p) {
        ServerRequest* r = nullptr;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            auto it = running.find(z0p::proof_tag(p));
            if (it == running.end()) return;
            r = it->second;
            running.erase(it);
        }
        r->status = z0p::proof_status(p);
        double engine_submit = 0;
        z0p::proof_times(p, engine_submit, r->admit_ms, r->done_ms);
        if (r->admit_ms == 0) r->admit_ms = r->done_ms;  // never admitted (over budget)
        s->complete(r);
    };
    e->on_done(on_done);
    while (true) {
        std::deque<ServerRequest*> arrived;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            cv_work.wait(lk, [&] { return s->stop || !incoming.empty() || e->active() + e->queued() > 0; });
            if (s->stop && incoming.empty() && e->active() + e->queued() == 0) break;
            arrived.swap(incoming);
        }
        for (ServerRequest* r : arrived) {
            try {
                r->proof = z0p::proof_new_air(*c, r->params, &r->tr);
            } catch (const z0p::Status& st) {
                r->status = st.code;
                s->complete(r);
                continue;
            } catch (...) {
                r->status = Z0P_ERR_DEVICE;
                s->complete(r);
                continue;
            }
            z0p::proof_set_tag(r->proof, r->id);
            {
                std::lock_guard<std::mutex> lk(s->mu);
                running[r->id] = r;
            }
            e->submit(r->proof);
        }
        // A device failure ends the proofs on the device with its status;
        // the scheduler is left empty and later requests run as usual.
        try {
            e->iterate();
        } catch (const z0p::Status& st) {
            e->fail_all(st.code);
        } catch (...) {
            e->fail_all(Z0P_ERR_DEVICE);
        }
    }
}

extern "C" {

int z0p_ctx_set_hbm_budget(z0p_ctx* ctx, uint64_t bytes) {
    if (!ctx) return Z0P_ERR_INVALID_ARGUMENT;
    ctx->c->hbm_budget = size_t(bytes);
    return Z0P_OK;
}

int z0p_ctx_set_waves(z0p_ctx* ctx, uint32_t min_active) {
    if (!ctx) return Z0P_ERR_INVALID_ARGUMENT;
    ctx->c->waves_min = min_active;
    return Z0P_OK;
}

int z0p_air_prove_batch(z0p_ctx* ctx, const z0p_air_params* params, z0p_transcript* const* transcripts,
                        size_t count, uint8_t** proofs, size_t* proof_lens, int* statuses) {
    if (!ctx || (count && (!params || !transcripts || !proofs || !proof_lens || !statuses)))
        return Z0P_ERR_INVALID_ARGUMENT;
    std::set<const z0p_transcript*> seen;
    for (size_t i = 0; i < count; ++i) {
        // Each instance advances its own transcript: no null and no sharing.
        if (!transcripts[i] || !seen.insert(transcripts[i]).second) return Z0P_ERR_INVALID_ARGUMENT;
        proofs[i] = nullptr;
        proof_lens[i] = 0;
        statuses[i] = Z0P_OK;
    }
    z0p::Ctx& c = *ctx->c;
    std::vector<std::unique_ptr<Proof, void (*)(Proof*)>> ps;
    ps.reserve(count);
    return guarded([&] {
        size_t want = 0;
        for (size_t i = 0; i < count; ++i) {
            Proof* p = nullptr;
            try {
                p = z0p::proof_new_air(c, params[i], &transcripts[i]->t);
            } catch (const z0p::Status& s) {
                statuses[i] = s.code;
            }
            ps.emplace_back(p, z0p::proof_free);
            if (p) want += z0p::proof_mem_bytes(p);
        }
        const size_t budget = c.hbm_budget ? c.hbm_budget : z0p::auto_budget(c, want);
        z0p::Engine e(c, budget, DEFAULT_MAX_ACTIVE);
        for (auto& p : ps)
            if (p) e.submit(p.get());
        // A device error that ends the batch has already given every
        // unfinished instance its status; instances that finished before it
        // still hand over their proofs, so statuses[i] == 0 always comes
        // with proofs[i].
        int ended = Z0P_OK;
        try {
            e.run();
        } catch (const z0p::Status& s) {
            ended = s.code;
        } catch (...) {
            ended = Z0P_ERR_DEVICE;
        }
        for (size_t i = 0; i < count; ++i) {
            if (!ps[i]) continue;
            statuses[i] = z0p::proof_status(ps[i].get());
            if (statuses[i] == Z0P_OK && z0p::proof_done(ps[i].get())) {
                proofs[i] = hand_over(z0p::proof_bytes(ps[i].get()), &proof_lens[i]);
                if (!proofs[i]) {
                    statuses[i] = Z0P_ERR_OUT_OF_MEMORY;
                    proof_lens[i] = 0;
                }
            } else if (statuses[i] == Z0P_OK) {
                statuses[i] = ended;
            }
        }
        if (ended != Z0P_OK) throw z0p::Status(ended);
    });
}

int z0p_server_create(z0p_ctx* ctx, uint64_t hbm_budget, uint32_t max_active, z0p_server** out) {
    return z0p_server_create_lanes(ctx, nullptr, hbm_budget, max_active, out);
}

int z0p_server_create_lanes(z0p_ctx* ctx, const z0p_schedule* large_schedule, uint64_t hbm_budget,
                            uint32_t max_active, z0p_server** out) {
    if (!ctx || !out) return Z0P_ERR_INVALID_ARGUMENT;
    z0p::Schedule large_sched = ctx->c->sched;
    if (large_schedule && !z0p::schedule_from_c(*large_schedule, large_sched)) return Z0P_ERR_INVALID_ARGUMENT;
    return guarded([&] {
        auto s = std::make_unique<z0p_server>();
        z0p::Ctx* small_ctx = ctx->c;
        if (hbm_budget == 0) {
            size_t free_b = 0, total_b = 0;
            z0p::hip_check(hipMemGetInfo(&free_b, &total_b));
            if (small_ctx->pool.in_use() == 0) free_b += small_ctx->pool.capacity();
            hbm_budget = free_b / 2;
        }
        s->max_active = max_active ? max_active : DEFAULT_MAX_ACTIVE;
        // The large lane gets its own context (stream, pool, tables) with the
        // caller's wave setting and kernel schedule, or the schedule given for
        // it (set before its scheduler thread starts). The small lane's share
        // of the budget is half of it, at most 1 GiB (a 2^15-step proof needs
        // ~32 MiB).
        Lane& small = s->lanes[0];
        Lane& large = s->lanes[1];
        small.c = small_ctx;
        large.c = z0p::ctx_create(small_ctx->device);
        large.owned = true;
        large.c->sched = large_sched;
        large.c->waves_min = small_ctx->waves_min;
        small.budget = std::min<size_t>(size_t(hbm_budget) / 2, size_t(1) << 30);
        large.budget = size_t(hbm_budget) - small.budget;
        for (Lane& l : s->lanes) {
            l.s = s.get();
            l.engine.reset(new z0p::Engine(*l.c, l.budget, s->max_active));
            // A request alone in its lane runs all its phases back to back.
            l.engine->set_pipelined(true);
        }
        z0p_server* raw = s.get();
        for (Lane& l : raw->lanes) {
            Lane* lp = &l;
            l.worker = std::thread([lp] { lp->loop(); });
        }
        *out = s.release();
    });
}

int z0p_server_submit_air(z0p_server* s, const z0p_air_params* params, const z0p_transcript* transcript,
                          uint64_t* request_id) {
    if (!s || !params || !transcript || !request_id) return Z0P_ERR_INVALID_ARGUMENT;
    return guarded([&] {
        // Parameter errors surface at submission, as from z0p_air_prove.
        z0p::validate_air(*params);
        z0p::validate_fri(z0p::air_fri_params(*params));
        auto r = std::make_unique<ServerRequest>();
        r->tr = transcript->t;
        r->params = *params;
        r->submit_ms = now_ms();
        Lane& lane = s->lanes[params->n_steps >= (uint64_t(1) << LARGE_LOG) ? 1 : 0];
        std::lock_guard<std::mutex> lk(s->mu);
        r->id = s->next_id++;
        *request_id = r->id;
        lane.incoming.push_back(r.release());
        ++s->pending;
        lane.cv_work.notify_one();
    });
}

int z0p_server_poll(z0p_server* s, int timeout_ms, uint64_t* request_id, int* status, uint8_t** proof,
                    size_t* proof_len, double* latency_ms, double* queue_ms) {
    if (!s || !request_id || !status || !proof || !proof_len) return Z0P_ERR_INVALID_ARGUMENT;
    std::unique_ptr<ServerRequest> r;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto ready = [&] { return !s->finished.empty(); };
        if (timeout_ms < 0) s->cv_done.wait(lk, ready);
        else if (timeout_ms > 0) s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
        if (s->finished.empty()) return Z0P_ERR_EMPTY_INPUT;
        r.reset(s->finished.front());
        s->finished.pop_front();
        --s->pending;
    }
    *request_id = r->id;
    *status = r->status;
    *proof = nullptr;
    *proof_len = 0;
    if (latency_ms) *latency_ms = r->done_ms - r->submit_ms;
    if (queue_ms) *queue_ms = r->admit_ms - r->submit_ms;
    if (r->status == Z0P_OK && r->proof) {
        *proof = hand_over(z0p::proof_bytes(r->proof), proof_len);
        if (!*proof) {
            *status = Z0P_ERR_OUT_OF_MEMORY;
            *proof_len = 0;
            return Z0P_OK;
        }
    }
    return Z0P_OK;
}

size_t z0p_server_pending(z0p_server* s) {
    if (!s) return 0;
    std::lock_guard<std::mutex> lk(s->mu);
    return s->pending;
}

void z0p_server_destroy(z0p_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        for (Lane& l : s->lanes) l.cv_work.notify_all();
    }
    for (Lane& l : s->lanes)
        if (l.worker.joinable()) l.worker.join();
    for (ServerRequest* r : s->finished) delete r;
    for (Lane& l : s->lanes)
        for (ServerRequest* r : l.incoming) delete r;
    delete s;
}

}  // extern "C"
