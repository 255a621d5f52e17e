// extern "C" boundary (include/z0p.h) for the entries that use the device:
// translate statuses, own outputs. The host-only entries are in capi_host.cpp.
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <set>
#include <string>
#include <thread>

#include "params.hpp"
#include "prover.hpp"

namespace {

using z0p::capi::guarded;
using z0p::capi::indices_ok;
using z0p::capi::put_indices;

// Hands the library's buffer to the caller (freed with z0p_free).
int emit(z0p::Bytes& v, uint8_t** out, size_t* len) {
    const size_t n = v.size();
    uint8_t* p = v.release();
    if (!p) return Z0P_ERR_OUT_OF_MEMORY;
    *out = p;
    *len = n;
    return Z0P_OK;
}

}  // namespace

bool z0p::schedule_from_c(const z0p_schedule& in, Schedule& s) {
    s.tail_max_log = in.tail_max_log;
    s.tail_group_max_log = in.tail_group_max_log;
    s.tail_spread_log = in.tail_spread_log;
    s.levels_fuse = in.levels_fuse;
    s.small_max_log = in.small_max_log;
    s.grind_blocks = in.grind_blocks;
    s.fold_fuse_max_log = in.fold_fuse_max_log;
    s.compose_fuse = in.compose_fuse;
    s.trace_fuse = in.trace_fuse;
    s.subtree_levels = in.subtree_levels;
    s.compose_subtree = in.compose_subtree;
    s.graph = in.graph;
    s.lde_split = in.lde_split;
    s.grind_pool = in.grind_pool;
    s.fold_commit_max_log = in.fold_commit_max_log;
    s.tree_top = in.tree_top;
    s.wave_streams = in.wave_streams;
    return s.valid();
}

extern "C" {

int z0p_ctx_create(int device, z0p_ctx** out) {
    if (!out) return Z0P_ERR_INVALID_ARGUMENT;
    return guarded([&] {
        z0p::Ctx* c = z0p::ctx_create(device);
        *out = new z0p_ctx{c};
    });
}

void z0p_ctx_destroy(z0p_ctx* ctx) {
    if (!ctx) return;
    delete ctx->c;
    delete ctx;
}

int z0p_ctx_set_profiling(z0p_ctx* ctx, int enabled) {
    if (!ctx) return Z0P_ERR_INVALID_ARGUMENT;
    ctx->c->prof.enabled = enabled != 0;
    return Z0P_OK;
}

int z0p_ctx_profile_only(z0p_ctx* ctx, const char* kernel) {
    if (!ctx) return Z0P_ERR_INVALID_ARGUMENT;
    ctx->c->prof.only = kernel ? kernel : "";
    return Z0P_OK;
}

size_t z0p_ctx_kernel_stats(z0p_ctx* ctx, char* buf, size_t cap) {
    if (!ctx) return 0;
    // Launches still timing: wait for them, so the figures are complete.
    if (!ctx->c->prof.pending.empty()) {
        (void)hipStreamSynchronize(ctx->c->alt_stream);
        (void)hipStreamSynchronize(ctx->c->stream);
        ctx->c->prof.resolve();
    }
    std::string s;
    for (auto& kv : ctx->c->prof.stats)
        s += kv.first + "=" + std::to_string(kv.second.launches) + "," + std::to_string(kv.second.total_ms) + ";";
    if (buf && cap) {
        size_t n = std::min(cap - 1, s.size());
        std::memcpy(buf, s.data(), n);
        buf[n] = 0;
    }
    return s.size() + 1;
}

void z0p_ctx_reset_stats(z0p_ctx* ctx) {
    if (!ctx) return;
    if (!ctx->c->prof.pending.empty()) {
        (void)hipStreamSynchronize(ctx->c->alt_stream);
        (void)hipStreamSynchronize(ctx->c->stream);
        ctx->c->prof.resolve();
    }
    ctx->c->prof.stats.clear();
}

namespace {
z0p_schedule to_c(const z0p::Schedule& s) {
    return z0p_schedule{s.tail_max_log, s.tail_group_max_log, s.tail_spread_log, s.levels_fuse, s.small_max_log,
                        s.grind_blocks, s.fold_fuse_max_log, s.compose_fuse, s.trace_fuse, s.subtree_levels,
                        s.compose_subtree, s.graph, s.lde_split, s.grind_pool, s.fold_commit_max_log, s.tree_top,
                        s.wave_streams};
}
}  // namespace

int z0p_ctx_get_schedule(z0p_ctx* ctx, z0p_schedule* out) {
    if (!ctx || !out) return Z0P_ERR_INVALID_ARGUMENT;
    *out = to_c(ctx->c->sched);
    return Z0P_OK;
}

int z0p_ctx_set_schedule(z0p_ctx* ctx, const z0p_schedule* in) {
    if (!ctx || !in) return Z0P_ERR_INVALID_ARGUMENT;
    z0p::Schedule s;
    if (!z0p::schedule_from_c(*in, s)) return Z0P_ERR_INVALID_ARGUMENT;
    ctx->c->sched = s;
    return Z0P_OK;
}

int z0p_ctx_autotune(z0p_ctx* ctx, const z0p_air_params* work, size_t count, uint32_t reps, z0p_schedule* out,
                     double* best_ms, double* start_ms) {
    if (!ctx || !work || count == 0 || reps == 0) return Z0P_ERR_INVALID_ARGUMENT;
    return guarded([&] {
        z0p::TuneResult r = z0p::autotune(*ctx->c, std::vector<z0p_air_params>(work, work + count), reps);
        ctx->tune_log = r.log;
        if (out) *out = to_c(r.best);
        if (best_ms) *best_ms = r.best_ms;
        if (start_ms) *start_ms = r.start_ms;
    });
}

size_t z0p_ctx_autotune_log(z0p_ctx* ctx, char* buf, size_t cap) {
    if (!ctx) return 0;
    const std::string& s = ctx->tune_log;
    if (buf && cap) {
        size_t n = std::min(cap - 1, s.size());
        std::memcpy(buf, s.data(), n);
        buf[n] = 0;
    }
    return s.size() + 1;
}

int z0p_transcript_grind(z0p_ctx* ctx, z0p_transcript* t, const char* label, uint32_t difficulty_bits,
                         uint64_t* nonce) {
    if (!ctx || !t || !label || !nonce) return Z0P_ERR_INVALID_ARGUMENT;
    return guarded([&] { *nonce = z0p::transcript_grind(*ctx->c, t->t, label, difficulty_bits); });
}

int z0p_transcript_grind_noalloc(z0p_ctx* ctx, z0p_transcript* t, const char* label, uint32_t difficulty_bits,
                                 uint64_t* nonce) {
    if (!ctx || !t || !label || !nonce) return Z0P_ERR_INVALID_ARGUMENT;
    uint8_t* dev = z0p::noalloc_scratch(ctx->c->device);
    if (!dev) return Z0P_ERR_DEVICE;  // registered when the ctx was created
    return guarded([&] { *nonce = z0p::transcript_grind(*ctx->c, t->t, label, difficulty_bits, dev); });
}

int z0p_air_prove(z0p_ctx* ctx, const z0p_air_params* params, z0p_transcript* transcript, uint8_t** proof,
                  size_t* proof_len) {
    if (!ctx || !params || !transcript || !proof || !proof_len) return Z0P_ERR_INVALID_ARGUMENT;
    int rc = Z0P_OK;
    int g = guarded([&] {
        z0p::Bytes bytes = z0p::air_prove(*ctx->c, *params, transcript->t);
        rc = emit(bytes, proof, proof_len);
    });
    return g != Z0P_OK ? g : rc;
}

int z0p_air_prove_with_trace(z0p_ctx* ctx, const z0p_air_params* params, const uint32_t* col_a, size_t len_a,
                             const uint32_t* col_b, size_t len_b, z0p_transcript* transcript, uint8_t** proof,
                             size_t* proof_len) {
    if (!ctx || !params || !transcript || !proof || !proof_len || (!col_a && len_a) || (!col_b && len_b))
        return Z0P_ERR_INVALID_ARGUMENT;
    int rc = Z0P_OK;
    int g = guarded([&] {
        z0p::Bytes bytes =
            z0p::air_prove_with_trace(*ctx->c, *params, col_a, len_a, col_b, len_b, transcript->t);
        rc = emit(bytes, proof, proof_len);
    });
    return g != Z0P_OK ? g : rc;
}

int z0p_air_prove_zk(z0p_ctx* ctx, const z0p_air_params* params, uint64_t n_rand, z0p_chacha20* rng,
                     z0p_transcript* transcript, uint8_t** proof, size_t* proof_len) {
    if (!ctx || !params || !rng || !transcript || !proof || !proof_len) return Z0P_ERR_INVALID_ARGUMENT;
    int rc = Z0P_OK;
    int g = guarded([&] {
        z0p::Bytes bytes = z0p::air_prove_zk(*ctx->c, *params, n_rand, *rng->r, transcript->t);
        rc = emit(bytes, proof, proof_len);
    });
    return g != Z0P_OK ? g : rc;
}

int z0p_range_prove(z0p_ctx* ctx, const z0p_range_params* params, z0p_transcript* transcript, uint8_t** proof,
                    size_t* proof_len) {
    if (!ctx || !params || !transcript || !proof || !proof_len) return Z0P_ERR_INVALID_ARGUMENT;
    int rc = Z0P_OK;
    int g = guarded([&] {
        z0p::Bytes bytes = z0p::range_prove(*ctx->c, *params, transcript->t);
        rc = emit(bytes, proof, proof_len);
    });
    return g != Z0P_OK ? g : rc;
}

int z0p_range_prove_with_columns(z0p_ctx* ctx, const z0p_range_params* params, const uint32_t* columns, size_t ncols,
                                 size_t len, z0p_transcript* transcript, uint8_t** proof, size_t* proof_len) {
    if (!ctx || !params || !transcript || !proof || !proof_len || (!columns && ncols && len))
        return Z0P_ERR_INVALID_ARGUMENT;
    int rc = Z0P_OK;
    int g = guarded([&] {
        z0p::Bytes bytes = z0p::range_prove_with_columns(*ctx->c, *params, columns, ncols, len, transcript->t);
        rc = emit(bytes, proof, proof_len);
    });
    return g != Z0P_OK ? g : rc;
}

int z0p_range_prove_zk(z0p_ctx* ctx, const z0p_range_params* params, uint64_t n_rand, z0p_chacha20* rng,
                       z0p_transcript* transcript, uint8_t** proof, size_t* proof_len) {
    if (!ctx || !params || !rng || !transcript || !proof || !proof_len) return Z0P_ERR_INVALID_ARGUMENT;
    int rc = Z0P_OK;
    int g = guarded([&] {
        z0p::Bytes bytes = z0p::range_prove_zk(*ctx->c, *params, n_rand, *rng->r, transcript->t);
        rc = emit(bytes, proof, proof_len);
    });
    return g != Z0P_OK ? g : rc;
}

int z0p_table_prove(z0p_ctx* ctx, const z0p_table_params* params, const uint32_t* table, size_t table_len,
                    z0p_transcript* transcript, uint8_t** proof, size_t* proof_len) {
    if (!ctx || !params || !transcript || !proof || !proof_len || (!table && table_len))
        return Z0P_ERR_INVALID_ARGUMENT;
    int rc = Z0P_OK;
    int g = guarded([&] {
        z0p::Bytes bytes = z0p::table_prove(*ctx->c, *params, table, table_len, transcript->t);
        rc = emit(bytes, proof, proof_len);
    });
    return g != Z0P_OK ? g : rc;
}

int z0p_table_prove_zk(z0p_ctx* ctx, const z0p_table_params* params, const uint32_t* table, size_t table_len,
                       uint64_t n_rand, z0p_chacha20* rng, z0p_transcript* transcript, uint8_t** proof,
                       size_t* proof_len) {
    if (!ctx || !params || !rng || !transcript || !proof || !proof_len || (!table && table_len))
        return Z0P_ERR_INVALID_ARGUMENT;
    int rc = Z0P_OK;
    int g = guarded([&] {
        z0p::Bytes bytes = z0p::table_prove_zk(*ctx->c, *params, table, table_len, n_rand, *rng->r, transcript->t);
        rc = emit(bytes, proof, proof_len);
    });
    return g != Z0P_OK ? g : rc;
}

int z0p_fold_acc_new(z0p_ctx* ctx, const z0p_fold_params* params, z0p_fold_acc** out) {
    if (!ctx || !params || !out) return Z0P_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    auto* h = new (std::nothrow) z0p_fold_acc;
    if (!h) return Z0P_ERR_OUT_OF_MEMORY;
    int g = guarded([&] { h->a = z0p::fold_acc_new(*ctx->c, *params); });
    if (g != Z0P_OK) {
        delete h;
        return g;
    }
    *out = h;
    return Z0P_OK;
}

void z0p_fold_acc_free(z0p_fold_acc* acc) {
    if (!acc) return;
    z0p::fold_acc_free(acc->a);
    delete acc;
}

size_t z0p_fold_acc_len(const z0p_fold_acc* acc) { return acc ? z0p::fold_acc_len(acc->a) : 0; }

int z0p_fold_acc_add_segment(z0p_fold_acc* acc, const uint32_t* col_a, size_t len_a, const uint32_t* col_b,
                             size_t len_b, z0p_transcript* transcript) {
    if (!acc || !transcript || (!col_a && len_a) || (!col_b && len_b)) return Z0P_ERR_INVALID_ARGUMENT;
    return guarded([&] { z0p::fold_acc_add_segment(acc->a, col_a, len_a, col_b, len_b, transcript->t); });
}

int z0p_fold_acc_finalize(z0p_fold_acc* acc, z0p_transcript* transcript, uint8_t** proof, size_t* proof_len) {
    if (!acc || !transcript || !proof || !proof_len) return Z0P_ERR_INVALID_ARGUMENT;
    int rc = Z0P_OK;
    int g = guarded([&] {
        z0p::Bytes bytes = z0p::fold_acc_finalize(acc->a, transcript->t);
        rc = emit(bytes, proof, proof_len);
    });
    return g != Z0P_OK ? g : rc;
}

namespace {
int fri_entry(z0p_ctx* ctx, const uint32_t* src, size_t count, int kind, uint64_t degree_bound,
              const z0p_fri_params* params, z0p_transcript* transcript, uint8_t** commitment, size_t* commitment_len,
              uint8_t** proof, size_t* proof_len, uint64_t* indices, size_t indices_cap) {
    if (!ctx || !params || !transcript || !commitment || !commitment_len || !proof || !proof_len ||
        (!src && count) || !indices_ok(indices, indices_cap, params))
        return Z0P_ERR_INVALID_ARGUMENT;
    int rc = Z0P_OK;
    int g = guarded([&] {
        z0p::Bytes c, p;
        std::vector<uint64_t> q;
        z0p::fri_commit_open(*ctx->c, src, count, kind, degree_bound, *params, transcript->t, c, p, q);
        rc = emit(c, commitment, commitment_len);
        if (rc == Z0P_OK) {
            rc = emit(p, proof, proof_len);
            if (rc != Z0P_OK) std::free(*commitment);
        }
        if (rc == Z0P_OK) put_indices(q, indices);
    });
    return g != Z0P_OK ? g : rc;
}

int fri_commit_entry(z0p_ctx* ctx, const uint32_t* src, size_t count, int kind, uint64_t degree_bound,
                     const z0p_fri_params* params, z0p_transcript* transcript, z0p_fri_state** out) {
    if (!ctx || !params || !transcript || !out || (!src && count)) return Z0P_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    auto* h = new (std::nothrow) z0p_fri_state;
    if (!h) return Z0P_ERR_OUT_OF_MEMORY;
    int g = guarded([&] { h->s = z0p::fri_commit(*ctx->c, src, count, kind, degree_bound, *params, transcript->t); });
    if (g != Z0P_OK) {
        delete h;
        return g;
    }
    *out = h;
    return Z0P_OK;
}
}  // namespace

int z0p_fri_commit_open(z0p_ctx* ctx, const uint32_t* coeffs, size_t n, const z0p_fri_params* params,
                        z0p_transcript* transcript, uint8_t** commitment, size_t* commitment_len, uint8_t** proof,
                        size_t* proof_len, uint64_t* indices, size_t indices_cap) {
    return fri_entry(ctx, coeffs, n, z0p::FRI_COEFFS, 0, params, transcript, commitment, commitment_len, proof,
                     proof_len, indices, indices_cap);
}

int z0p_fri_commit_open_ext(z0p_ctx* ctx, const uint32_t* coeffs4, size_t n, const z0p_fri_params* params,
                            z0p_transcript* transcript, uint8_t** commitment, size_t* commitment_len, uint8_t** proof,
                            size_t* proof_len, uint64_t* indices, size_t indices_cap) {
    return fri_entry(ctx, coeffs4, n, z0p::FRI_COEFFS_EXT, 0, params, transcript, commitment, commitment_len, proof,
                     proof_len, indices, indices_cap);
}

int z0p_fri_commit_open_evals(z0p_ctx* ctx, const uint32_t* evals, size_t m, uint64_t degree_bound,
                              const z0p_fri_params* params, z0p_transcript* transcript, uint8_t** commitment,
                              size_t* commitment_len, uint8_t** proof, size_t* proof_len, uint64_t* indices,
                              size_t indices_cap) {
    return fri_entry(ctx, evals, m, z0p::FRI_EVALS, degree_bound, params, transcript, commitment, commitment_len,
                     proof, proof_len, indices, indices_cap);
}

int z0p_fri_commit_open_ext_evals(z0p_ctx* ctx, const uint32_t* evals4, size_t m, uint64_t degree_bound,
                                  const z0p_fri_params* params, z0p_transcript* transcript, uint8_t** commitment,
                                  size_t* commitment_len, uint8_t** proof, size_t* proof_len, uint64_t* indices,
                                  size_t indices_cap) {
    return fri_entry(ctx, evals4, m, z0p::FRI_EVALS_EXT, degree_bound, params, transcript, commitment, commitment_len,
                     proof, proof_len, indices, indices_cap);
}

int z0p_fri_commit(z0p_ctx* ctx, const uint32_t* coeffs, size_t n, const z0p_fri_params* params,
                   z0p_transcript* transcript, z0p_fri_state** out) {
    return fri_commit_entry(ctx, coeffs, n, z0p::FRI_COEFFS, 0, params, transcript, out);
}

int z0p_fri_commit_ext(z0p_ctx* ctx, const uint32_t* coeffs4, size_t n, const z0p_fri_params* params,
                       z0p_transcript* transcript, z0p_fri_state** out) {
    return fri_commit_entry(ctx, coeffs4, n, z0p::FRI_COEFFS_EXT, 0, params, transcript, out);
}

int z0p_fri_commit_evals(z0p_ctx* ctx, const uint32_t* evals, size_t m, uint64_t degree_bound,
                         const z0p_fri_params* params, z0p_transcript* transcript, z0p_fri_state** out) {
    return fri_commit_entry(ctx, evals, m, z0p::FRI_EVALS, degree_bound, params, transcript, out);
}

int z0p_fri_commit_ext_evals(z0p_ctx* ctx, const uint32_t* evals4, size_t m, uint64_t degree_bound,
                             const z0p_fri_params* params, z0p_transcript* transcript, z0p_fri_state** out) {
    return fri_commit_entry(ctx, evals4, m, z0p::FRI_EVALS_EXT, degree_bound, params, transcript, out);
}

int z0p_fri_state_commitment(const z0p_fri_state* state, uint8_t** commitment, size_t* commitment_len) {
    if (!state || !state->s || !commitment || !commitment_len) return Z0P_ERR_INVALID_ARGUMENT;
    const z0p::Bytes& b = z0p::fri_state_commitment(*state->s);
    auto* p = static_cast<uint8_t*>(std::malloc(std::max<size_t>(b.size(), 1)));
    if (!p) return Z0P_ERR_OUT_OF_MEMORY;
    std::memcpy(p, b.data(), b.size());
    *commitment = p;
    *commitment_len = b.size();
    return Z0P_OK;
}

int z0p_fri_open(z0p_fri_state* state, z0p_transcript* transcript, uint8_t** proof, size_t* proof_len,
                 uint64_t* indices, size_t indices_cap) {
    if (!state || !state->s || !transcript || !proof || !proof_len) return Z0P_ERR_INVALID_ARGUMENT;
    if (indices && indices_cap < z0p::fri_state_queries(*state->s)) return Z0P_ERR_INVALID_ARGUMENT;
    int rc = Z0P_OK;
    int g = guarded([&] {
        std::vector<uint64_t> q;
        z0p::Bytes bytes = z0p::fri_open(*state->s, transcript->t, q);
        rc = emit(bytes, proof, proof_len);
        if (rc == Z0P_OK) put_indices(q, indices);
    });
    return g != Z0P_OK ? g : rc;
}

void z0p_fri_state_free(z0p_fri_state* state) {
    if (!state) return;
    z0p::fri_state_free(state->s);
    delete state;
}

int z0p_air_verify_batch_device(z0p_ctx* ctx, const z0p_air_params* params, z0p_transcript* const* transcripts,
                                const uint8_t* const* proofs, const size_t* proof_lens, size_t count,
                                uint32_t threads, int* statuses) {
    if (!ctx) return Z0P_ERR_INVALID_ARGUMENT;
    if (count && (!params || !transcripts || !proofs || !proof_lens || !statuses)) return Z0P_ERR_INVALID_ARGUMENT;
    std::set<const z0p_transcript*> seen;
    for (size_t i = 0; i < count; ++i)
        if (!transcripts[i] || !seen.insert(transcripts[i]).second || (!proofs[i] && proof_lens[i]))
            return Z0P_ERR_INVALID_ARGUMENT;
    std::vector<z0p::Transcript*> trs(count);
    for (size_t i = 0; i < count; ++i) trs[i] = &transcripts[i]->t;
    return guarded([&] {
        z0p::air_verify_batch_device(*ctx->c, params, trs.data(), proofs, proof_lens, count, threads, statuses);
    });
}

}  // extern "C"
