// extern "C" boundary (include/z0p.h): translate statuses, own outputs.
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
    if (ctx) ctx->c->prof.stats.clear();
}

const char* z0p_status_string(int s) {
    switch (s) {
        case Z0P_OK: return "ok";
        case Z0P_ERR_EMPTY_INPUT: return "empty input: need at least one element";
        case Z0P_ERR_INDEX_OUT_OF_BOUNDS: return "index out of bounds";
        case Z0P_ERR_INVALID_DOMAIN_SIZE: return "invalid domain size: must be a power of two";
        case Z0P_ERR_INVALID_BLOWUP: return "invalid blowup: must be a power of two";
        case Z0P_ERR_NO_QUERIES: return "query count must be non-zero";
        case Z0P_ERR_BAD_MERKLE_PATH: return "merkle authentication path failed";
        case Z0P_ERR_FRI_VERIFICATION_FAILED: return "FRI verification failed";
        case Z0P_ERR_MALFORMED_INPUT: return "malformed input";
        case Z0P_ERR_RNG_UNAVAILABLE: return "operating-system RNG unavailable";
        case Z0P_ERR_DEVICE: return "HIP device error";
        case Z0P_ERR_OUT_OF_MEMORY: return "device out of memory";
        case Z0P_ERR_INVALID_ARGUMENT: return "invalid argument";
        case Z0P_ERR_NO_DEVICE: return "no usable GPU device";
        default: return "unknown status";
    }
}

void z0p_free(void* p) { std::free(p); }

namespace {
z0p_schedule to_c(const z0p::Schedule& s) {
    return z0p_schedule{s.tail_max_log, s.tail_group_max_log, s.tail_spread_log, s.levels_fuse, s.small_max_log,
                        s.grind_blocks, s.fold_fuse_max_log, s.compose_fuse, s.trace_fuse, s.subtree_levels,
                        s.compose_subtree};
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
    s.tail_max_log = in->tail_max_log;
    s.tail_group_max_log = in->tail_group_max_log;
    s.tail_spread_log = in->tail_spread_log;
    s.levels_fuse = in->levels_fuse;
    s.small_max_log = in->small_max_log;
    s.grind_blocks = in->grind_blocks;
    s.fold_fuse_max_log = in->fold_fuse_max_log;
    s.compose_fuse = in->compose_fuse;
    s.trace_fuse = in->trace_fuse;
    s.subtree_levels = in->subtree_levels;
    s.compose_subtree = in->compose_subtree;
    if (!s.valid()) return Z0P_ERR_INVALID_ARGUMENT;
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

z0p_transcript* z0p_transcript_new(const uint8_t* label, size_t label_len) {
    auto* t = new (std::nothrow) z0p_transcript;
    if (t) t->t = z0p::Transcript(label, label_len);
    return t;
}

z0p_transcript* z0p_transcript_clone(const z0p_transcript* t) {
    if (!t) return nullptr;
    return new (std::nothrow) z0p_transcript(*t);
}

void z0p_transcript_free(z0p_transcript* t) { delete t; }

void z0p_transcript_absorb(z0p_transcript* t, const char* label, const uint8_t* data, size_t len) {
    if (t && label) t->t.absorb(label, data, len);
}

void z0p_transcript_challenge_bytes(z0p_transcript* t, const char* label, uint8_t out[32]) {
    if (t && label && out) t->t.squeeze(label, out);
}

uint32_t z0p_transcript_challenge_bb(z0p_transcript* t, const char* label) {
    if (!t || !label) return 0;
    uint8_t d[32];
    t->t.squeeze(label, d);
    return uint32_t(z0p::load_le64(d) % z0p::P);
}

void z0p_transcript_challenge_ext(z0p_transcript* t, const char* label, uint32_t out[4]) {
    if (!t || !label || !out) return;
    const z0p::E4 e = t->t.challenge_ext(label);
    for (int i = 0; i < 4; ++i) out[i] = e.v[i];
}

int z0p_transcript_challenge_index(z0p_transcript* t, const char* label, uint64_t domain_size, uint64_t* out) {
    if (!t || !label || !out) return Z0P_ERR_INVALID_ARGUMENT;
    if (!z0p::pow2(domain_size)) return Z0P_ERR_INVALID_DOMAIN_SIZE;
    *out = t->t.challenge_index(label, domain_size);
    return Z0P_OK;
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

int z0p_transcript_verify_grind(z0p_transcript* t, const char* label, uint64_t nonce, uint32_t difficulty_bits) {
    if (!t || !label || difficulty_bits > 256) return 0;
    const std::vector<uint8_t> head = t->t.pow_head(label);
    uint8_t d[32];
    z0p::pow_digest(head, nonce, d);
    if (z0p::leading_zero_bits(d) < difficulty_bits) return 0;
    uint8_t nb[8];
    z0p::store_le64(nb, nonce);
    t->t.absorb(label, nb, 8);
    return 1;
}

int z0p_air_params_new(uint64_t n_steps, uint64_t blowup, uint64_t num_queries, uint64_t final_size, uint32_t a0,
                       uint32_t b0, uint32_t b_last, z0p_air_params* out) {
    if (!out) return Z0P_ERR_INVALID_ARGUMENT;
    using namespace z0p;
    if (n_steps < 2 || !pow2(n_steps)) return Z0P_ERR_INVALID_DOMAIN_SIZE;
    if (blowup < 4 || !pow2(blowup)) return Z0P_ERR_INVALID_BLOWUP;
    if (num_queries == 0) return Z0P_ERR_NO_QUERIES;
    if (final_size < 2 || !pow2(final_size)) return Z0P_ERR_INVALID_DOMAIN_SIZE;
    if (n_steps > MAX_DOMAIN || blowup > MAX_DOMAIN || n_steps * blowup > MAX_DOMAIN) return Z0P_ERR_INVALID_DOMAIN_SIZE;
    std::memset(out, 0, sizeof(*out));
    out->n_steps = n_steps;
    out->blowup = blowup;
    out->num_queries = num_queries;
    out->final_size = final_size;
    out->a0 = a0 % P;
    out->b0 = b0 % P;
    out->b_last = b_last % P;
    return Z0P_OK;
}

int z0p_air_params_for_security(uint64_t n_steps, uint32_t a0, uint32_t b0, uint32_t b_last, uint32_t target_bits,
                                z0p_air_params* out) {
    if (!out) return Z0P_ERR_INVALID_ARGUMENT;
    using namespace z0p;
    if (n_steps < 2 || !pow2(n_steps)) return Z0P_ERR_INVALID_DOMAIN_SIZE;
    const uint64_t blowup = 16;
    if (n_steps > MAX_DOMAIN / blowup) return Z0P_ERR_INVALID_DOMAIN_SIZE;
    const uint64_t lde = n_steps * blowup;
    auto strength = tune_strength(target_bits, composition_millibits(lde, n_steps - 1));
    auto shape = tune_shape(lde, strength.first, true, lde, 24);
    std::memset(out, 0, sizeof(*out));
    out->n_steps = n_steps;
    out->blowup = blowup;
    out->num_queries = strength.first;
    out->final_size = shape.first;
    out->grinding_bits = strength.second;
    out->cap_height = shape.second;
    out->a0 = a0 % P;
    out->b0 = b0 % P;
    out->b_last = b_last % P;
    return Z0P_OK;
}

int z0p_fri_params_new(uint64_t blowup, uint64_t num_queries, uint64_t final_size, z0p_fri_params* out) {
    if (!out) return Z0P_ERR_INVALID_ARGUMENT;
    using namespace z0p;
    if (blowup < 2 || !pow2(blowup)) return Z0P_ERR_INVALID_BLOWUP;
    if (num_queries == 0) return Z0P_ERR_NO_QUERIES;
    if (final_size < 2 || !pow2(final_size) || final_size > MAX_FINAL_SIZE) return Z0P_ERR_INVALID_DOMAIN_SIZE;
    std::memset(out, 0, sizeof(*out));
    out->blowup = blowup;
    out->num_queries = num_queries;
    out->final_size = final_size;
    return Z0P_OK;
}

int z0p_fri_params_for_security(uint64_t poly_len, uint32_t target_bits, z0p_fri_params* out) {
    if (!out) return Z0P_ERR_INVALID_ARGUMENT;
    using namespace z0p;
    if (poly_len == 0) return Z0P_ERR_EMPTY_INPUT;
    if (!pow2(poly_len)) return Z0P_ERR_INVALID_DOMAIN_SIZE;
    const uint64_t blowup = 16;
    if (poly_len > MAX_DOMAIN / blowup) return Z0P_ERR_INVALID_DOMAIN_SIZE;
    const uint64_t lde = poly_len * blowup;
    auto strength = tune_strength(target_bits, 2000);
    auto shape = tune_shape(lde, strength.first, false, 0, 0);
    std::memset(out, 0, sizeof(*out));
    out->blowup = blowup;
    out->num_queries = strength.first;
    out->final_size = shape.first;
    out->grinding_bits = strength.second;
    out->cap_height = shape.second;
    return Z0P_OK;
}

uint32_t z0p_air_b_last(uint32_t a0, uint32_t b0, uint64_t n_steps) {
    if (n_steps == 0) return 0;
    return z0p::fib_b_last(a0 % z0p::P, b0 % z0p::P, n_steps);
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

int z0p_range_params_new(uint64_t n_rows, uint64_t n_bits, uint64_t blowup, uint64_t num_queries,
                         uint64_t final_size, uint32_t value, z0p_range_params* out) {
    if (!out) return Z0P_ERR_INVALID_ARGUMENT;
    using namespace z0p;
    if (n_rows < 2 || !pow2(n_rows)) return Z0P_ERR_INVALID_DOMAIN_SIZE;
    if (n_bits == 0 || n_bits > 24) return Z0P_ERR_MALFORMED_INPUT;
    if (blowup < 2 || !pow2(blowup)) return Z0P_ERR_INVALID_BLOWUP;
    if (num_queries == 0) return Z0P_ERR_NO_QUERIES;
    if (final_size < 2 || !pow2(final_size)) return Z0P_ERR_INVALID_DOMAIN_SIZE;
    if (n_rows > MAX_DOMAIN || blowup > MAX_DOMAIN || n_rows * blowup > MAX_DOMAIN) return Z0P_ERR_INVALID_DOMAIN_SIZE;
    std::memset(out, 0, sizeof(*out));
    out->n_rows = n_rows;
    out->n_bits = n_bits;
    out->blowup = blowup;
    out->num_queries = num_queries;
    out->final_size = final_size;
    out->value = value % P;
    return Z0P_OK;
}

int z0p_range_params_for_security(uint64_t n_rows, uint64_t n_bits, uint32_t value, uint32_t target_bits,
                                  z0p_range_params* out) {
    if (!out) return Z0P_ERR_INVALID_ARGUMENT;
    using namespace z0p;
    if (n_rows < 2 || !pow2(n_rows)) return Z0P_ERR_INVALID_DOMAIN_SIZE;
    if (n_bits == 0 || n_bits > 24) return Z0P_ERR_MALFORMED_INPUT;
    const uint64_t blowup = 16;
    if (n_rows > MAX_DOMAIN / blowup) return Z0P_ERR_INVALID_DOMAIN_SIZE;
    const uint64_t lde = n_rows * blowup;
    auto strength = tune_strength(target_bits, composition_millibits(lde, n_rows - 1));
    // one column opening: 8 index bytes plus 4 per column
    auto shape = tune_shape(lde, strength.first, true, lde, 8 + 4 * (n_bits + 1));
    std::memset(out, 0, sizeof(*out));
    out->n_rows = n_rows;
    out->n_bits = n_bits;
    out->blowup = blowup;
    out->num_queries = strength.first;
    out->final_size = shape.first;
    out->grinding_bits = strength.second;
    out->cap_height = shape.second;
    out->value = value % P;
    return Z0P_OK;
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

int z0p_range_verify(const z0p_range_params* params, z0p_transcript* transcript, const uint8_t* proof,
                     size_t proof_len) {
    if (!params || !transcript || (!proof && proof_len)) return Z0P_ERR_INVALID_ARGUMENT;
    return guarded([&] { z0p::range_verify(*params, transcript->t, proof, proof_len); });
}

int z0p_table_params_new(uint64_t n_rows, uint64_t blowup, uint64_t num_queries, uint64_t final_size, uint32_t value,
                         z0p_table_params* out) {
    if (!out) return Z0P_ERR_INVALID_ARGUMENT;
    using namespace z0p;
    if (n_rows < 2 || !pow2(n_rows)) return Z0P_ERR_INVALID_DOMAIN_SIZE;
    if (blowup < 2 || !pow2(blowup)) return Z0P_ERR_INVALID_BLOWUP;
    if (num_queries == 0) return Z0P_ERR_NO_QUERIES;
    if (final_size < 2 || !pow2(final_size)) return Z0P_ERR_INVALID_DOMAIN_SIZE;
    if (n_rows > MAX_DOMAIN || blowup > MAX_DOMAIN || n_rows * blowup > MAX_DOMAIN) return Z0P_ERR_INVALID_DOMAIN_SIZE;
    std::memset(out, 0, sizeof(*out));
    out->n_rows = n_rows;
    out->blowup = blowup;
    out->num_queries = num_queries;
    out->final_size = final_size;
    out->value = value % P;
    return Z0P_OK;
}

int z0p_table_params_for_security(uint64_t n_rows, uint64_t table_len, uint32_t value, uint32_t target_bits,
                                  z0p_table_params* out) {
    if (!out) return Z0P_ERR_INVALID_ARGUMENT;
    using namespace z0p;
    if (n_rows < 2 || !pow2(n_rows)) return Z0P_ERR_INVALID_DOMAIN_SIZE;
    if (table_len < 2) return Z0P_ERR_MALFORMED_INPUT;
    const unsigned __int128 prod = (unsigned __int128)table_len * (n_rows - 1);
    if (prod > UINT64_MAX || prod < n_rows) return Z0P_ERR_MALFORMED_INPUT;
    const uint64_t bound = uint64_t(prod) - n_rows;
    const uint64_t need = bound == UINT64_MAX ? bound : bound + 1;
    // the smallest blowup from 4 with need <= lde / 4
    uint64_t blowup = 4, lde = 0;
    while (true) {
        if (n_rows > UINT64_MAX / blowup) return Z0P_ERR_INVALID_DOMAIN_SIZE;
        lde = n_rows * blowup;
        if (need <= lde / 4 || blowup >= (uint64_t(1) << 27)) break;
        blowup *= 2;
    }
    if (lde > MAX_DOMAIN || need > lde / 4) return Z0P_ERR_INVALID_DOMAIN_SIZE;
    auto strength = tune_strength(target_bits, composition_millibits(lde, bound));
    auto shape = tune_shape(lde, strength.first, true, lde, 12);  // 8 index bytes + one field
    std::memset(out, 0, sizeof(*out));
    out->n_rows = n_rows;
    out->blowup = blowup;
    out->num_queries = strength.first;
    out->final_size = shape.first;
    out->grinding_bits = strength.second;
    out->cap_height = shape.second;
    out->value = value % P;
    return Z0P_OK;
}

uint64_t z0p_table_membership_bound(uint64_t table_len, uint64_t n_rows, uint64_t n_rand) {
    return z0p::membership_bound(table_len, n_rows, n_rand);
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

int z0p_table_verify(const z0p_table_params* params, const uint32_t* table, size_t table_len,
                     z0p_transcript* transcript, const uint8_t* proof, size_t proof_len) {
    if (!params || !transcript || (!proof && proof_len) || (!table && table_len)) return Z0P_ERR_INVALID_ARGUMENT;
    return guarded([&] { z0p::table_verify(*params, table, table_len, transcript->t, proof, proof_len); });
}

int z0p_fold_params_new(uint64_t seg_steps, uint64_t blowup, uint64_t num_queries, uint64_t final_size,
                        uint32_t in_a, uint32_t in_b, uint32_t out_b, z0p_fold_params* out) {
    if (!out) return Z0P_ERR_INVALID_ARGUMENT;
    using namespace z0p;
    if (seg_steps < 2 || !pow2(seg_steps)) return Z0P_ERR_INVALID_DOMAIN_SIZE;
    if (blowup < 4 || !pow2(blowup)) return Z0P_ERR_INVALID_BLOWUP;
    if (num_queries == 0) return Z0P_ERR_NO_QUERIES;
    if (final_size < 2 || !pow2(final_size)) return Z0P_ERR_INVALID_DOMAIN_SIZE;
    if (seg_steps > MAX_DOMAIN || blowup > MAX_DOMAIN || seg_steps * blowup > MAX_DOMAIN)
        return Z0P_ERR_INVALID_DOMAIN_SIZE;
    std::memset(out, 0, sizeof(*out));
    out->seg_steps = seg_steps;
    out->blowup = blowup;
    out->num_queries = num_queries;
    out->final_size = final_size;
    out->in_a = in_a % P;
    out->in_b = in_b % P;
    out->out_b = out_b % P;
    return Z0P_OK;
}

int z0p_fold_params_for_security(uint64_t seg_steps, uint32_t in_a, uint32_t in_b, uint32_t out_b,
                                 uint32_t target_bits, z0p_fold_params* out) {
    if (!out) return Z0P_ERR_INVALID_ARGUMENT;
    using namespace z0p;
    if (seg_steps < 2 || !pow2(seg_steps)) return Z0P_ERR_INVALID_DOMAIN_SIZE;
    const uint64_t blowup = 16;
    if (seg_steps > MAX_DOMAIN / blowup) return Z0P_ERR_INVALID_DOMAIN_SIZE;
    const uint64_t lde = seg_steps * blowup;
    auto strength = tune_strength(target_bits, composition_millibits(lde, seg_steps - 1));
    auto shape = tune_shape(lde, strength.first, true, lde, 24);  // 8 index bytes + 4 fields
    std::memset(out, 0, sizeof(*out));
    out->seg_steps = seg_steps;
    out->blowup = blowup;
    out->num_queries = strength.first;
    out->final_size = shape.first;
    out->grinding_bits = strength.second;
    out->cap_height = shape.second;
    out->in_a = in_a % P;
    out->in_b = in_b % P;
    out->out_b = out_b % P;
    return Z0P_OK;
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

int z0p_fold_verify(const z0p_fold_params* params, z0p_transcript* transcript, const uint8_t* proof,
                    size_t proof_len) {
    if (!params || !transcript || (!proof && proof_len)) return Z0P_ERR_INVALID_ARGUMENT;
    return guarded([&] { z0p::fold_verify(*params, transcript->t, proof, proof_len); });
}

z0p_chacha20* z0p_chacha20_new(const uint8_t key[32], const uint8_t nonce[12]) {
    if (!key || !nonce) return nullptr;
    auto* h = new (std::nothrow) z0p_chacha20;
    if (!h) return nullptr;
    h->r = new (std::nothrow) z0p::ChaCha20(key, nonce);
    if (!h->r) {
        delete h;
        return nullptr;
    }
    return h;
}

int z0p_chacha20_from_os(z0p_chacha20** out) {
    if (!out) return Z0P_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    return guarded([&] {
        auto h = std::make_unique<z0p_chacha20>();
        h->r = z0p::ChaCha20::from_os();
        *out = h.release();
    });
}

void z0p_chacha20_free(z0p_chacha20* rng) {
    if (!rng) return;
    delete rng->r;
    delete rng;
}

void z0p_chacha20_fill_bytes(z0p_chacha20* rng, uint8_t* out, size_t len) {
    if (rng && (out || !len)) rng->r->fill(out, len);
}
uint32_t z0p_chacha20_next_u32(z0p_chacha20* rng) { return rng ? rng->r->next_u32() : 0; }
uint64_t z0p_chacha20_next_u64(z0p_chacha20* rng) { return rng ? rng->r->next_u64() : 0; }
uint32_t z0p_chacha20_next_field(z0p_chacha20* rng) { return rng ? rng->r->next_field() : 0; }

int z0p_air_verify(const z0p_air_params* params, z0p_transcript* transcript, const uint8_t* proof,
                   size_t proof_len) {
    if (!params || !transcript || (!proof && proof_len)) return Z0P_ERR_INVALID_ARGUMENT;
    return guarded([&] { z0p::air_verify(*params, transcript->t, proof, proof_len); });
}

int z0p_air_verify_batch(const z0p_air_params* params, z0p_transcript* const* transcripts,
                         const uint8_t* const* proofs, const size_t* proof_lens, size_t count, uint32_t threads,
                         int* statuses) {
    if (count && (!params || !transcripts || !proofs || !proof_lens || !statuses)) return Z0P_ERR_INVALID_ARGUMENT;
    std::set<const z0p_transcript*> seen;
    for (size_t i = 0; i < count; ++i)
        if (!transcripts[i] || !seen.insert(transcripts[i]).second || (!proofs[i] && proof_lens[i]))
            return Z0P_ERR_INVALID_ARGUMENT;
    size_t workers = threads ? threads : std::max(1u, std::thread::hardware_concurrency());
    workers = std::min(workers, count);
    std::atomic<size_t> next{0};
    auto work = [&] {
        for (size_t i = next.fetch_add(1); i < count; i = next.fetch_add(1))
            statuses[i] = guarded([&] { z0p::air_verify(params[i], transcripts[i]->t, proofs[i], proof_lens[i]); });
    };
    std::vector<std::thread> pool;
    for (size_t w = 1; w < workers; ++w) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    return Z0P_OK;
}

namespace {
bool indices_ok(const uint64_t* indices, size_t cap, const z0p_fri_params* params) {
    return !indices || cap >= params->num_queries;
}

void put_indices(const std::vector<uint64_t>& q, uint64_t* indices) {
    if (indices) std::copy(q.begin(), q.end(), indices);
}

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

int z0p_fri_verify(const uint8_t* commitment, size_t commitment_len, const uint8_t* proof, size_t proof_len,
                   const z0p_fri_params* params, z0p_transcript* transcript, uint64_t* indices, size_t indices_cap) {
    if (!params || !transcript || (!commitment && commitment_len) || (!proof && proof_len) ||
        !indices_ok(indices, indices_cap, params))
        return Z0P_ERR_INVALID_ARGUMENT;
    return guarded([&] {
        std::vector<uint64_t> q = z0p::fri_verify(commitment, commitment_len, proof, proof_len, *params, transcript->t);
        put_indices(q, indices);
    });
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

}  // extern "C"
