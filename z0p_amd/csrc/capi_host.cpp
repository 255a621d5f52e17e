// extern "C" boundary (include/z0p.h) for the entries that need no device:
// parameter constructors, transcripts, ChaCha20, the verifiers and the
// generic Merkle helpers. It includes no HIP header, so these entries and
// verifier.cpp also build with a plain host compiler (the sanitizer builds).
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <set>
#include <thread>

#include "host.hpp"
#include "params.hpp"

namespace z0p {

namespace {

// (F(k), F(k+1)) canonical by fast doubling.
void fib_pair(uint64_t k, uint32_t& f, uint32_t& g) {
    f = 0;
    g = 1;
    for (int b = 63; b >= 0; --b) {
        uint32_t f2 = mul(f, sub(dbl(g), f));
        uint32_t g2 = add(mul(f, f), mul(g, g));
        if ((k >> b) & 1) {
            f = g2;
            g = add(f2, g2);
        } else {
            f = f2;
            g = g2;
        }
    }
}

}  // namespace

uint32_t fib_b_last(uint32_t a0, uint32_t b0, uint64_t n) {
    // B[n-1] = F(n-1) a0 + F(n) b0
    uint32_t f, g;
    fib_pair(n - 1, f, g);
    return add(mul(f, a0), mul(g, b0));
}

}  // namespace z0p

namespace {
using z0p::capi::guarded;
using z0p::capi::indices_ok;
using z0p::capi::put_indices;
}  // namespace

extern "C" {

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

void z0p_merkle_hash_leaf(const uint8_t* data, size_t len, uint8_t out[32]) {
    if (out && (data || !len)) z0p::merkle_hash_leaf(data, len, out);
}

void z0p_merkle_hash_node(const uint8_t left[32], const uint8_t right[32], uint8_t out[32]) {
    if (left && right && out) z0p::merkle_hash_node(left, right, out);
}

int z0p_merkle_verify(const uint8_t root[32], const uint8_t* leaf, size_t leaf_len, uint64_t index,
                      const uint8_t* path, size_t steps) {
    if (!root || (!leaf && leaf_len) || (!path && steps)) return 0;
    uint8_t d[32];
    z0p::merkle_hash_leaf(leaf, leaf_len, d);
    return z0p::merkle_verify_digest(root, d, index, path, steps) ? 1 : 0;
}

int z0p_merkle_verify_digest(const uint8_t root[32], const uint8_t leaf_digest[32], uint64_t index,
                             const uint8_t* path, size_t steps) {
    if (!root || !leaf_digest || (!path && steps)) return 0;
    return z0p::merkle_verify_digest(root, leaf_digest, index, path, steps) ? 1 : 0;
}

}  // extern "C"
