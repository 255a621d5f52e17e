/*
 * z0p-mi355x: C ABI of the MI355X (gfx950) STARK prover for the z0p
 * Fibonacci AIR and FRI polynomial commitment.
 *
 * Drop-in boundary for the reference crate's public proving API
 * (FutureProofsTech/z0p v0.3.0). Each entry point names the reference item it
 * replaces. Proofs and commitments are the reference's canonical VERSION = 6
 * wire bytes (src/codec.rs:21), byte-identical to the reference CPU prover's
 * output for the same parameters and transcript.
 *
 * Conventions
 *  - Status: 0 = Ok; 1..9 = the reference `Error` variants in declaration
 *    order (src/error.rs:9-39); >= 100 = device/library errors.
 *  - Ownership: inputs are borrowed for the call. Byte outputs are allocated
 *    by the library and released with z0p_free.
 *  - Transcripts are advanced exactly as the reference advances its
 *    `&mut Transcript` (src/transcript.rs:37-40): callers may pre-absorb and
 *    keep using the transcript after a call.
 *  - A z0p_ctx owns one HIP stream on one device plus its device memory.
 *    Calls on one ctx are synchronous and must not overlap; use one ctx per
 *    host thread. A ctx held by a z0p_server is used by the server only.
 *  - Field elements cross the boundary as canonical u32 (< p = 2013265921).
 */
#ifndef Z0P_H
#define Z0P_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    Z0P_OK = 0,
    Z0P_ERR_EMPTY_INPUT = 1,               /* Error::EmptyInput            error.rs:9  */
    Z0P_ERR_INDEX_OUT_OF_BOUNDS = 2,       /* Error::IndexOutOfBounds      error.rs:11 */
    Z0P_ERR_INVALID_DOMAIN_SIZE = 3,       /* Error::InvalidDomainSize     error.rs:18 */
    Z0P_ERR_INVALID_BLOWUP = 4,            /* Error::InvalidBlowup         error.rs:23 */
    Z0P_ERR_NO_QUERIES = 5,                /* Error::NoQueries             error.rs:28 */
    Z0P_ERR_BAD_MERKLE_PATH = 6,           /* Error::BadMerklePath         error.rs:30 */
    Z0P_ERR_FRI_VERIFICATION_FAILED = 7,   /* Error::FriVerificationFailed error.rs:32 */
    Z0P_ERR_MALFORMED_INPUT = 8,           /* Error::MalformedInput        error.rs:34 */
    Z0P_ERR_RNG_UNAVAILABLE = 9,           /* Error::RngUnavailable        error.rs:39 */
    Z0P_ERR_DEVICE = 100,                  /* HIP runtime failure */
    Z0P_ERR_OUT_OF_MEMORY = 101,           /* device allocation failed or over budget */
    Z0P_ERR_INVALID_ARGUMENT = 102,        /* null pointer or bad handle */
    Z0P_ERR_NO_DEVICE = 103                /* no usable gfx950 device */
};

/* air::Params (src/air.rs:31-54). a0, b0, b_last are canonical field values. */
typedef struct z0p_air_params {
    uint64_t n_steps;
    uint64_t blowup;
    uint64_t num_queries;
    uint64_t final_size;
    uint32_t grinding_bits;
    uint32_t reserved0;
    uint64_t cap_height;
    uint32_t a0;
    uint32_t b0;
    uint32_t b_last;
    uint32_t reserved1;
} z0p_air_params;

/* fri::Params (src/fri.rs:105-133). */
typedef struct z0p_fri_params {
    uint64_t blowup;
    uint64_t num_queries;
    uint64_t final_size;
    uint32_t grinding_bits;
    uint32_t reserved0;
    uint64_t cap_height;
} z0p_fri_params;

/* Opaque Fiat-Shamir transcript (src/transcript.rs:54-57). */
typedef struct z0p_transcript z0p_transcript;

/* Opaque device context: one HIP stream, device buffers, cached tables. */
typedef struct z0p_ctx z0p_ctx;

/* ---- library / context (no reference counterpart: the reference has no device) ---- */
int z0p_ctx_create(int device, z0p_ctx** out);
void z0p_ctx_destroy(z0p_ctx* ctx);
/* Per-kernel HIP-event timing on the ctx stream (off by default). */
int z0p_ctx_set_profiling(z0p_ctx* ctx, int enabled);
/* Times only launches named `kernel` (as z0p_ctx_kernel_stats names them)
 * while profiling is on; NULL or "" times every launch. */
int z0p_ctx_profile_only(z0p_ctx* ctx, const char* kernel);
/* "name=launches,total_ms;..." for the launches since the last reset; returns
 * the length needed including the terminating NUL. */
size_t z0p_ctx_kernel_stats(z0p_ctx* ctx, char* buf, size_t cap);
void z0p_ctx_reset_stats(z0p_ctx* ctx);
const char* z0p_status_string(int status);
void z0p_free(void* p);

/* ---- transcript: replaces transcript::Transcript (src/transcript.rs) ---- */
/* Transcript::new (transcript.rs:62-69). */
z0p_transcript* z0p_transcript_new(const uint8_t* label, size_t label_len);
/* Transcript: Clone (transcript.rs:53). */
z0p_transcript* z0p_transcript_clone(const z0p_transcript* t);
void z0p_transcript_free(z0p_transcript* t);
/* Transcript::absorb (transcript.rs:72-77). */
void z0p_transcript_absorb(z0p_transcript* t, const char* label, const uint8_t* data, size_t len);
/* Transcript::challenge_bytes (transcript.rs:100-102). */
void z0p_transcript_challenge_bytes(z0p_transcript* t, const char* label, uint8_t out[32]);
/* Transcript::challenge_bb (transcript.rs:124-129): u64 LE of the squeeze
 * mod p, canonical. */
uint32_t z0p_transcript_challenge_bb(z0p_transcript* t, const char* label);
/* Transcript::challenge_ext (transcript.rs:136-141): the four LE words of
 * the squeeze, each reduced mod p (coordinate 0 first). */
void z0p_transcript_challenge_ext(z0p_transcript* t, const char* label, uint32_t out[4]);
/* Transcript::challenge_index (transcript.rs:148-160): *out = low u64 of the
 * squeeze masked below domain_size. Z0P_ERR_INVALID_DOMAIN_SIZE (the
 * reference's None) unless domain_size is a nonzero power of two; the
 * transcript is then left unchanged. */
int z0p_transcript_challenge_index(z0p_transcript* t, const char* label, uint64_t domain_size, uint64_t* out);
/* Transcript::grind (transcript.rs:206-228): the smallest u64 nonce whose
 * BLAKE3(BLAKE3(history) || u64 len || label || u64 nonce) has at least
 * difficulty_bits leading zero bits, searched on ctx's device, then absorbed
 * under label. difficulty_bits > 32 absorbs and returns 0, as the reference
 * does. Labels up to 976 bytes (head and nonce in one BLAKE3 chunk), else
 * Z0P_ERR_MALFORMED_INPUT. */
int z0p_transcript_grind(z0p_ctx* ctx, z0p_transcript* t, const char* label, uint32_t difficulty_bits,
                         uint64_t* nonce);
/* z0p_transcript_grind that allocates nothing: no hipMalloc, hipFree or
 * other allocating routine runs, and no profiling events are recorded. It
 * has no counterpart in the reference; it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation). The grind head and the running minimum nonce live
 * in one static 256-byte device array per GPU, shared by every call on that
 * GPU and keeping the last call's contents (the minimum is reset by a kernel
 * before each search window), so calls that overlap on one GPU race, from
 * any ctx. Same limits and results as z0p_transcript_grind: labels up to
 * 976 bytes, else Z0P_ERR_MALFORMED_INPUT. */
int z0p_transcript_grind_noalloc(z0p_ctx* ctx, z0p_transcript* t, const char* label, uint32_t difficulty_bits,
                                 uint64_t* nonce);
/* Transcript::verify_grind (transcript.rs:235-246): 1 when the nonce meets the
 * difficulty (absorbed as grind absorbs it), 0 otherwise (nothing absorbed;
 * also for difficulty_bits > 256 or a null argument). Host only. */
int z0p_transcript_verify_grind(z0p_transcript* t, const char* label, uint64_t nonce, uint32_t difficulty_bits);

/* ---- parameters ---- */
/* air::Params::new (air.rs:63-101) with grinding_bits = cap_height = 0. */
int z0p_air_params_new(uint64_t n_steps, uint64_t blowup, uint64_t num_queries, uint64_t final_size, uint32_t a0,
                       uint32_t b0, uint32_t b_last, z0p_air_params* out);
/* air::Params::for_security (air.rs:129-162). */
int z0p_air_params_for_security(uint64_t n_steps, uint32_t a0, uint32_t b0, uint32_t b_last, uint32_t target_bits,
                                z0p_air_params* out);
/* fri::Params::new (fri.rs:141-158) with grinding_bits = cap_height = 0. */
int z0p_fri_params_new(uint64_t blowup, uint64_t num_queries, uint64_t final_size, z0p_fri_params* out);
/* fri::Params::for_security (fri.rs:232-255). */
int z0p_fri_params_for_security(uint64_t poly_len, uint32_t target_bits, z0p_fri_params* out);
/* air::build_trace(a0, b0, n).1[n-1] (air.rs:341-354): the honest output claim. */
uint32_t z0p_air_b_last(uint32_t a0, uint32_t b0, uint64_t n_steps);

/* ---- proving: air.rs / fri.rs ---- */
/* air::prove (air.rs:365-372) followed by Proof::to_bytes (air.rs:262-281). */
int z0p_air_prove(z0p_ctx* ctx, const z0p_air_params* params, z0p_transcript* transcript, uint8_t** proof,
                  size_t* proof_len);
/* Proof::from_bytes + air::verify (air.rs:289-327, 607-752). Host only. */
int z0p_air_verify(const z0p_air_params* params, z0p_transcript* transcript, const uint8_t* proof, size_t proof_len);
/* fri::commit (fri.rs:408-433) + fri::public_commitment (fri.rs:863-870) +
 * fri::open (fri.rs:645-666); outputs Commitment::to_bytes and
 * Proof::to_bytes (fri.rs:875-880, 331-336). indices (may be NULL) receives
 * the num_queries sampled leaf indices that fri::open returns beside the
 * proof (bit-reversed layer-0 order); indices_cap must be at least
 * num_queries when it is not NULL (else Z0P_ERR_INVALID_ARGUMENT). */
int z0p_fri_commit_open(z0p_ctx* ctx, const uint32_t* coeffs, size_t n, const z0p_fri_params* params,
                        z0p_transcript* transcript, uint8_t** commitment, size_t* commitment_len, uint8_t** proof,
                        size_t* proof_len, uint64_t* indices, size_t indices_cap);
/* air::prove_with_trace (air.rs:385-406): explicit columns of len_a = len_b =
 * n_steps canonical values (else Z0P_ERR_MALFORMED_INPUT), both interpolated
 * and extended on the device; then Proof::to_bytes. */
int z0p_air_prove_with_trace(z0p_ctx* ctx, const z0p_air_params* params, const uint32_t* col_a, size_t len_a,
                             const uint32_t* col_b, size_t len_b, z0p_transcript* transcript, uint8_t** proof,
                             size_t* proof_len);

/* ---- hiding: replaces rng::ChaCha20 (src/rng.rs) and air::prove_zk ---- */
/* Opaque ChaCha20 stream (rng.rs:82-88). Not clonable, as in the reference. */
typedef struct z0p_chacha20 z0p_chacha20;
/* ChaCha20::new(key, nonce) (rng.rs:107-116); NULL on a null argument. */
z0p_chacha20* z0p_chacha20_new(const uint8_t key[32], const uint8_t nonce[12]);
/* rng::fresh_stream (rng.rs:185-209): key from /dev/urandom, zero nonce;
 * Z0P_ERR_RNG_UNAVAILABLE when it cannot be read. */
int z0p_chacha20_from_os(z0p_chacha20** out);
void z0p_chacha20_free(z0p_chacha20* rng);
/* ChaCha20::fill_bytes / next_u32 / next_u64 / next_field (rng.rs:118-168). */
void z0p_chacha20_fill_bytes(z0p_chacha20* rng, uint8_t* out, size_t len);
uint32_t z0p_chacha20_next_u32(z0p_chacha20* rng);
uint64_t z0p_chacha20_next_u64(z0p_chacha20* rng);
uint32_t z0p_chacha20_next_field(z0p_chacha20* rng);
/* air::prove_zk (air.rs:439-464): both columns masked by
 * blind::mask_coeffs (blind.rs:38-79) with n_rand >= 16 * num_queries
 * maskers each, drawn from rng (advanced as the reference advances it);
 * the masked LDE has next_pow2(n_steps + n_rand) * blowup points, so
 * final_size must leave a power-of-four fold ratio over it. Verified by the
 * same z0p_air_verify. */
int z0p_air_prove_zk(z0p_ctx* ctx, const z0p_air_params* params, uint64_t n_rand, z0p_chacha20* rng,
                     z0p_transcript* transcript, uint8_t** proof, size_t* proof_len);

/* fri::commit_ext (fri.rs:444-461) with ntt::evaluate_on_coset_ext
 * (ntt.rs:394-420) + public_commitment + open: n Ext4 coefficients, 4 words
 * each (coordinate 0 first). Outputs as z0p_fri_commit_open. */
int z0p_fri_commit_open_ext(z0p_ctx* ctx, const uint32_t* coeffs4, size_t n, const z0p_fri_params* params,
                            z0p_transcript* transcript, uint8_t** commitment, size_t* commitment_len, uint8_t** proof,
                            size_t* proof_len, uint64_t* indices, size_t indices_cap);
/* fri::commit_evals (fri.rs:473-500) + public_commitment + open: m base-field
 * LDE values in natural order on the 31-coset, declared degree_bound < m. */
int z0p_fri_commit_open_evals(z0p_ctx* ctx, const uint32_t* evals, size_t m, uint64_t degree_bound,
                              const z0p_fri_params* params, z0p_transcript* transcript, uint8_t** commitment,
                              size_t* commitment_len, uint8_t** proof, size_t* proof_len, uint64_t* indices,
                              size_t indices_cap);
/* fri::commit_ext_evals (fri.rs:510-549) + public_commitment + open: m Ext4
 * LDE values (4 words each) in natural order, declared degree_bound < m. */
int z0p_fri_commit_open_ext_evals(z0p_ctx* ctx, const uint32_t* evals4, size_t m, uint64_t degree_bound,
                                  const z0p_fri_params* params, z0p_transcript* transcript, uint8_t** commitment,
                                  size_t* commitment_len, uint8_t** proof, size_t* proof_len, uint64_t* indices,
                                  size_t indices_cap);

/* Commitment/Proof::from_bytes + fri::verify (fri.rs:686-859). Host only.
 * indices (may be NULL; indices_cap >= num_queries) receives the sampled
 * query indices fri::verify returns (fri.rs:686), filled only on Z0P_OK. */
int z0p_fri_verify(const uint8_t* commitment, size_t commitment_len, const uint8_t* proof, size_t proof_len,
                   const z0p_fri_params* params, z0p_transcript* transcript, uint64_t* indices, size_t indices_cap);

/* Generic byte-leaf Merkle trees (merkle.rs), host only. Paths are `steps`
 * entries of 33 bytes: the sibling digest, then the direction byte (1: the
 * running node is the left child, 0: the right; any other value fails), as
 * in the wire format's paths (codec.rs:87-99).
 * merkle::hash_leaf (merkle.rs:16-18): BLAKE3(0x00 || data). */
void z0p_merkle_hash_leaf(const uint8_t* data, size_t len, uint8_t out[32]);
/* merkle::hash_node (merkle.rs:24-26): BLAKE3(0x01 || left || right). */
void z0p_merkle_hash_node(const uint8_t left[32], const uint8_t right[32], uint8_t out[32]);
/* MerkleTree::verify (merkle.rs:165-189): 1 when the path opens the leaf
 * data at `index` under `root` (index < 2^steps, direction bits equal to
 * index's bits), else 0. */
int z0p_merkle_verify(const uint8_t root[32], const uint8_t* leaf, size_t leaf_len, uint64_t index,
                      const uint8_t* path, size_t steps);
/* MerkleTree::verify_digest (merkle.rs:196-216): the same for a leaf given
 * by its digest. */
int z0p_merkle_verify_digest(const uint8_t root[32], const uint8_t leaf_digest[32], uint64_t index,
                             const uint8_t* path, size_t steps);

/* fri::ProverState (fri.rs:276-294): the committed layers and their trees,
 * kept in device memory that the state owns, so fri::open can run later
 * (and more than once) over the same commitment. A state uses its ctx: calls
 * on it must not overlap other calls on that ctx, and it must be freed
 * before the ctx is destroyed. */
typedef struct z0p_fri_state z0p_fri_state;
/* fri::commit (fri.rs:408-433): coefficients as z0p_fri_commit_open. */
int z0p_fri_commit(z0p_ctx* ctx, const uint32_t* coeffs, size_t n, const z0p_fri_params* params,
                   z0p_transcript* transcript, z0p_fri_state** out);
/* fri::commit_ext (fri.rs:444-461). */
int z0p_fri_commit_ext(z0p_ctx* ctx, const uint32_t* coeffs4, size_t n, const z0p_fri_params* params,
                       z0p_transcript* transcript, z0p_fri_state** out);
/* fri::commit_evals (fri.rs:473-500). */
int z0p_fri_commit_evals(z0p_ctx* ctx, const uint32_t* evals, size_t m, uint64_t degree_bound,
                         const z0p_fri_params* params, z0p_transcript* transcript, z0p_fri_state** out);
/* fri::commit_ext_evals (fri.rs:510-549). */
int z0p_fri_commit_ext_evals(z0p_ctx* ctx, const uint32_t* evals4, size_t m, uint64_t degree_bound,
                             const z0p_fri_params* params, z0p_transcript* transcript, z0p_fri_state** out);
/* fri::public_commitment (fri.rs:863-870) + Commitment::to_bytes. */
int z0p_fri_state_commitment(const z0p_fri_state* state, uint8_t** commitment, size_t* commitment_len);
/* fri::open (fri.rs:645-666) + Proof::to_bytes: the proof-of-work, the
 * queries and their openings, advancing transcript; indices as in
 * z0p_fri_commit_open. */
int z0p_fri_open(z0p_fri_state* state, z0p_transcript* transcript, uint8_t** proof, size_t* proof_len,
                 uint64_t* indices, size_t indices_cap);
void z0p_fri_state_free(z0p_fri_state* state);

/* ---- range-check AIR (src/range.rs) on the same device path: the value
 * column V and n_bits bit columns extended on the 31-coset, one joint row
 * tree (leaf = BLAKE3(0x00 || V || C_0 || ...) per LDE row), the composition
 * of range.rs:767-823 committed by fri::commit_ext_evals ---- */
/* range::Params (range.rs:29-52); value is reduced mod p. */
typedef struct z0p_range_params {
    uint64_t n_rows;
    uint64_t n_bits;
    uint64_t blowup;
    uint64_t num_queries;
    uint64_t final_size;
    uint32_t grinding_bits;
    uint32_t reserved0;
    uint64_t cap_height;
    uint32_t value;
    uint32_t reserved1;
} z0p_range_params;
/* range::Params::new (range.rs:59-98): grinding 0, cap height 0. */
int z0p_range_params_new(uint64_t n_rows, uint64_t n_bits, uint64_t blowup, uint64_t num_queries,
                         uint64_t final_size, uint32_t value, z0p_range_params* out);
/* range::Params::for_security (range.rs:124-158). */
int z0p_range_params_for_security(uint64_t n_rows, uint64_t n_bits, uint32_t value, uint32_t target_bits,
                                  z0p_range_params* out);
/* range::prove (range.rs:334-342): constant columns of value and its bits. */
int z0p_range_prove(z0p_ctx* ctx, const z0p_range_params* params, z0p_transcript* transcript, uint8_t** proof,
                    size_t* proof_len);
/* range::prove_with_columns (range.rs:357-382): ncols = n_bits + 1 columns of
 * len = n_rows canonical values each, V first, contiguous. */
int z0p_range_prove_with_columns(z0p_ctx* ctx, const z0p_range_params* params, const uint32_t* columns, size_t ncols,
                                 size_t len, z0p_transcript* transcript, uint8_t** proof, size_t* proof_len);
/* range::prove_zk (range.rs:401-428): every column masked with n_rand >=
 * num_queries maskers from rng (V's first); degree bound n_rows + 2 n_rand - 2. */
int z0p_range_prove_zk(z0p_ctx* ctx, const z0p_range_params* params, uint64_t n_rand, z0p_chacha20* rng,
                       z0p_transcript* transcript, uint8_t** proof, size_t* proof_len);
/* Proof::from_bytes + range::verify (range.rs:280-306, 521-721). Host only. */
int z0p_range_verify(const z0p_range_params* params, z0p_transcript* transcript, const uint8_t* proof,
                     size_t proof_len);

/* ---- table-membership AIR (src/table.rs) on the same device path: the
 * value column extended on the 31-coset, one tree leaf per LDE value
 * (BLAKE3(0x00 || V)), the composition of table.rs:597-644 committed by
 * fri::commit_ext_evals; the public table is table_len >= 2 entries, each
 * reduced mod p ---- */
/* table::Params (table.rs:27-46); value is reduced mod p. */
typedef struct z0p_table_params {
    uint64_t n_rows;
    uint64_t blowup;
    uint64_t num_queries;
    uint64_t final_size;
    uint32_t grinding_bits;
    uint32_t reserved0;
    uint64_t cap_height;
    uint32_t value;
    uint32_t reserved1;
} z0p_table_params;
/* table::Params::new (table.rs:53-87): grinding 0, cap height 0. */
int z0p_table_params_new(uint64_t n_rows, uint64_t blowup, uint64_t num_queries, uint64_t final_size, uint32_t value,
                         z0p_table_params* out);
/* table::Params::for_security (table.rs:114-168): the smallest blowup >= 4
 * with rate <= 1/4 at this table length. */
int z0p_table_params_for_security(uint64_t n_rows, uint64_t table_len, uint32_t value, uint32_t target_bits,
                                  z0p_table_params* out);
/* table::membership_bound (table.rs:357-363), saturating. */
uint64_t z0p_table_membership_bound(uint64_t table_len, uint64_t n_rows, uint64_t n_rand);
/* table::prove (table.rs:305-318). */
int z0p_table_prove(z0p_ctx* ctx, const z0p_table_params* params, const uint32_t* table, size_t table_len,
                    z0p_transcript* transcript, uint8_t** proof, size_t* proof_len);
/* table::prove_zk (table.rs:329-355): the column masked with n_rand >=
 * num_queries maskers from rng. */
int z0p_table_prove_zk(z0p_ctx* ctx, const z0p_table_params* params, const uint32_t* table, size_t table_len,
                       uint64_t n_rand, z0p_chacha20* rng, z0p_transcript* transcript, uint8_t** proof,
                       size_t* proof_len);
/* Proof::from_bytes + table::verify (table.rs:265-288, 430-538). Host only. */
int z0p_table_verify(const z0p_table_params* params, const uint32_t* table, size_t table_len,
                     z0p_transcript* transcript, const uint8_t* proof, size_t proof_len);

/* ---- IVC accumulation (src/fold.rs): hash-chained Fibonacci segments, each
 * committed with its own row tree and six alphas, and one batched FRI over
 * the gamma-combination of their compositions ---- */
/* fold::Params (fold.rs:41-65); in_a, in_b, out_b are reduced mod p. */
typedef struct z0p_fold_params {
    uint64_t seg_steps;
    uint64_t blowup;
    uint64_t num_queries;
    uint64_t final_size;
    uint32_t grinding_bits;
    uint32_t reserved0;
    uint64_t cap_height;
    uint32_t in_a;
    uint32_t in_b;
    uint32_t out_b;
    uint32_t reserved1;
} z0p_fold_params;
/* fold::Params::new (fold.rs:72-111): grinding 0, cap height 0. */
int z0p_fold_params_new(uint64_t seg_steps, uint64_t blowup, uint64_t num_queries, uint64_t final_size,
                        uint32_t in_a, uint32_t in_b, uint32_t out_b, z0p_fold_params* out);
/* fold::Params::for_security (fold.rs:136-170). */
int z0p_fold_params_for_security(uint64_t seg_steps, uint32_t in_a, uint32_t in_b, uint32_t out_b,
                                 uint32_t target_bits, z0p_fold_params* out);
/* fold::Accumulator (fold.rs:361-400): segments stay in device memory on
 * ctx until finalize; their buffers then return to the ctx for the next
 * accumulator and are freed with the ctx. Free every accumulator before its
 * ctx. Calls on one accumulator and other calls on its ctx must not overlap. */
typedef struct z0p_fold_acc z0p_fold_acc;
int z0p_fold_acc_new(z0p_ctx* ctx, const z0p_fold_params* params, z0p_fold_acc** out);
void z0p_fold_acc_free(z0p_fold_acc* acc);
/* Accumulator::len (fold.rs:388-391). */
size_t z0p_fold_acc_len(const z0p_fold_acc* acc);
/* Accumulator::add_segment (fold.rs:407-492): columns of seg_steps canonical
 * values each (else Z0P_ERR_MALFORMED_INPUT). */
int z0p_fold_acc_add_segment(z0p_fold_acc* acc, const uint32_t* col_a, size_t len_a, const uint32_t* col_b,
                             size_t len_b, z0p_transcript* transcript);
/* Accumulator::finalize (fold.rs:494-570) + IvcProof::to_bytes (fold.rs:281-310).
 * Consumes the segments: the accumulator is empty afterwards. */
int z0p_fold_acc_finalize(z0p_fold_acc* acc, z0p_transcript* transcript, uint8_t** proof, size_t* proof_len);
/* IvcProof::from_bytes + fold::verify (fold.rs:312-359, 572-667). Host only. */
int z0p_fold_verify(const z0p_fold_params* params, z0p_transcript* transcript, const uint8_t* proof,
                    size_t proof_len);

/* ---- batched proving (no reference counterpart: the reference proves one
 * instance per air::prove call; this is `count` such calls at once) ---- */
/* Device-memory budget of the batch and server schedulers on this ctx, in
 * bytes; 0 (the default) lets the library choose from free device memory. */
int z0p_ctx_set_hbm_budget(z0p_ctx* ctx, uint64_t bytes);
/* Batch and server schedulers on this ctx split the active proofs into two
 * alternating waves once at least `min_active` are active (0: never), so the
 * host's transcript steps for one wave overlap the other wave's kernels. The
 * proofs are the same either way. Default 8. A server started on the ctx
 * uses the setting in both of its lanes. */
int z0p_ctx_set_waves(z0p_ctx* ctx, uint32_t min_active);
/* Runs air::prove for every params[i] with transcripts[i] (each advanced as
 * that single call would advance it). The instances share every kernel
 * launch and are admitted as device memory allows. statuses[i] receives
 * instance i's status; proofs[i] / proof_lens[i] are set when it is 0.
 * Returns Z0P_OK when the batch ran (per-instance failures are only in
 * statuses) or a device error that ended the whole batch. */
int z0p_air_prove_batch(z0p_ctx* ctx, const z0p_air_params* params, z0p_transcript* const* transcripts,
                        size_t count, uint8_t** proofs, size_t* proof_lens, int* statuses);

/* ---- batched verification (no reference counterpart: `count` calls of
 * air::verify, air.rs:607-752, on host threads) ---- */
/* Verifies proofs[i] under params[i] and transcripts[i] (each advanced as its
 * own z0p_air_verify would; null or shared transcripts are rejected with
 * Z0P_ERR_INVALID_ARGUMENT). statuses[i] receives instance i's status.
 * threads 0: one per hardware thread. */
/* z0p_air_verify_batch with the Merkle path hashing on ctx's device: host
 * threads parse each proof and replay its transcript with every other check,
 * then one kernel per chunk of 64 proofs walks all their authentication
 * paths (one per lane). statuses[i] equals what z0p_air_verify returns for
 * instance i (a failed path check recorded before a host-side failure is
 * reported as Z0P_ERR_BAD_MERKLE_PATH, as the serial order would). On a
 * rejected instance the transcript may be advanced past the failing check.
 * The batched verification service of the survey's ranked extensions over
 * air::verify (air.rs:607-752, fri.rs:686-859). */
int z0p_air_verify_batch_device(z0p_ctx* ctx, const z0p_air_params* params, z0p_transcript* const* transcripts,
                                const uint8_t* const* proofs, const size_t* proof_lens, size_t count,
                                uint32_t threads, int* statuses);
int z0p_air_verify_batch(const z0p_air_params* params, z0p_transcript* const* transcripts,
                         const uint8_t* const* proofs, const size_t* proof_lens, size_t count, uint32_t threads,
                         int* statuses);

/* ---- kernel schedule and autotuner (no reference counterpart: launch
 * shapes of this library; proof bytes never depend on them) ---- */
typedef struct z0p_schedule {
    int32_t tail_max_log;       /* tree levels of at most 2^k digests use the quad-lane tail (2..24) */
    int32_t tail_group_max_log; /* tail workgroup: at most 2^k digests (2..9) */
    int32_t tail_spread_log;    /* smaller tail groups so a level spans ~2^k workgroups (0..16) */
    int32_t levels_fuse;        /* wide tree levels per launch (1 or 2) */
    int32_t small_max_log;      /* fused fold + commit for FRI layers of at most 2^k values (0 = off, 4..10) */
    int32_t grind_blocks;       /* proof-of-work workgroups per launch (16..2^20) */
    int32_t fold_fuse_max_log;  /* fold + next layer's leaves in one kernel for layers of at most 2^k values (0 = off, 3..27) */
    int32_t compose_fuse;       /* composition + FRI layer 0's leaves in one kernel (0 or 1) */
    int32_t trace_fuse;         /* linked column + trace leaves in one kernel (0 or 1) */
    int32_t subtree_levels;     /* tree levels above level 1 reduced in LDS by each leaf workgroup (0 = off, 1..8) */
    int32_t compose_subtree;    /* the same inside the composition kernel (0 = off, 1..8) */
    int32_t graph;              /* launch each scheduler iteration as one captured HIP graph (0 or 1) */
    int32_t lde_split;          /* LDE row and column passes in kernels of their own, else in the shared NTT kernel (0 or 1) */
    int32_t grind_pool;         /* workgroups of the pooled proof-of-work search when several proofs grind in one launch
                                 * (0 = each proof on workgroups of its own; 0..8192) */
    int32_t fold_commit_max_log; /* FRI layers of 2^8..2^k values committed in one launch: fold, leaves and the
                                  * whole tree, its upper levels by the last workgroup of each group (0 = off, 8..27) */
    int32_t tree_top;           /* a proof tree's narrow levels in one staged launch instead of tail launches (0 or 1) */
    int32_t wave_streams;       /* the two waves of a batch or server lane on two HIP streams, so that one wave's
                                 * kernels run beside the other's (0 = one stream; 0 or 1) */
} z0p_schedule;
int z0p_ctx_get_schedule(z0p_ctx* ctx, z0p_schedule* out);
/* Z0P_ERR_INVALID_ARGUMENT for a value outside the ranges above. */
int z0p_ctx_set_schedule(z0p_ctx* ctx, const z0p_schedule* s);
/* Proves `work` (one pipelined proof when count is 1, else one batch) under
 * candidate schedules, coordinate descent over the knobs, `reps` timed runs
 * per candidate (median), and leaves the fastest set on the ctx. Outputs
 * (each may be NULL): the schedule, its and the starting schedule's
 * milliseconds per run of the workload. */
int z0p_ctx_autotune(z0p_ctx* ctx, const z0p_air_params* work, size_t count, uint32_t reps, z0p_schedule* out,
                     double* best_ms, double* start_ms);
/* "knob=value:ms;..." for every trial of the last autotune on this ctx;
 * returns the length needed including the terminating NUL. */
size_t z0p_ctx_autotune_log(z0p_ctx* ctx, char* buf, size_t cap);

/* ---- serving (no reference counterpart) ----
 * An asynchronous request loop that takes over the ctx until
 * z0p_server_destroy. Requests below 2^16 steps and the larger ones run in two
 * lanes, each a scheduler on its own host thread and HIP stream (the large
 * lane on a second context the server creates on the same device with the
 * ctx's kernel schedule), so short requests are not held behind long ones.
 * Requests arriving while others run are admitted into their lane's running
 * batch as device memory and slots free up, so their device work overlaps;
 * a request alone in its lane runs all its phases back to back. Every result
 * is byte-identical to the proof z0p_air_prove returns for the same request. */
typedef struct z0p_server z0p_server;
/* hbm_budget 0: half of the free device memory; the small lane gets half of
 * the budget, at most 1 GiB. max_active 0: 256 per lane. */
int z0p_server_create(z0p_ctx* ctx, uint64_t hbm_budget, uint32_t max_active, z0p_server** out);
/* z0p_server_create with a kernel schedule of its own for the large lane
 * (NULL: the ctx's, as z0p_server_create): the two lanes see different proof
 * sizes, and a schedule tuned on 2^16..2^20-step batches runs the large lane
 * faster than one tuned on the whole mix. Z0P_ERR_INVALID_ARGUMENT when a
 * field is out of range. */
int z0p_server_create_lanes(z0p_ctx* ctx, const z0p_schedule* large_schedule, uint64_t hbm_budget,
                            uint32_t max_active, z0p_server** out);
/* Queues one AIR proof. The transcript is copied: the request proves over
 * its own copy and the caller's transcript is left unchanged. */
int z0p_server_submit_air(z0p_server* s, const z0p_air_params* params, const z0p_transcript* transcript,
                          uint64_t* request_id);
/* Takes one finished request, waiting up to timeout_ms (0: no wait, < 0:
 * until one finishes). Returns Z0P_OK with *request_id set, or
 * Z0P_ERR_EMPTY_INPUT when none finished in time. The request's own status
 * goes to *status; *proof (free with z0p_free) is set when it is 0.
 * Latencies are milliseconds from submission to completion and to
 * admission onto the device. */
int z0p_server_poll(z0p_server* s, int timeout_ms, uint64_t* request_id, int* status, uint8_t** proof,
                    size_t* proof_len, double* latency_ms, double* queue_ms);
/* Requests submitted and not yet polled (queued, running or finished). */
size_t z0p_server_pending(z0p_server* s);
/* Finishes every running request, drops unpolled results, releases the ctx. */
void z0p_server_destroy(z0p_server* s);

#ifdef __cplusplus
}
#endif

#endif /* Z0P_H */
