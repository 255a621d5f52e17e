/*
 * Kernel-level diagnostic entry points of libz0p_amd.so. They have no
 * counterpart in the reference crate: each runs one device stage of the
 * prover on host-supplied data so parity tests can localise a mismatch to a
 * single kernel. Same status conventions as z0p.h; all buffers are host
 * memory, copied in and out by the call.
 */
#ifndef Z0P_TESTING_H
#define Z0P_TESTING_H

#include "z0p.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Inverse NTT of n natural-order evaluations; coefficients returned in
 * bit-reversed order (position q holds coefficient brev(q)). */
int z0p_test_intt(z0p_ctx* ctx, const uint32_t* evals, size_t n, uint32_t* coeffs_br);

/* Coset LDE (shift 31, size n*blowup) of n natural-order coefficients,
 * returned in bit-reversed order over the whole domain. */
int z0p_test_lde(z0p_ctx* ctx, const uint32_t* coeffs, size_t n, size_t blowup, uint32_t* lde_br);

/* Merkle tree over 64-byte leaves (leaves a power of two): all levels leaf
 * to root, concatenated, 32 bytes per digest. */
int z0p_test_merkle64(z0p_ctx* ctx, const uint8_t* flat, size_t leaves, uint8_t* levels);

/* Trace tree leaves+levels from natural-order LDE columns a, b (m = n *
 * blowup each): all levels leaf to root, 32 bytes per digest. */
int z0p_test_trace_tree(z0p_ctx* ctx, const uint32_t* a, const uint32_t* b, size_t n, size_t blowup,
                        uint8_t* levels);

/* Linked LDE column B from natural-order A (air.rs:806-844). */
int z0p_test_linked(z0p_ctx* ctx, const uint32_t* a, size_t n, size_t blowup, uint32_t actual_b_last, uint32_t a0,
                    uint32_t* b);

/* Composition over natural-order columns; out is m Ext4 (4 words each) in
 * natural order (air.rs:847-929). alphas: 5 Ext4 = 20 words. With levels01
 * non-null, the kernel that also hashes FRI layer 0 runs instead and
 * levels01 receives that layer's quad-leaf digests (m/4) then their parents
 * (m/8), 32 bytes each. */
int z0p_test_compose(z0p_ctx* ctx, const uint32_t* a, const uint32_t* b, size_t n, size_t blowup, uint32_t a0,
                     uint32_t b0, uint32_t b_last, const uint32_t* alphas, uint32_t* out, uint8_t* levels01);

/* One arity-4 fold of a bit-reversed Ext4 layer of length len at fold depth
 * `depth` of an LDE of size lde_len; challenges: 3 Ext4 = 12 words; out: the
 * len/4 folded values. With levels01 non-null (len >= 32), the kernel that
 * also hashes the folded layer runs instead and levels01 receives its quad-
 * leaf digests (len/16) then their parents (len/32), 32 bytes each. */
int z0p_test_fold(z0p_ctx* ctx, const uint32_t* layer, size_t len, size_t lde_len, size_t depth,
                  const uint32_t* challenges, uint32_t* out, uint8_t* levels01);

/* Proof-of-work: smallest nonce whose BLAKE3(head || u64 nonce) has at least
 * `bits` (<= 32) leading zero bits; head_len + 8 <= 1024 (one chunk). */
int z0p_test_grind(z0p_ctx* ctx, const uint8_t* head, size_t head_len, uint32_t bits, uint64_t* nonce);

/* Plain device copy of `bytes` (multiple of 16) between two arena buffers;
 * *ms receives its event-timed duration. The calibration kernel for the
 * FETCH_SIZE / WRITE_SIZE counters. */
int z0p_test_copy(z0p_ctx* ctx, size_t bytes, float* ms);

/* Polls of host memory that a pre-enqueued phase's gate makes before giving
 * up (default 2^22). 0 makes every gate give up at once: a lone proof then
 * fails with Z0P_ERR_DEVICE while every kernel still runs to completion, the
 * way a host that never answers would end it. */
int z0p_test_set_gate_spin(z0p_ctx* ctx, uint32_t polls);

#ifdef __cplusplus
}
#endif

#endif /* Z0P_TESTING_H */
