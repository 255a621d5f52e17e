// Kernel launchers (host-callable). Every launcher enqueues on `st` and
// performs no synchronisation; device pointers come from the context arena.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace z0p {

struct Profiler;  // per-kernel event timing, see prover.cpp

struct Launch {
    hipStream_t stream;
    Profiler* prof;  // may be null
};

// Two-level power tables for omega_m^{+-e}, e < m (m = lde size):
// w^e = lo[e & 4095] * hi[e >> 12], entries in Montgomery form.
struct PowTable {
    const uint32_t* lo;
    const uint32_t* hi;
};

// Per-coset constants of an LDE with blowup B (all Montgomery form):
//   shift[j] = s_j = 31 * Omega^j, used for the point x = s_j * omega_n^t.
struct CosetTables {
    const uint32_t* shift;     // [B]
    const uint32_t* small;     // single-pass NTT: [B][n] s_j^k
    const uint32_t* row;       // two-pass: [B][n2] (s_j^{n1})^r
    const uint32_t* col;       // two-pass: [B][n1] s_j^c
};

// NTT geometry: n = 2^logn; two-pass split n = n1 * n2 (n1 strided length).
struct NttShape {
    int logn, log1, log2;
};
NttShape ntt_shape(int logn);

void launch_fib_trace(const Launch& L, uint32_t* col, uint64_t n, uint32_t a0, uint32_t b0);

// Inverse NTT: natural-order evaluations -> coefficients in bit-reversed
// order, scaled by 1/n. scratch must hold n words (two-pass only).
void launch_intt_to_br(const Launch& L, const uint32_t* evals, uint32_t* coeffs_br, uint32_t* scratch, int logn,
                       const uint32_t* tw4096_inv, PowTable winv_m, int log_blowup, uint32_t ninv_m);

// Coset LDE: bit-reversed coefficients (n) -> m = B*n evaluations on
// 31*<Omega_m> in bit-reversed order (coset j lands in block brev(j)).
// scratch must hold m words (two-pass only).
void launch_lde_br(const Launch& L, const uint32_t* coeffs_br, uint32_t* lde_br, uint32_t* scratch, int logn,
                   int log_blowup, const uint32_t* tw4096, PowTable w_m, CosetTables ct);

// In-place bit-reversal permutation of n words into out (out != in).
void launch_bitrev(const Launch& L, const uint32_t* in, uint32_t* out, int logn);

// Linked second column (air.rs:806-844) in bit-reversed layout.
void launch_linked_b(const Launch& L, const uint32_t* a_br, uint32_t* b_br, int logn, int log_blowup,
                     PowTable w_m, const uint32_t* shift_m, const uint32_t* coeff_van, uint32_t w_last_m);

// Composition (air.rs:847-929) written as FRI layer 0 (bit-reversed Ext4).
struct ComposeConsts {
    uint32_t w_last_m;
    uint32_t a0, b0, b_last;
    uint32_t alpha_m[5][4];
};
void launch_compose(const Launch& L, const uint32_t* a_br, const uint32_t* b_br, uint32_t* layer0, int logn,
                    int log_blowup, PowTable w_m, const uint32_t* shift_m, const uint32_t* zinv_m,
                    const ComposeConsts& c);

// Trace Merkle leaves (8 joint rows A||B per 64-byte leaf) from the
// bit-reversed columns, written in natural leaf order into `levels`
// (contiguous level arrays). With B = 16 the two leaves of a trace step are
// also reduced to their parent (level 1) when max_level >= 1. Returns the
// highest level written.
int launch_trace_leaves(const Launch& L, const uint32_t* a_br, const uint32_t* b_br, uint32_t* levels, int logn,
                        int log_blowup, int max_level);

// Generic fused subtree levels: from `count` digests at level_base produce
// up to `nlevels` (<= 9) parent levels, stored contiguously after it.
void launch_merkle_levels(const Launch& L, uint32_t* level_base, uint64_t count, int nlevels);

// FRI layer leaves: each leaf is 4 consecutive Ext4 values (64 bytes); base
// layers hold one u32 per value with the upper coordinates zero. Writes the
// leaf level and up to `nlevels` more fused levels contiguously.
void launch_fri_leaves(const Launch& L, const uint32_t* layer, bool base_field, uint32_t* levels, uint64_t leaves,
                       int nlevels);

// Arity-4 fold (fri.rs:1273-1331) of a bit-reversed layer of length len.
struct FoldConsts {
    uint32_t ch_m[3][4];
    uint32_t unit_m;
    uint32_t ginv_m;      // 31^{-4^d}, Montgomery
    int log_len;          // log2 of this layer's length
    int depth;            // fold depth d
};
void launch_fold(const Launch& L, const uint32_t* layer, bool base_field, uint32_t* next, PowTable winv_m,
                 const FoldConsts& c);

// Proof-of-work: smallest nonce in [0, limit) whose digest of head||nonce has
// >= bits leading zero bits; result (u64, ~0 when none) written to *result.
void launch_grind(const Launch& L, const uint32_t head_words[16], uint32_t head_len, uint32_t bits, uint64_t limit,
                  unsigned long long* result);

// Streaming copy of `words` (a multiple of 4) u32: counter calibration.
void launch_copy(const Launch& L, const uint32_t* in, uint32_t* out, uint64_t words);

// Gather words: for item i copy len[i] words from base + src[i] to out + dst[i].
void launch_gather(const Launch& L, const uint32_t* base, const uint64_t* src, const uint32_t* dst,
                   const uint32_t* len, uint32_t items, uint32_t* out);

}  // namespace z0p
