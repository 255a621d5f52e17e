// Kernel launchers (host-callable).
//
// Every kernel is job-table driven: one launch covers one or many proofs
// ("jobs") of possibly different sizes. Each job descriptor records the
// first workgroup it owns (block_begin); a workgroup finds its job by binary
// search over the table. A single job travels as a by-value kernel argument,
// so the one-proof path needs no descriptor upload. Launchers enqueue on the
// given stream and never synchronise.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "field.hpp"

namespace z0p {

struct Profiler;

// Copies a host descriptor table into device memory for one launch.
struct TableSink {
    virtual const void* push(const void* host, size_t bytes) = 0;
    virtual ~TableSink() = default;
};

// dry: compute and push the descriptor tables only, launch nothing. A
// scheduler runs its launch sequence once dry, uploads every table with one
// copy, then replays the sequence, which pushes the same tables in the same
// order and so receives the same device addresses.
struct Launch {
    hipStream_t stream;
    Profiler* prof;   // may be null
    TableSink* sink;  // required for launches with more than one job
    bool dry = false;
    bool lde_split = true;  // 2^16..2^24-point LDE passes in kernels of their own (else in k_ntt)
    uint32_t grind_pool = 0;  // workgroups of the pooled search for launches of several grinds (0: per-job workgroups)
};

// Two-level power tables for omega_m^{+-e}, e < m:
// w^e = lo[e & 4095] * hi[e >> 12], entries in Montgomery form.
struct PowTable {
    const uint32_t* lo;
    const uint32_t* hi;
};

// Per-coset constants of an LDE with blowup B (Montgomery form).
struct CosetTables {
    const uint32_t* shift;  // [B] s_j = 31 * Omega^j
    const uint32_t* small;  // single-pass NTT: [B][n] s_j^k
    const uint32_t* row;    // two-pass: [B][n2] (s_j^{n1})^brev(r)
    const uint32_t* col;    // two-pass: [B][n1] s_j^c
};

// NTT geometry: n = 2^logn; two-pass split n = n1 * n2 (n1 strided length).
struct NttShape {
    int logn, log1, log2;
};
NttShape ntt_shape(int logn);

// ---- descriptors ----------------------------------------------------------
// Interpolant coefficients of a column c1 phi^i + c2 psi^i (i < n), phi and
// psi the roots of x^2 = x + 1: coefficient k is
// k1 / (1 - phi w^-k) + k2 / (1 - psi w^-k), with k1 = c1 (1 - phi^n) / n,
// k2 = c2 (1 - psi^n) / n and w = omega_n (a geometric sum per root). Written
// bit-reversed, as the inverse NTT writes them. Requires phi^n != 1 and
// psi^n != 1 (no denominator vanishes).
struct FibCoeffJob {
    uint32_t* out;
    PowTable winv;  // omega_m^-e; omega_n^-k = omega_m^-(k << logb)
    uint32_t phi_m, psi_m, k1_m, k2_m;
    int logn, logb;
    uint32_t block_begin;
};

enum NttKind : int {
    NTT_INTT_SMALL = 0,  // natural evals -> bit-reversed coeffs * 1/n (n <= 4096)
    NTT_INTT_COLS = 1,   // pass A of the two-pass inverse
    NTT_INTT_ROWS = 2,   // pass B of the two-pass inverse
    NTT_LDE_SMALL = 3,   // bit-reversed coeffs -> coset evals (n <= 4096)
    NTT_LDE_ROWS = 4,    // pass A of the two-pass coset LDE
    NTT_LDE_COLS = 5,    // pass B of the two-pass coset LDE
};
struct NttJob {
    const uint32_t* in;
    uint32_t* out;
    const uint32_t* tw4096;  // omega_4096^{+-e}, e < 2048
    PowTable pt;             // omega_m^{+-e}
    const uint32_t* tab_a;   // LDE_SMALL: scale; LDE_ROWS: rowtab
    const uint32_t* tab_b;   // LDE_ROWS: coltab
    uint32_t ninv_m;
    int kind, logn, log1, log2, logb, cols;
    // LDE kinds, hiding mask (blind.rs:38-79): coefficient k of coset j
    // gains sum_b (z_j - 1) z_j^b mask[b n + k] for b n + k < mask_len, with
    // z_j = s_j^n and zvan[j] = z_j - 1 (Montgomery). mask is canonical.
    const uint32_t* mask;
    const uint32_t* zvan;
    uint32_t mask_len;
    uint32_t block_begin;
};

// Per-point inverses of an AIR geometry (trace length n, m LDE points),
// Montgomery form, in the bit-reversed LDE order, for x_p the point at
// position p and w_last = omega_n^(n-1):
//   i1[p] = 1/(x_p - 1), iw[p] = 1/(x_p - w_last), iz[p] = (x_p - w_last)/(x_p^n - 1).
// They depend on the domain only, like the twiddle tables.
struct PointInverses {
    const uint32_t* i1 = nullptr;
    const uint32_t* iw = nullptr;
    const uint32_t* iz = nullptr;
};
struct PointInvJob {
    uint32_t *i1, *iw, *iz;
    const uint32_t* xn;
    const uint32_t* shift_m;
    const uint32_t* zinv_m;
    uint32_t w_last_m;
    int logn, logb;
};
void launch_point_inv(const Launch& L, const PointInvJob& job);

struct LinkedJob {
    const uint32_t* a;
    uint32_t* b;
    const uint32_t* iw;      // [m] 1/(x_p - w_last), Montgomery (PointInverses)
    const uint32_t* zvan_m;  // [B] s_j^n - 1, Montgomery
    uint32_t coeff;          // (b_last - a0) w_last / n, canonical
    int logn, logb;
    uint32_t block_begin;
};

struct ComposeConsts {
    uint32_t w_last_m;
    uint32_t a0, b0, b_last;
    uint32_t alpha_m[5][4];
};
struct ComposeJob {
    const uint32_t* a;
    const uint32_t* b;
    uint32_t* layer0;
    PointInverses pinv;      // 1/(x - 1), 1/(x - w_last), 1/Z per position
    const ComposeConsts* c;  // device memory: the alphas come from the transcript
    int logn, logb;
    uint32_t block_begin;
    // Non-null: also hash layer 0's quad leaves into level 0 of this tree
    // (and their parents into level 1 when with_parent), one leaf pair per
    // thread, so the values are hashed while still in registers.
    uint32_t* tree = nullptr;
    int with_parent = 0;
    int sub_levels = 0;  // levels above level 1 reduced in LDS per workgroup
};

// Range-check composition (range.rs:767-823), written as FRI layer 0:
//   a_d (V - sum_c 2^c C_c)/Z + sum_c a_c C_c (C_c - 1)/Z + a_b (V - value)/(x - 1)
// over the columns V, C_0..C_{nbits-1} (bit-reversed LDEs, col_stride apart).
constexpr int RANGE_MAX_BITS = 24;
struct RangeConsts {
    uint32_t value, nbits;
    uint32_t alpha_m[RANGE_MAX_BITS + 2][4];  // decomposition, bits, boundary (Montgomery)
};
struct RangeComposeJob {
    const uint32_t* cols;
    uint64_t col_stride;
    uint32_t* layer0;
    const uint32_t* xn;      // [n] omega_n^brev(q), Montgomery
    const uint32_t* shift_m;
    const uint32_t* zinv_m;  // [B] 1/(s_j^n - 1)
    const RangeConsts* c;    // device memory: the alphas come from the transcript
    int logn, logb;
    uint32_t block_begin;
};
void launch_range_compose(const Launch& L, std::vector<RangeComposeJob> jobs);

// Table-membership composition (table.rs:597-644), written as FRI layer 0:
//   a_m prod_k (V - t_k) / Z + a_b (V - value) / (x - 1).
struct TableConsts {
    uint32_t value, reserved;
    uint32_t alpha_m[2][4];  // membership, boundary (Montgomery)
};
struct TableComposeJob {
    const uint32_t* v;        // bit-reversed LDE of the value column
    uint32_t* layer0;
    const uint32_t* xn;       // [n] omega_n^brev(q), Montgomery
    const uint32_t* shift_m;
    const uint32_t* zinv_m;   // [B] 1/(s_j^n - 1)
    const TableConsts* c;     // device memory: the alphas come from the transcript
    const uint32_t* table_m;  // the public table, Montgomery
    uint32_t table_len;
    int logn, logb;
    uint32_t block_begin;
};
void launch_table_compose(const Launch& L, std::vector<TableComposeJob> jobs);

// IVC accumulation (fold.rs:494-570): one committed segment of the
// Fibonacci relation. a, b are its bit-reversed LDE columns (m words each),
// tree its row tree (one leaf per LDE row); beta_m[t] = gamma * alpha_t
// (Montgomery), so sum_k gamma_k C_k = sum_k sum_t beta_{k,t} s_{k,t}.
struct FoldStepDesc {
    const uint32_t* a;
    const uint32_t* b;
    const uint32_t* tree;
    uint32_t in_a, in_b, out_a, out_b;
    uint32_t beta_m[6][4];
};
// Layer 0 of the batched FRI: the gamma-combination of every step's
// composition (fold.rs:785-854), written in bit-reversed order.
struct FoldCombineJob {
    const FoldStepDesc* steps;
    uint32_t nsteps;
    uint32_t* layer0;
    const uint32_t* xn;      // [n] omega_n^brev(q), Montgomery
    const uint32_t* shift_m;
    const uint32_t* zinv_m;  // [B] 1/(s_j^n - 1)
    uint32_t w_last_m;
    int logn, logb;
    uint32_t block_begin;
};
void launch_fold_combine(const Launch& L, std::vector<FoldCombineJob> jobs);
// Step openings (fold.rs:540-566): for step k and query q, written to
// out + (k nq + q) words_per_query: A, B at the natural index brev(query)
// and at the next row (index + m/n), then both row paths (target digests
// each, leaf level first).
struct FoldOpenJob {
    const FoldStepDesc* steps;
    uint32_t nsteps, nq;
    const uint64_t* queries;  // device memory
    int logn, logb, target;
    uint32_t words_per_query;
    uint32_t* out;
};
void launch_fold_open(const Launch& L, const FoldOpenJob& job);

// Merkle leaves. Kinds: trace rows (8 joint A||B rows per leaf from the
// bit-reversed columns), FRI Ext4 quads, FRI base-field quads (upper
// coordinates zero), raw 64-byte leaves.
// LEAF_TRACE_LINKED: as LEAF_TRACE, but column B is computed from A in the
// same pass (the linked column, air.rs:806-844) and written to b_out.
// LEAF_ROWS: one leaf per LDE row of ncols bit-reversed columns (col_stride
// words apart), BLAKE3(0x00 || row words LE) at the row's natural index
// (merkle.rs:62-79 with hash.rs:352-382): the range AIR's joint row tree.
// LEAF_NODES: the "leaves" are the nodes of a tree level above level h: leaf
// i = node(src digest 2i, src digest 2i + 1), src being level h; `levels`
// then points at level h + 1 and `leaves` counts its nodes.
enum LeafKind : int {
    LEAF_TRACE = 0,
    LEAF_EXT = 1,
    LEAF_BASE = 2,
    LEAF_RAW = 3,
    LEAF_TRACE_LINKED = 4,
    LEAF_ROWS = 5,
    LEAF_NODES = 6
};
// Kinds whose workgroups own contiguous leaves (subtree fusion possible).
constexpr bool leaf_contiguous(int kind) { return kind == LEAF_EXT || kind == LEAF_BASE || kind == LEAF_RAW || kind == LEAF_NODES; }
// A workgroup reduces at most this many levels above level 1 in LDS (its 256
// level-1 digests down to one).
constexpr int SUBTREE_MAX_LEVELS = 8;
struct LeavesJob {
    const uint32_t* src;
    const uint32_t* src_b;  // trace column B
    uint32_t* levels;       // level 0 then level 1, contiguous
    uint64_t leaves;
    int kind;
    int with_parent;        // also write level 1 (needs leaves >= 2)
    int logn, logb;         // trace geometry
    uint32_t block_begin;
    // Contiguous kinds with with_parent: levels 2 .. 1 + sub_levels too, each
    // workgroup reducing its 256 level-1 digests in LDS (0: none).
    int sub_levels = 0;
    // LEAF_TRACE_LINKED (leaves >= 2): B = A[next row] + coeff (x^n - 1)/(x - w_last).
    uint32_t* b_out = nullptr;
    const uint32_t* iw = nullptr;       // [m] 1/(x_p - w_last), Montgomery (PointInverses)
    const uint32_t* zvan_m = nullptr;   // [B] s_j^n - 1, Montgomery
    uint32_t coeff = 0;                 // (b_last - a0) w_last / n, canonical
    // LEAF_ROWS: src holds ncols columns of col_stride words each.
    int ncols = 0;
    uint64_t col_stride = 0;
};

// One or two parent levels from a level of `count` digests, one node (or
// one node pair plus their parent) per thread: the wide levels.
struct LevelsJob {
    uint32_t* level;  // this level; parents follow contiguously
    uint64_t count;
    int fuse;         // 1 or 2 levels
    uint32_t block_begin;
};

// Up to group_log levels in LDS per 2^group_log-digest group (one
// workgroup of 2^(group_log+1) threads): the narrow tail, one four-lane quad
// per parent node.
struct TailJob {
    uint32_t* level;
    uint64_t count;
    int nlevels;
    int group_log;  // TAIL_GROUP_MIN_LOG..TAIL_GROUP_MAX_LOG
    uint32_t block_begin;
};

// Fold of layer d with challenges r0, r1, r2: out = M/4 + x^-1 (O r0/4 +
// x^-1 (E r1/4 + x^-1 S r2/4)) with (M, O, E, S) the 4-point inverse
// transform of the quad without its 1/4 (fri.rs:1273-1331).
struct FoldConsts {
    E4Mat cm[3];      // r_k / 4 as multiplication matrices (Montgomery)
    uint32_t inv4_m;  // 1/4, Montgomery
    uint32_t unit_m;  // two_adic_generator(2), Montgomery
    uint32_t ginv_m;  // 31^{-4^d}, Montgomery
    int log_len;      // log2 of this layer's length
    int depth;        // fold depth d
};
inline FoldConsts make_fold_consts(const E4 (&r)[3], int log_len, int depth) {
    FoldConsts fc{};
    const uint32_t inv4 = inv_c(4);
    for (int k = 0; k < 3; ++k) fc.cm[k] = e4_mat_m(e4_mul_base(r[k], inv4));
    fc.inv4_m = to_mont(inv4);
    fc.unit_m = to_mont(root_of_unity(2));
    fc.ginv_m = to_mont(pow_c(inv_c(GENERATOR), uint64_t(1) << (2 * depth)));
    fc.log_len = log_len;
    fc.depth = depth;
    return fc;
}
struct FoldJob {
    const uint32_t* layer;
    uint32_t* next;
    // [m/4] Omega_m^-brev(i) over the layer-0 quads (Montgomery); layer d
    // uses its first m/4^(d+1) entries: brev_k(i) 4^d = brev_{k+2d}(i).
    const uint32_t* finv;
    const FoldConsts* c;  // device memory: the challenges come from the transcript
    int log_len;          // = c->log_len, for the grid
    int base_field;
    uint32_t block_begin;
    // Non-null: also hash layer d+1's quad leaves into level 0 of this tree
    // (parents into level 1 when with_parent), eight outputs per thread.
    uint32_t* tree = nullptr;
    int with_parent = 0;
    int sub_levels = 0;  // levels above level 1 reduced in LDS per workgroup
    // Non-null (with tree; also when no counter is needed): the whole
    // commitment of layer d+1 in one launch
    // (k_fold_commit): one output per thread, the quad leaves, six levels per
    // workgroup, then levels above by the last workgroup of each group of 64
    // (FOLD_COMMIT_GROUP) up to level `target`, the cap level. The counters
    // (tree_stage_counters(leaves, 0, target) of them) are zero on entry and
    // are left zero.
    uint32_t* counters = nullptr;
    int target = 0;
};
constexpr int FOLD_COMMIT_GROUP_LOG = 6;  // digests a workgroup reduces per stage: 2^6
// Words per stage counter (1: adjacent; padding each counter to 256 bytes
// measured no faster).
constexpr uint32_t STAGE_COUNTER_WORDS = 1;
// Words of counters a staged tree top uses from level `lvl` (count digests
// there, 64 per workgroup, each reducing up to six levels) to level
// `target`: one counter per group of 64 digests at each stage above the
// workgroups' own levels. A one-launch layer commit starts at its leaves
// (lvl 0).
inline uint64_t tree_stage_counters(uint64_t count, int lvl, int target) {
    uint64_t total = 0;
    int own = target - lvl < FOLD_COMMIT_GROUP_LOG ? target - lvl : FOLD_COMMIT_GROUP_LOG;
    lvl += own;
    count >>= own;
    while (lvl < target) {
        total += (count + 63) / 64;
        const int nl = target - lvl < FOLD_COMMIT_GROUP_LOG ? target - lvl : FOLD_COMMIT_GROUP_LOG;
        lvl += nl;
        count >>= nl;
    }
    return total * STAGE_COUNTER_WORDS;
}

// Clears `words` words of device memory (a proof's staged-top counters,
// before its first tree): a kernel, not a memset, so that captured launch
// sequences hold kernels only.
struct ZeroJob {
    uint32_t* dst;
    uint64_t words;
    uint32_t block_begin;
};

// Tree levels above `lvl` up to `target` in one launch (k_tree_top): 64
// digests of level lvl per workgroup, then the staged top; the counters
// (tree_stage_counters of them) are zero on entry and are left zero.
struct TreeTopJob {
    uint32_t* tree;  // level-contiguous, level 0 has `leaves` digests
    uint64_t leaves;
    int lvl, target;
    uint32_t* counters;
    uint32_t block_begin;
};

// Fold of layer d into a small layer d+1 plus its whole Merkle commitment
// (quad leaves and parent levels up to the cap) in one workgroup.
constexpr uint32_t SMALL_MAX_VALUES = 1024;  // values of layer d+1 (one compute unit)
struct SmallCommitJob {
    const uint32_t* layer;  // layer d
    uint32_t* next;         // layer d+1 (Ext4)
    const uint32_t* finv;
    const FoldConsts* c;
    uint32_t* tree;         // layer d+1's tree, level-contiguous
    uint32_t leaves;        // = values / 4 (<= SMALL_MAX_VALUES / 4)
    int target;             // levels above the leaves to build
    int base_field;
    uint32_t block_begin;
};
void launch_small_commit(const Launch& L, std::vector<SmallCommitJob> jobs);

// Proof-of-work preimage head || u64 nonce (one BLAKE3 chunk, at most 1024
// bytes): the blocks before the one the nonce starts in do not depend on the
// nonce, so the host compresses them into `cv`; the device compresses the
// one or two tail blocks per nonce.
struct GrindHead {
    uint32_t cv[8];     // chaining value entering the first tail block
    uint32_t tail[32];  // tail blocks as LE words, the nonce bytes zero
};
constexpr uint32_t GRIND_MAX_PREIMAGE = 1024;

struct GrindJob {
    const GrindHead* head;  // device memory
    uint32_t bits;
    uint32_t nblocks;
    uint64_t start, limit;  // nonces scanned: [start, limit)
    unsigned long long* result;  // [0] running minimum, [1] pooled chunk counter
    uint32_t nonce_off;     // byte offset of the nonce in the tail (< 64)
    uint32_t len0, flags0;  // first tail block: byte length and flags
    uint32_t len1;          // second tail block's byte length (0: one tail block)
    uint32_t block_begin;
};

// Head (BLAKE3(history) || u64 len || label) of `head_len` bytes: fills `h`
// and the job's tail fields; false when head || nonce exceeds one chunk.
bool grind_prepare(const uint8_t* head, size_t head_len, GrindHead& h, GrindJob& j);

// Device buffer of a transcript grind: the GrindHead, then the running
// minimum nonce at GRIND_BUF_RESULT and the pooled search's chunk counter
// after it (GrindJob::result points at both).
constexpr size_t GRIND_BUF_RESULT = 192, GRIND_BUF_BYTES = 256;
static_assert(sizeof(GrindHead) <= GRIND_BUF_RESULT && GRIND_BUF_RESULT + 16 <= GRIND_BUF_BYTES, "grind buffer layout");

// The allocation-free grind's static device array (GRIND_BUF_BYTES per
// device): noalloc_scratch_init looks up its address on `device` (which must
// be current) once per device; noalloc_scratch returns it (null before).
constexpr int NOALLOC_MAX_DEVICES = 64;
void noalloc_scratch_init(int device);
uint8_t* noalloc_scratch(int device);

// ---- launchers ---------------------------------------------------------------
void launch_fib_coeffs(const Launch& L, std::vector<FibCoeffJob> jobs);
void launch_ntt(const Launch& L, std::vector<NttJob> jobs);
void launch_linked(const Launch& L, std::vector<LinkedJob> jobs);
void launch_compose(const Launch& L, std::vector<ComposeJob> jobs);
void launch_leaves(const Launch& L, std::vector<LeavesJob> jobs);
void launch_levels(const Launch& L, std::vector<LevelsJob> jobs);
void launch_tail(const Launch& L, std::vector<TailJob> jobs);
void launch_tree_top(const Launch& L, std::vector<TreeTopJob> jobs);
void launch_zero(const Launch& L, std::vector<ZeroJob> jobs);
void launch_fold(const Launch& L, std::vector<FoldJob> jobs);
// Initialises every job's result to ~0, then searches.
void launch_grind(const Launch& L, std::vector<GrindJob> jobs);
// out[i] = in[brev(i)] mod p: raw coefficients to the bit-reversed,
// canonical order the coset LDE consumes.
// out[i * width + k] = in[brev(i) * in_stride + k] mod p, i < 2^logn, k < width:
// base values (1, 1), one coordinate of Ext4 values (4, 1), Ext4 values (4, 4).
struct BitrevJob {
    const uint32_t* in;
    uint32_t* out;
    int logn;
    uint32_t block_begin;
    int in_stride = 1;
    int width = 1;
};
void launch_bitrev(const Launch& L, std::vector<BitrevJob> jobs);

// Ext4 values from four coordinate planes: out[4i + c] = planes[c * count + i].
struct InterleaveJob {
    const uint32_t* planes;
    uint32_t* out;
    uint64_t count;
    uint32_t block_begin;
};
void launch_interleave(const Launch& L, std::vector<InterleaveJob> jobs);
// Copies of small regions (caps, final layers, nonces of many proofs) into
// the host-visible staging area, by one workgroup, which then stores `seq`
// to `flag` (if not null) after a system-scope fence: the host waits on the
// flag instead of synchronising the stream.
struct CollectItem {
    const uint32_t* src;
    uint32_t* dst;
    uint32_t words;
    uint32_t first;  // words of the items before this one (set by the launcher)
};
// Optional gate run by the same kernel after the flag: wait (bounded by
// spin_limit polls) until the host-mapped *go reaches `want`, then copy
// `words` constants from host-mapped src to device dst; a wait that
// expires sets *timeout. This is how work enqueued ahead of its transcript
// inputs is released (go == nullptr: no gate).
struct HostGate {
    const uint64_t* go;
    uint64_t want;
    const uint32_t* src;
    uint32_t* dst;
    uint32_t words;
    uint32_t spin_limit;
    uint32_t* timeout;
};
void launch_collect(const Launch& L, std::vector<CollectItem> items, uint64_t* flag, uint64_t seq,
                    const HostGate& gate);
void launch_copy(const Launch& L, const uint32_t* in, uint32_t* out, uint64_t words);

// Batched verification's Merkle path checks: `count` records of
// PATH_CHECK_WORDS words (16 leaf words, 8 expected digest words, steps,
// first sibling, direction bits low and high), siblings as 8 words each;
// ok[i] = 1 when record i's path reaches its expected digest, else 0.
constexpr int PATH_CHECK_WORDS = 28;
void launch_verify_paths(hipStream_t s, const uint32_t* checks, const uint32_t* siblings, uint32_t count,
                         uint8_t* ok);

// out[i] = w^(brev_bits(i) << shift) for i < 2^bits, w's powers from `pt`
// (Montgomery form): the bit-reversed point tables of a geometry.
void launch_pow_brev(const Launch& L, PowTable pt, uint32_t* out, int bits, int shift);

// NTT job builders: the passes of one inverse / coset LDE. pass2 is empty
// for single-pass sizes; otherwise pass1 must complete before pass2.
struct NttPlan {
    std::vector<NttJob> pass1, pass2;
};
NttPlan plan_intt(const uint32_t* evals, uint32_t* coeffs_br, uint32_t* scratch, int logn, const uint32_t* tw4096_inv,
                  PowTable winv_m, int logb, uint32_t ninv_m);
NttPlan plan_lde(const uint32_t* coeffs_br, uint32_t* lde_br, uint32_t* scratch, int logn, int logb,
                 const uint32_t* tw4096, PowTable w_m, CosetTables ct);

// A level of at most this many digests is reduced by the LDS tail kernel
// (latency-bound: a quad of lanes per node); wider levels use the
// one-node-per-thread kernel (throughput-bound).
constexpr int TAIL_GROUP_MIN_LOG = 2;
constexpr int TAIL_GROUP_MAX_LOG = 9;  // 512 digests, 1024 threads

}  // namespace z0p

namespace z0p {

// Openings of one proof, query q written to out + q * words_per_query:
// for each FRI layer d, the opened quad (16 words; 4 for a base-field layer)
// then target[d] sibling digests (8 words each, leaf level first); for AIR
// proofs then the two 16-word trace leaves of natural rows idx and idx + B
// (8 rows of A||B each) and their trace_target-step paths, where idx is the
// bit-reversed query index (air.rs:567-591).
constexpr int MAX_OPEN_LAYERS = 14;
struct OpenJob {
    const uint32_t* layer[MAX_OPEN_LAYERS];
    const uint32_t* tree[MAX_OPEN_LAYERS];
    uint64_t tree_leaves[MAX_OPEN_LAYERS];
    int target[MAX_OPEN_LAYERS];
    int base[MAX_OPEN_LAYERS];
    int nlayers;
    // AIR trace (a == nullptr for a FRI-only proof)
    const uint32_t* a;
    const uint32_t* b;
    const uint32_t* trace_tree;
    uint64_t trace_leaves;
    int trace_target;
    int logn, logb;
    // Range AIR rows (cols != nullptr, instead of a/b): ncols words at the
    // query's position, then the row tree path (trace_tree, one leaf per row).
    const uint32_t* cols = nullptr;
    int ncols = 0;
    uint64_t col_stride = 0;
    const uint64_t* queries;
    uint32_t nq;
    uint32_t words_per_query;
    uint32_t* out;
    uint32_t block_begin;
    // Non-null: the job was enqueued before its queries were known; it runs
    // only if *live (written with the queries) is non-zero.
    const uint64_t* live = nullptr;
};
void launch_open(const Launch& L, std::vector<OpenJob> jobs);

}  // namespace z0p
