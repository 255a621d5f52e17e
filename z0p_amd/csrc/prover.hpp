// Prover entry points behind the C ABI.
#pragma once
#include <cstdint>
#include <deque>
#include <functional>
#include <string>
#include <variant>
#include <vector>

#include "context.hpp"
#include "host.hpp"

namespace z0p {

Ctx* ctx_create(int device);

uint32_t brev_host(uint64_t x, int bits);

// A Merkle tree stored level-contiguously from the leaves up to the cap
// level `target = depth - drop` (levels above the cap are never needed).
struct TreeGeom {
    uint64_t leaves = 0;
    int depth = 0, drop = 0, target = 0;
    uint64_t words() const {
        uint64_t d = 0;
        for (int l = 0; l <= target; ++l) d += leaves >> l;
        return d * 8;
    }
    uint64_t level_offset_words(int l) const {
        uint64_t d = 0;
        for (int k = 0; k < l; ++k) d += leaves >> k;
        return d * 8;
    }
};
TreeGeom tree_geom(uint64_t leaves, uint64_t cap_height);

// One kernel step of one proof; a proof's steps of one iteration form a
// chain that must run in order.
using Step = std::variant<ZeroJob, FibCoeffJob, BitrevJob, InterleaveJob, NttJob, LinkedJob, ComposeJob, RangeComposeJob,
                          TableComposeJob, FoldCombineJob, LeavesJob, LevelsJob, TailJob, TreeTopJob, FoldJob,
                          SmallCommitJob, GrindJob, OpenJob>;

// Leaves, then wide levels, then the narrow levels up to the cap level: LDS
// tail launches, or with `counters` (tree_stage_counters(leaves, 0, target)
// zeroed words) and the schedule's tree_top, one staged launch.
void tree_steps(std::vector<Step>& chain, uint32_t* tree, const TreeGeom& g, LeavesJob leaves, const Schedule& s,
                uint32_t* counters = nullptr);
// The tree levels above `have` (levels 0..have already written) up to the cap.
void tree_levels(std::vector<Step>& chain, uint32_t* tree, const TreeGeom& g, int have, const Schedule& s,
                 uint32_t* counters = nullptr);

// Launches step r of every chain for r = 0, 1, ... (one launch per kernel
// type per r), then the collect items, which end by storing `seq` to
// `flag` (when not null) once every earlier write is visible to the host.
void launch_chains(const Launch& L, const std::vector<const std::vector<Step>*>& chains,
                   const std::vector<CollectItem>& collect, uint64_t* flag = nullptr, uint64_t seq = 0,
                   const HostGate& gate = HostGate{});

// One proof request and its state (defined in prover.cpp).
struct Proof;
Proof* proof_new_air(Ctx& c, const z0p_air_params& p, Transcript* tr);
Proof* proof_new_air_trace(Ctx& c, const z0p_air_params& p, const uint32_t* col_a, uint64_t len_a,
                          const uint32_t* col_b, uint64_t len_b, Transcript* tr);
Proof* proof_new_air_zk(Ctx& c, const z0p_air_params& p, uint64_t n_rand, ChaCha20& rng, Transcript* tr);
// range::prove / prove_with_columns / prove_zk (range.rs:334-428).
Proof* proof_new_range(Ctx& c, const z0p_range_params& p, Transcript* tr);
Proof* proof_new_range_cols(Ctx& c, const z0p_range_params& p, const uint32_t* cols, uint64_t ncols, uint64_t len,
                            Transcript* tr);
Proof* proof_new_range_zk(Ctx& c, const z0p_range_params& p, uint64_t n_rand, ChaCha20& rng, Transcript* tr);
// table::prove / prove_zk (table.rs:305-355): table_len >= 2 entries.
Proof* proof_new_table(Ctx& c, const z0p_table_params& p, const uint32_t* table, uint64_t table_len,
                       Transcript* tr);
Proof* proof_new_table_zk(Ctx& c, const z0p_table_params& p, const uint32_t* table, uint64_t table_len,
                          uint64_t n_rand, ChaCha20& rng, Transcript* tr);
// What a FRI commitment starts from (fri.rs:408-549).
enum FriInput : int {
    FRI_COEFFS = 0,     // commit: base-field coefficients
    FRI_COEFFS_EXT,     // commit_ext: Ext4 coefficients (4 words each)
    FRI_EVALS,          // commit_evals: base-field LDE values, natural order
    FRI_EVALS_EXT,      // commit_ext_evals: Ext4 LDE values, natural order
    FRI_FOLD_COMBINE,   // commit_ext_evals of the IVC steps' gamma-combination (fold.rs:494-524)
};
Proof* proof_new_fri(Ctx& c, const uint32_t* src, uint64_t count, int kind, uint64_t degree_bound,
                     const z0p_fri_params& fp, Transcript* tr);
// The batched FRI of fold::Accumulator::finalize: layer 0 is combined on the
// device from `nsteps` committed segments (device table), seg_steps rows
// each, extended by `blowup`; degree bound seg_steps - 1.
Proof* proof_new_fold_fri(Ctx& c, const FoldStepDesc* steps, uint32_t nsteps, uint64_t seg_steps, uint64_t blowup,
                          const z0p_fri_params& fp, Transcript* tr);
// The FRI query indices of a finished proof (layer-0 positions).
const std::vector<uint64_t>& proof_queries(const Proof* p);
void proof_free(Proof* p);
int proof_status(const Proof* p);
// True once the proof finished its last phase (its bytes are ready).
bool proof_done(const Proof* p);
Bytes& proof_bytes(Proof* p);
Bytes& proof_commitment(Proof* p);
size_t proof_mem_bytes(const Proof* p);
void proof_times(const Proof* p, double& submit_ms, double& admit_ms, double& done_ms);
void proof_set_tag(Proof* p, uint64_t tag);
uint64_t proof_tag(const Proof* p);

// The scheduler: admits queued proofs while pool memory allows (budget 0:
// one proof at a time, the pool grown to fit it) and advances all admitted
// proofs by one transcript phase per iteration.
class Engine {
   public:
    Engine(Ctx& c, size_t budget, size_t max_active);
    ~Engine();
    void submit(Proof* p);
    // One iteration; false once nothing is queued or active. Throws Status
    // on a device failure (call fail_all, then drop the engine).
    bool iterate();
    void run();
    // Ends every queued and active proof with `status`.
    void fail_all(int status);
    void on_done(std::function<void(Proof*)> f) { on_done_ = std::move(f); }
    // Latency mode: a proof alone in the scheduler has all its phases up to
    // the proof-of-work enqueued at once, each gated on the device by the
    // host's transcript step (no launch latency between phases).
    void set_pipelined(bool on) { pipelined_ = on; }
    size_t active() const { return active_.size(); }
    size_t queued() const { return queue_.size(); }
    uint64_t iterations() const { return iterations_; }

   private:
    void admit();
    void finish(Proof* p, int status);
    void pipeline(Proof& p);
    uint64_t launch(const std::vector<Proof*>& ws);
    void complete(const std::vector<Proof*>& ws, uint64_t seq, bool reset_ring);
    Ctx& c_;
    size_t budget_, max_active_;
    std::deque<Proof*> queue_;
    std::vector<Proof*> active_;
    std::vector<Proof*> flight_;  // launched wave awaiting its host steps
    uint64_t flight_seq_ = 0;
    int ring_half_ = 0;
    uint64_t iterations_ = 0;
    std::function<void(Proof*)> on_done_;
    bool pipelined_ = false;
};

// Budget for proofs needing `want` bytes in total: at most 80% of the device
// memory that is free (counting an idle pool as free).
size_t auto_budget(Ctx& c, size_t want);

// Autotuner result (tuner.cpp): the chosen schedule (also left set on the
// context), its time and the starting schedule's, and every trial as
// "knob=value:ms;".
struct TuneResult {
    Schedule best;
    double start_ms = 0, best_ms = 0;
    unsigned trials = 0;
    std::string log;
};
TuneResult autotune(Ctx& c, const std::vector<z0p_air_params>& work, unsigned reps);
// The C ABI's schedule struct as the scheduler's; false when a field is out
// of range (capi.cpp).
bool schedule_from_c(const z0p_schedule& in, Schedule& out);

Bytes air_prove(Ctx& ctx, const z0p_air_params& p, Transcript& tr);
Bytes air_prove_with_trace(Ctx& ctx, const z0p_air_params& p, const uint32_t* col_a, uint64_t len_a,
                                          const uint32_t* col_b, uint64_t len_b, Transcript& tr);
Bytes air_prove_zk(Ctx& ctx, const z0p_air_params& p, uint64_t n_rand, ChaCha20& rng, Transcript& tr);
// fri::commit* + public_commitment + open; `indices` receives the sampled
// leaf indices fri::open returns (fri.rs:645-666).
void fri_commit_open(Ctx& ctx, const uint32_t* src, uint64_t count, int kind, uint64_t degree_bound,
                     const z0p_fri_params& fp, Transcript& tr, Bytes& commitment, Bytes& proof,
                     std::vector<uint64_t>& indices);
// One proof alone through the scheduler (latency mode); throws its status.
void run_proof(Ctx& ctx, Proof* p);

// fri::commit* (fri.rs:408-549) and fri::open (fri.rs:645-666) as separate
// calls: the state holds its ctx and a proof whose device memory (layers and
// trees) is its own until fri_state_free.
struct FriState {
    Ctx* ctx = nullptr;
    Proof* proof = nullptr;
};
FriState* fri_commit(Ctx& ctx, const uint32_t* src, uint64_t count, int kind, uint64_t degree_bound,
                     const z0p_fri_params& fp, Transcript& tr);
// Commitment::to_bytes of fri::public_commitment (fri.rs:863-880).
const Bytes& fri_state_commitment(const FriState& s);
uint64_t fri_state_queries(const FriState& s);
// Proof::to_bytes; `indices` receives the sampled leaf indices.
Bytes fri_open(FriState& s, Transcript& tr, std::vector<uint64_t>& indices);
void fri_state_free(FriState* s);

// Transcript::grind (transcript.rs:206-228): the smallest nonce, searched on
// the device, then absorbed under `label`. `dev` is a device buffer of
// GRIND_BUF_BYTES (the head, then the running minimum at GRIND_BUF_RESULT);
// null uses the ctx's own buffer. Throws MalformedInput when head || nonce
// exceeds one BLAKE3 chunk (labels above 976 bytes). With `dev` given the
// call allocates nothing and records no profiling events.
uint64_t transcript_grind(Ctx& c, Transcript& tr, const char* label, uint32_t bits, uint8_t* dev = nullptr);
// The search alone: the smallest passing nonce of the prepared head.
uint64_t grind_search(Ctx& c, uint8_t* dev, const GrindHead& h, GrindJob j, uint32_t bits, bool profile = true);
Bytes range_prove(Ctx& ctx, const z0p_range_params& p, Transcript& tr);
Bytes range_prove_with_columns(Ctx& ctx, const z0p_range_params& p, const uint32_t* cols, uint64_t ncols,
                               uint64_t len, Transcript& tr);
Bytes range_prove_zk(Ctx& ctx, const z0p_range_params& p, uint64_t n_rand, ChaCha20& rng, Transcript& tr);
Bytes table_prove(Ctx& ctx, const z0p_table_params& p, const uint32_t* table, uint64_t table_len, Transcript& tr);
Bytes table_prove_zk(Ctx& ctx, const z0p_table_params& p, const uint32_t* table, uint64_t table_len, uint64_t n_rand,
                     ChaCha20& rng, Transcript& tr);

// Batched air::verify with the Merkle path hashing on the device
// (verify_batch.cpp); statuses[i] is what air_verify would return for
// instance i.
void air_verify_batch_device(Ctx& c, const z0p_air_params* params, Transcript* const* trs,
                             const uint8_t* const* proofs, const size_t* lens, size_t count, unsigned threads,
                             int* statuses);

// IVC accumulator (fold.rs:361-570; fold.cpp).
struct FoldAcc;
FoldAcc* fold_acc_new(Ctx& c, const z0p_fold_params& p);
void fold_acc_free(FoldAcc* a);
size_t fold_acc_len(const FoldAcc* a);
void fold_acc_add_segment(FoldAcc* a, const uint32_t* col_a, uint64_t len_a, const uint32_t* col_b, uint64_t len_b,
                          Transcript& tr);
Bytes fold_acc_finalize(FoldAcc* a, Transcript& tr);

}  // namespace z0p

// Opaque handles of the C ABI that hold device state.
struct z0p_ctx {
    z0p::Ctx* c;
    std::string tune_log;  // trials of the last z0p_ctx_autotune
};
struct z0p_fold_acc {
    z0p::FoldAcc* a = nullptr;
};
struct z0p_fri_state {
    z0p::FriState* s = nullptr;
};
