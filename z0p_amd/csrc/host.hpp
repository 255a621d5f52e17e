// Host-only half of the library: the verifiers, parameter constructors,
// transcript and ChaCha20 handles, and the helpers the extern "C" entries
// share. Nothing here includes HIP, so these sources also build with a
// plain host compiler (the sanitizer builds of the verifier).
#pragma once
#include <cstdint>
#include <new>
#include <vector>

#include "rng.hpp"
#include "transcript.hpp"

namespace z0p {

// air::build_trace(a0, b0, n).1[n-1] by fast doubling (air.rs:341-354).
uint32_t fib_b_last(uint32_t a0, uint32_t b0, uint64_t n);

// Host verifier (air.rs:607-752, fri.rs:686-859); throws Status on rejection.
void air_verify(const z0p_air_params& p, Transcript& tr, const uint8_t* proof, size_t len);
// Returns the sampled query indices (fri.rs:686).
std::vector<uint64_t> fri_verify(const uint8_t* cm, size_t cm_len, const uint8_t* pf, size_t pf_len,
                                 const z0p_fri_params& fp, Transcript& tr);
void range_verify(const z0p_range_params& p, Transcript& tr, const uint8_t* proof, size_t len);
void fold_verify(const z0p_fold_params& p, Transcript& tr, const uint8_t* proof, size_t len);
void table_verify(const z0p_table_params& p, const uint32_t* table, uint64_t table_len, Transcript& tr,
                  const uint8_t* proof, size_t len);
// table.rs:357-363: the saturating bound |T| (n + n_rand - 1) - n.
inline uint64_t membership_bound(uint64_t table_len, uint64_t n, uint64_t n_rand) {
    const uint64_t rows = n_rand > UINT64_MAX - n ? UINT64_MAX : n + n_rand;
    const unsigned __int128 prod = (unsigned __int128)table_len * (rows == 0 ? 0 : rows - 1);
    const uint64_t sat = prod > UINT64_MAX ? UINT64_MAX : uint64_t(prod);
    return sat < n ? 0 : sat - n;
}

// Merkle path checks of a batched verification that run on the device: a
// 64-byte leaf hashed up `steps` levels (bit l of `dirs` set: the running
// node is the left child at level l, siblings[sib + l] the other child)
// must reach `expected`. Words are little-endian, as on the wire.
struct PathCheck {
    uint32_t leaf[16];
    uint32_t expected[8];
    uint32_t steps;
    uint32_t sib;
    uint64_t dirs;
};
struct PathChecks {
    std::vector<PathCheck> checks;
    std::vector<uint32_t> siblings;  // 8 words per digest
};
// air::verify with its Merkle path checks appended to `defer` instead of
// run (every other check runs here, in the reference's order). The proof is
// accepted iff this returns and every recorded check holds; if this throws,
// a recorded check that fails came first, so the proof's status is then
// BadMerklePath, else the thrown status.
void air_verify_deferred(const z0p_air_params& p, Transcript& tr, const uint8_t* proof, size_t len,
                         PathChecks& defer);

// Generic Merkle trees over byte leaves (merkle.rs:16-30, 165-235):
// leaf H(0x00 || data), node H(0x01 || left || right).
void merkle_hash_leaf(const uint8_t* data, size_t len, uint8_t out[32]);
void merkle_hash_node(const uint8_t left[32], const uint8_t right[32], uint8_t out[32]);
// MerkleTree::verify_digest: the path's direction bits must spell `index`
// (is_left at level l <=> bit l of index is 0) and index < 2^steps.
bool merkle_verify_digest(const uint8_t root[32], const uint8_t leaf_digest[32], uint64_t index,
                          const uint8_t* path, size_t steps);

namespace capi {

// Runs f, mapping a thrown Status to its code and anything else to a
// status (no exception crosses the C boundary).
template <typename F>
int guarded(F&& f) {
    try {
        f();
        return Z0P_OK;
    } catch (const Status& s) {
        return s.code;
    } catch (const std::bad_alloc&) {
        return Z0P_ERR_OUT_OF_MEMORY;
    } catch (...) {
        return Z0P_ERR_DEVICE;
    }
}

// Query-index outputs of the FRI entries: null, or room for every query.
inline bool indices_ok(const uint64_t* indices, size_t cap, const z0p_fri_params* params) {
    return !indices || cap >= params->num_queries;
}
inline void put_indices(const std::vector<uint64_t>& q, uint64_t* indices) {
    if (indices)
        for (size_t i = 0; i < q.size(); ++i) indices[i] = q[i];
}

}  // namespace capi
}  // namespace z0p

// Opaque handles of the C ABI that need no device.
struct z0p_transcript {
    z0p::Transcript t;
};
struct z0p_chacha20 {
    z0p::ChaCha20* r = nullptr;
};
