// Prover entry points behind the C ABI.
#pragma once
#include <cstdint>
#include <vector>

#include "context.hpp"
#include "transcript.hpp"

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
void build_levels(Ctx& ctx, uint32_t* tree, const TreeGeom& g, int have);

// air::build_trace(a0, b0, n).1[n-1] by fast doubling (air.rs:341-354).
uint32_t fib_b_last(uint32_t a0, uint32_t b0, uint64_t n);

std::vector<uint8_t> air_prove(Ctx& ctx, const z0p_air_params& p, Transcript& tr);
void fri_commit_open(Ctx& ctx, const uint32_t* coeffs, uint64_t n, const z0p_fri_params& fp, Transcript& tr,
                     std::vector<uint8_t>& commitment, std::vector<uint8_t>& proof);

// Host verifier (air.rs:607-752, fri.rs:686-859); throws Status on rejection.
void air_verify(const z0p_air_params& p, Transcript& tr, const uint8_t* proof, size_t len);
void fri_verify(const uint8_t* cm, size_t cm_len, const uint8_t* pf, size_t pf_len, const z0p_fri_params& fp,
                Transcript& tr);

}  // namespace z0p

// Opaque handles of the C ABI.
struct z0p_transcript {
    z0p::Transcript t;
};
struct z0p_ctx {
    z0p::Ctx* c;
};
