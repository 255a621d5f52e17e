// Protocol parameters: validation and the size/strength tuners.
//
// The tuners decide the production shape (queries, grinding, final layer,
// cap height) that is absorbed into the transcript and fixes the wire size,
// so they follow the reference formulas exactly:
//   fri.rs:963-975 (layer sizes), 983-999 (FRI byte estimate), 1007-1021
//   (column byte estimate), 1009-1044 (tune_shape), 1051-1057
//   (composition_millibits), 1066-1074 (tune_strength), 232-255 and
//   air.rs:129-162 (for_security).
#pragma once
#include <algorithm>
#include <cstdint>
#include <utility>
#include <vector>

#include "field.hpp"
#include "status.hpp"

namespace z0p {

inline bool pow2(uint64_t x) { return x != 0 && (x & (x - 1)) == 0; }
// Trailing zeros with tz(0) = 64, like usize::trailing_zeros on 64-bit
// targets (the final-size enumeration can reach an empty layer).
inline uint32_t ilog2(uint64_t x) { return x == 0 ? 64u : uint32_t(__builtin_ctzll(x)); }
constexpr uint64_t MAX_DOMAIN = uint64_t(1) << TWO_ADICITY;
constexpr uint64_t MAX_FINAL_SIZE = 1024;

// Estimated proof bytes for a FRI layer schedule at a given cap height.
inline uint64_t fri_bytes_estimate(uint64_t lde_len, uint64_t final_size, uint64_t queries, uint64_t cap) {
    uint64_t total = 32;  // three dims (24) + nonce (8)
    for (uint64_t len = lde_len;; len /= 4) {
        uint64_t leaves = len % 4 == 0 ? len / 4 : len;
        uint64_t depth = ilog2(leaves);
        uint64_t drop = std::min(cap, depth);
        total += 32ull << std::min<uint64_t>(drop, 16);
        bool last = len <= final_size;
        if (!last) total += queries * (64 + 33 * (depth - drop));
        if (last) break;
    }
    return total;
}

inline uint64_t column_bytes_estimate(uint64_t lde_len, uint64_t queries, uint64_t value_bytes, uint64_t cap) {
    uint64_t depth = ilog2(lde_len);
    uint64_t drop = std::min(cap, depth);
    return (32ull << std::min<uint64_t>(drop, 16)) + queries * (value_bytes + 33 * (depth - drop));
}

// (final_size, cap_height) with the smallest estimate; finals are lde/4^f
// capped at 32, caps 0..4; strict improvement keeps the earlier (larger
// final, smaller cap) candidate on ties.
inline std::pair<uint64_t, uint64_t> tune_shape(uint64_t lde_len, uint64_t queries, bool with_column,
                                                uint64_t col_lde, uint64_t col_value_bytes) {
    uint64_t fin = std::max<uint64_t>(lde_len / 4, 2);
    while (fin > 32) fin /= 4;
    std::pair<uint64_t, uint64_t> best{fin, 0};
    uint64_t best_bytes = ~uint64_t(0);
    while (true) {
        for (uint64_t cap = 0; cap <= 4; ++cap) {
            uint64_t t = fri_bytes_estimate(lde_len, fin, queries, cap);
            if (with_column) t += column_bytes_estimate(col_lde, queries, col_value_bytes, cap);
            if (t < best_bytes) {
                best_bytes = t;
                best = {fin, cap};
            }
        }
        if (fin <= 1) break;
        fin /= 4;
        if (fin == 0) break;
    }
    return best;
}

inline uint32_t composition_millibits(uint64_t lde_len, uint64_t degree_bound) {
    uint64_t den = 1;
    while (den < degree_bound + 1) den <<= 1;
    if (lde_len < den) return 500;
    return 500 * (ilog2(lde_len) - ilog2(den));
}

// (queries, grinding): grinding takes min(20, target/2) bits, queries cover
// the rest at the given milli-bits each, rounded up, at least one.
inline std::pair<uint64_t, uint32_t> tune_strength(uint32_t target_bits, uint32_t millibits) {
    uint32_t grinding = std::min<uint32_t>(20, target_bits / 2);
    uint64_t remaining = uint64_t(target_bits - grinding) * 1000;
    uint64_t per = std::max<uint32_t>(millibits, 1);
    uint64_t q = (remaining + per - 1) / per;
    q = std::min<uint64_t>(std::max<uint64_t>(q, 1), 1000000);
    return {q, grinding};
}

// fri::Params::validate (fri.rs:163-196).
inline void validate_fri(const z0p_fri_params& p) {
    require(p.blowup >= 2 && pow2(p.blowup), Z0P_ERR_INVALID_BLOWUP);
    require(p.num_queries != 0, Z0P_ERR_NO_QUERIES);
    require(p.num_queries <= (1u << 20), Z0P_ERR_MALFORMED_INPUT);
    require(p.final_size >= 2 && pow2(p.final_size) && p.final_size <= MAX_FINAL_SIZE, Z0P_ERR_INVALID_DOMAIN_SIZE);
    require(p.grinding_bits <= 32, Z0P_ERR_MALFORMED_INPUT);
    require(p.cap_height <= 8, Z0P_ERR_MALFORMED_INPUT);
}

// air::Params::validate (air.rs:181-211).
inline void validate_air(const z0p_air_params& p) {
    require(p.n_steps >= 2 && pow2(p.n_steps), Z0P_ERR_INVALID_DOMAIN_SIZE);
    require(p.blowup >= 4 && pow2(p.blowup), Z0P_ERR_INVALID_BLOWUP);
    require(p.num_queries != 0, Z0P_ERR_NO_QUERIES);
    require(p.final_size >= 2 && pow2(p.final_size), Z0P_ERR_INVALID_DOMAIN_SIZE);
    require(p.cap_height <= 8, Z0P_ERR_MALFORMED_INPUT);
    require(p.n_steps <= MAX_DOMAIN && p.blowup <= MAX_DOMAIN && p.n_steps * p.blowup <= MAX_DOMAIN,
            Z0P_ERR_INVALID_DOMAIN_SIZE);
}

// air::Params::fri_params (air.rs:170-176): fri::Params::new then builders.
inline z0p_fri_params air_fri_params(const z0p_air_params& a) {
    require(a.blowup >= 2 && pow2(a.blowup), Z0P_ERR_INVALID_BLOWUP);
    require(a.num_queries != 0, Z0P_ERR_NO_QUERIES);
    require(a.final_size >= 2 && pow2(a.final_size) && a.final_size <= MAX_FINAL_SIZE, Z0P_ERR_INVALID_DOMAIN_SIZE);
    z0p_fri_params f{};
    f.blowup = a.blowup;
    f.num_queries = a.num_queries;
    f.final_size = a.final_size;
    f.grinding_bits = a.grinding_bits;
    f.cap_height = a.cap_height;
    return f;
}

// range::Params::validate (range.rs:182-215).
inline void validate_range(const z0p_range_params& p) {
    require(p.n_rows >= 2 && pow2(p.n_rows), Z0P_ERR_INVALID_DOMAIN_SIZE);
    require(p.n_bits >= 1 && p.n_bits <= 24, Z0P_ERR_MALFORMED_INPUT);
    require(p.blowup >= 2 && pow2(p.blowup), Z0P_ERR_INVALID_BLOWUP);
    require(p.num_queries != 0, Z0P_ERR_NO_QUERIES);
    require(p.final_size >= 2 && pow2(p.final_size), Z0P_ERR_INVALID_DOMAIN_SIZE);
    require(p.cap_height <= 8, Z0P_ERR_MALFORMED_INPUT);
    require(p.n_rows <= MAX_DOMAIN && p.blowup <= MAX_DOMAIN && p.n_rows * p.blowup <= MAX_DOMAIN,
            Z0P_ERR_INVALID_DOMAIN_SIZE);
}

// range::Params::fri_params (range.rs:172-179).
inline z0p_fri_params range_fri_params(const z0p_range_params& a) {
    require(a.blowup >= 2 && pow2(a.blowup), Z0P_ERR_INVALID_BLOWUP);
    require(a.num_queries != 0, Z0P_ERR_NO_QUERIES);
    require(a.final_size >= 2 && pow2(a.final_size) && a.final_size <= MAX_FINAL_SIZE, Z0P_ERR_INVALID_DOMAIN_SIZE);
    z0p_fri_params f{};
    f.blowup = a.blowup;
    f.num_queries = a.num_queries;
    f.final_size = a.final_size;
    f.grinding_bits = a.grinding_bits;
    f.cap_height = a.cap_height;
    return f;
}

// table::Params::validate (table.rs:184-212).
inline void validate_table(const z0p_table_params& p) {
    require(p.n_rows >= 2 && pow2(p.n_rows), Z0P_ERR_INVALID_DOMAIN_SIZE);
    require(p.blowup >= 2 && pow2(p.blowup), Z0P_ERR_INVALID_BLOWUP);
    require(p.num_queries != 0, Z0P_ERR_NO_QUERIES);
    require(p.final_size >= 2 && pow2(p.final_size), Z0P_ERR_INVALID_DOMAIN_SIZE);
    require(p.cap_height <= 8, Z0P_ERR_MALFORMED_INPUT);
    require(p.n_rows <= MAX_DOMAIN && p.blowup <= MAX_DOMAIN && p.n_rows * p.blowup <= MAX_DOMAIN,
            Z0P_ERR_INVALID_DOMAIN_SIZE);
}

// table::Params::fri_params (table.rs:175-181).
inline z0p_fri_params table_fri_params(const z0p_table_params& a) {
    require(a.blowup >= 2 && pow2(a.blowup), Z0P_ERR_INVALID_BLOWUP);
    require(a.num_queries != 0, Z0P_ERR_NO_QUERIES);
    require(a.final_size >= 2 && pow2(a.final_size) && a.final_size <= MAX_FINAL_SIZE, Z0P_ERR_INVALID_DOMAIN_SIZE);
    z0p_fri_params f{};
    f.blowup = a.blowup;
    f.num_queries = a.num_queries;
    f.final_size = a.final_size;
    f.grinding_bits = a.grinding_bits;
    f.cap_height = a.cap_height;
    return f;
}

// fold::Params::validate (fold.rs:186-218).
inline void validate_fold(const z0p_fold_params& p) {
    require(p.seg_steps >= 2 && pow2(p.seg_steps), Z0P_ERR_INVALID_DOMAIN_SIZE);
    require(p.blowup >= 4 && pow2(p.blowup), Z0P_ERR_INVALID_BLOWUP);
    require(p.num_queries != 0, Z0P_ERR_NO_QUERIES);
    require(p.final_size >= 2 && pow2(p.final_size), Z0P_ERR_INVALID_DOMAIN_SIZE);
    require(p.cap_height <= 8, Z0P_ERR_MALFORMED_INPUT);
    require(p.seg_steps <= MAX_DOMAIN && p.blowup <= MAX_DOMAIN && p.seg_steps * p.blowup <= MAX_DOMAIN,
            Z0P_ERR_INVALID_DOMAIN_SIZE);
}

// fold::Params::fri_params (fold.rs:177-183).
inline z0p_fri_params fold_fri_params(const z0p_fold_params& a) {
    require(a.blowup >= 2 && pow2(a.blowup), Z0P_ERR_INVALID_BLOWUP);
    require(a.num_queries != 0, Z0P_ERR_NO_QUERIES);
    require(a.final_size >= 2 && pow2(a.final_size) && a.final_size <= MAX_FINAL_SIZE, Z0P_ERR_INVALID_DOMAIN_SIZE);
    z0p_fri_params f{};
    f.blowup = a.blowup;
    f.num_queries = a.num_queries;
    f.final_size = a.final_size;
    f.grinding_bits = a.grinding_bits;
    f.cap_height = a.cap_height;
    return f;
}

// FRI layer-shape checks shared by commit (fri.rs:556-569, 1125-1140).
inline size_t fri_num_folds(uint64_t lde_len, uint64_t final_size, uint64_t degree_bound) {
    require(final_size < lde_len && lde_len % final_size == 0, Z0P_ERR_INVALID_DOMAIN_SIZE);
    uint64_t ratio = lde_len / final_size;
    require(pow2(ratio) && ilog2(ratio) % 2 == 0, Z0P_ERR_INVALID_DOMAIN_SIZE);
    require(lde_len >= 2 && degree_bound + 1 <= lde_len / 2, Z0P_ERR_MALFORMED_INPUT);
    size_t folds = ilog2(ratio) / 2;
    uint64_t residual = degree_bound >> (2 * folds);
    require(residual + 1 < final_size, Z0P_ERR_MALFORMED_INPUT);
    return folds;
}

// Leaves of FRI layer d's tree: quads when the layer length is a multiple of
// four, otherwise one leaf per value (fri.rs:1401-1407).
inline uint64_t fri_tree_leaves(uint64_t layer_len) { return layer_len % 4 == 0 ? layer_len / 4 : layer_len; }
inline uint64_t cap_drop(uint64_t leaves, uint64_t cap_height) { return std::min<uint64_t>(cap_height, ilog2(leaves)); }

}  // namespace z0p
