// ORACLE — test infrastructure only (see bb.hpp).
//
// Binary BLAKE3 Merkle tree, restated from src/merkle.rs.
#pragma once
#include <utility>
#include <vector>

#include "blake3.hpp"
#include "par.hpp"

namespace oracle {

using PathStep = std::pair<Digest, bool>;  // (sibling, current_is_left)
using Path = std::vector<PathStep>;

inline Digest hash_leaf(const uint8_t* data, size_t n) { return leaf_digest(data, n); }
inline Digest hash_node(const Digest& l, const Digest& r) { return node_digest(l, r); }

// merkle.rs:359-376 — one level, parallel over parents (min chunk 1024).
inline std::vector<Digest> hash_level(const std::vector<Digest>& prev) {
    std::vector<Digest> next(prev.size() / 2);
    for_each_indexed(next.data(), next.size(), 1024, [&](size_t base, Digest* piece, size_t n) {
        for (size_t j = 0; j < n; ++j) piece[j] = node_digest(prev[2 * (base + j)], prev[2 * (base + j) + 1]);
    });
    return next;
}

struct MerkleTree {
    std::vector<std::vector<Digest>> layers;

    // merkle.rs:350-361 — pad by duplicating the last leaf, then hash up.
    static MerkleTree finish(std::vector<Digest> current) {
        size_t target = 1;
        while (target < current.size()) target <<= 1;
        while (current.size() < target) current.push_back(current.back());
        MerkleTree t;
        t.layers.push_back(std::move(current));
        while (t.layers.back().size() > 1) t.layers.push_back(hash_level(t.layers.back()));
        return t;
    }

    // merkle.rs:93-110
    static MerkleTree from_flat64(const uint8_t* flat, size_t len) {
        if (len == 0) throw Z0pError(ERR_EMPTY_INPUT);
        if (len % 64 != 0) throw Z0pError(ERR_MALFORMED_INPUT, "64-byte leaves must tile the buffer exactly");
        size_t count = len / 64;
        std::vector<Digest> cur(count);
        for_each_indexed(cur.data(), count, 1024, [&](size_t base, Digest* piece, size_t n) {
            for (size_t j = 0; j < n; ++j) piece[j] = leaf64_digest(flat + 64 * (base + j));
        });
        return finish(std::move(cur));
    }

    // merkle.rs:62-79
    static MerkleTree from_flat(const uint8_t* flat, size_t len, size_t leaf_len) {
        if (len == 0 || leaf_len == 0) throw Z0pError(ERR_EMPTY_INPUT);
        if (len % leaf_len != 0) throw Z0pError(ERR_MALFORMED_INPUT, "flat leaves must tile the buffer exactly");
        size_t count = len / leaf_len;
        std::vector<Digest> cur(count);
        for_each_indexed(cur.data(), count, 1024, [&](size_t base, Digest* piece, size_t n) {
            for (size_t j = 0; j < n; ++j) piece[j] = leaf_digest(flat + leaf_len * (base + j), leaf_len);
        });
        return finish(std::move(cur));
    }

    Digest root() const { return layers.back()[0]; }
    size_t num_leaves() const { return layers[0].size(); }

    // merkle.rs:139-158
    Path prove(size_t index) const {
        if (index >= num_leaves()) throw Z0pError(ERR_INDEX_OUT_OF_BOUNDS);
        Path path;
        for (size_t l = 0; l + 1 < layers.size(); ++l) {
            bool is_left = index % 2 == 0;
            path.emplace_back(is_left ? layers[l][index + 1] : layers[l][index - 1], is_left);
            index /= 2;
        }
        return path;
    }

    // merkle.rs:229-232
    std::vector<Digest> cap(size_t dropped) const {
        size_t depth = layers.size() - 1;
        return layers[depth - std::min(dropped, depth)];
    }

    // merkle.rs:240-245
    Path prove_capped(size_t index, size_t drop_top) const {
        Path full = prove(index);
        size_t drop = std::min(drop_top, full.size());
        full.resize(full.size() - drop);
        return full;
    }
};

// merkle.rs:320-324
inline bool directions_match_index(size_t index, const Path& path) {
    for (size_t level = 0; level < path.size(); ++level)
        if (path[level].second != (((index >> level) & 1) == 0)) return false;
    return true;
}

// merkle.rs:254-277
inline bool verify_capped64(const Digest& cap_entry, const uint8_t* leaf, size_t index, size_t total_depth,
                            const Path& path) {
    if (total_depth >= 64 || path.size() > total_depth || (index >> total_depth) != 0 ||
        !directions_match_index(index, path))
        return false;
    Digest cur = leaf64_digest(leaf);
    for (auto& step : path) cur = step.second ? hash_node(cur, step.first) : hash_node(step.first, cur);
    return cur == cap_entry;
}

// merkle.rs:286-309 — a variable-length leaf (hashed as BLAKE3(0x00 || data)).
inline bool verify_capped(const Digest& cap_entry, const uint8_t* leaf, size_t leaf_len, size_t index,
                          size_t total_depth, const Path& path) {
    if (total_depth >= 64 || path.size() > total_depth || (index >> total_depth) != 0 ||
        !directions_match_index(index, path))
        return false;
    Digest cur = hash_leaf(leaf, leaf_len);
    for (auto& step : path) cur = step.second ? hash_node(cur, step.first) : hash_node(step.first, cur);
    return cur == cap_entry;
}

// merkle.rs:333-337 — (drop, depth).
inline std::pair<size_t, size_t> capped_shape(size_t num_leaves, size_t cap_height) {
    size_t depth = size_t(__builtin_ctzll(num_leaves));
    return {std::min(cap_height, depth), depth};
}

// merkle.rs:340-346
inline std::vector<uint8_t> cap_bytes(const std::vector<Digest>& cap) {
    std::vector<uint8_t> out;
    for (auto& d : cap) out.insert(out.end(), d.begin(), d.end());
    return out;
}

}  // namespace oracle
