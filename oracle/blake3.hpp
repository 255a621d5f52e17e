// ORACLE — test infrastructure only (see bb.hpp).
//
// BLAKE3-256 (unkeyed), restated from src/hash.rs (scalar path).
#pragma once
#include <array>
#include <cstdint>
#include <cstring>
#include <vector>

#include "bb.hpp"

namespace oracle {

using Digest = std::array<uint8_t, 32>;

// hash.rs:16-25
static const uint32_t B3_IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                                  0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
// hash.rs:28
static const int B3_PERM[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
// hash.rs:31-37
constexpr uint32_t CHUNK_START = 1, CHUNK_END = 2, PARENT = 4, ROOT = 8;
constexpr size_t BLOCK_LEN = 64;
constexpr size_t CHUNK_LEN = 1024;
constexpr size_t MAX_STACK = 54;

inline uint32_t rotr32(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

// hash.rs:45-66
inline void quarter_round(uint32_t* s, int a, int b, int c, int d, uint32_t x, uint32_t y) {
    s[a] = s[a] + s[b] + x;
    s[d] = rotr32(s[d] ^ s[a], 16);
    s[c] = s[c] + s[d];
    s[b] = rotr32(s[b] ^ s[c], 12);
    s[a] = s[a] + s[b] + y;
    s[d] = rotr32(s[d] ^ s[a], 8);
    s[c] = s[c] + s[d];
    s[b] = rotr32(s[b] ^ s[c], 7);
}

// hash.rs:70-79
inline void b3_round(uint32_t* s, const uint32_t* m) {
    quarter_round(s, 0, 4, 8, 12, m[0], m[1]);
    quarter_round(s, 1, 5, 9, 13, m[2], m[3]);
    quarter_round(s, 2, 6, 10, 14, m[4], m[5]);
    quarter_round(s, 3, 7, 11, 15, m[6], m[7]);
    quarter_round(s, 0, 5, 10, 15, m[8], m[9]);
    quarter_round(s, 1, 6, 11, 12, m[10], m[11]);
    quarter_round(s, 2, 7, 8, 13, m[12], m[13]);
    quarter_round(s, 3, 4, 9, 14, m[14], m[15]);
}

// hash.rs:93-133 — 16 output words.
inline std::array<uint32_t, 16> compress(const uint32_t chaining[8], const uint32_t block[16], uint64_t counter,
                                         uint32_t block_len, uint32_t flags) {
    uint32_t s[16];
    for (int i = 0; i < 8; ++i) s[i] = chaining[i];
    for (int i = 0; i < 4; ++i) s[8 + i] = B3_IV[i];
    s[12] = uint32_t(counter);
    s[13] = uint32_t(counter >> 32);
    s[14] = block_len;
    s[15] = flags;
    uint32_t m[16];
    std::memcpy(m, block, sizeof(m));
    for (int r = 0; r < 7; ++r) {
        b3_round(s, m);
        if (r < 6) {
            uint32_t t[16];
            for (int i = 0; i < 16; ++i) t[i] = m[B3_PERM[i]];
            std::memcpy(m, t, sizeof(m));
        }
    }
    std::array<uint32_t, 16> out;
    for (int i = 0; i < 8; ++i) out[i] = s[i] ^ s[i + 8];
    for (int i = 0; i < 8; ++i) out[i + 8] = s[i + 8] ^ chaining[i];
    return out;
}

// hash.rs:136-142
inline Digest chaining_bytes(const uint32_t* out) {
    Digest d;
    for (int i = 0; i < 8; ++i) put_u32(d.data() + 4 * i, out[i]);
    return d;
}

// hash.rs:144-150
inline void words_from_block(const uint8_t* block, uint32_t w[16]) {
    for (int i = 0; i < 16; ++i) w[i] = get_u32(block + 4 * i);
}

inline void words_from_digest(const Digest& d, uint32_t w[8]) {
    for (int i = 0; i < 8; ++i) w[i] = get_u32(d.data() + 4 * i);
}

// hash.rs:156-168
inline Digest parent_cv(const Digest& l, const Digest& r, uint32_t extra_flags = 0) {
    uint8_t block[64];
    std::memcpy(block, l.data(), 32);
    std::memcpy(block + 32, r.data(), 32);
    uint32_t w[16];
    words_from_block(block, w);
    auto out = compress(B3_IV, w, 0, 64, PARENT | extra_flags);
    return chaining_bytes(out.data());
}
// hash.rs:173-185
inline Digest parent_root(const Digest& l, const Digest& r) { return parent_cv(l, r, ROOT); }

// hash.rs:189-225 — 1..=16 blocks of one chunk, last possibly partial.
inline Digest chunk_output(const uint8_t* data, size_t len, uint64_t counter, bool root) {
    uint32_t chaining[8];
    std::memcpy(chaining, B3_IV, sizeof(chaining));
    size_t block_count = (len + BLOCK_LEN - 1) / BLOCK_LEN;
    for (size_t i = 0; i < block_count; ++i) {
        const uint8_t* block = data + i * BLOCK_LEN;
        size_t blen = std::min(BLOCK_LEN, len - i * BLOCK_LEN);
        uint8_t padded[64] = {0};
        std::memcpy(padded, block, blen);
        uint32_t w[16];
        words_from_block(padded, w);
        uint32_t flags = 0;
        if (i == 0) flags |= CHUNK_START;
        if (i + 1 == block_count) {
            flags |= CHUNK_END;
            if (root) flags |= ROOT;
        }
        auto out = compress(chaining, w, counter, uint32_t(blen), flags);
        std::memcpy(chaining, out.data(), 32);
    }
    return chaining_bytes(chaining);
}

// hash.rs:233-324 — incremental hasher with a Bao chaining-value stack.
struct Hasher {
    std::vector<uint8_t> buf;
    size_t start = 0;
    Digest stack[MAX_STACK];
    size_t stack_len = 0;
    uint64_t next_index = 0;
    size_t total_len = 0;

    void update(const uint8_t* input, size_t n) {
        total_len += n;
        buf.insert(buf.end(), input, input + n);
        while (buf.size() - start >= 2 * CHUNK_LEN) {
            size_t end = start + CHUNK_LEN;
            Digest cv = chunk_output(buf.data() + start, CHUNK_LEN, next_index, false);
            push_merged(cv);
            start = end;
        }
        if (start >= 65536) {
            buf.erase(buf.begin(), buf.begin() + long(start));
            start = 0;
        }
    }
    void update(const std::vector<uint8_t>& v) { update(v.data(), v.size()); }

    Digest finalize() const {
        Hasher h = *this;
        return h.finalize_consume();
    }

    Digest finalize_consume() {
        if (total_len == 0) {
            uint32_t zero[16] = {0};
            auto out = compress(B3_IV, zero, 0, 0, CHUNK_START | CHUNK_END | ROOT);
            return chaining_bytes(out.data());
        }
        size_t tail_len = buf.size() - start;
        if (tail_len > CHUNK_LEN) {
            size_t end = start + CHUNK_LEN;
            Digest cv = chunk_output(buf.data() + start, CHUNK_LEN, next_index, false);
            push_merged(cv);
            start = end;
        }
        const uint8_t* partial = buf.data() + start;
        size_t plen = buf.size() - start;
        if (stack_len == 0) return chunk_output(partial, plen, next_index, true);
        Digest current = chunk_output(partial, plen, next_index, false);
        for (size_t i = stack_len - 1; i >= 1; --i) current = parent_cv(stack[i], current);
        return parent_root(stack[0], current);
    }

    void push_merged(const Digest& cv) {
        stack[stack_len++] = cv;
        next_index += 1;
        uint64_t total = next_index;
        while (total % 2 == 0) {
            Digest right = stack[stack_len - 1];
            Digest left = stack[stack_len - 2];
            stack_len -= 2;
            stack[stack_len++] = parent_cv(left, right);
            total /= 2;
        }
    }
};

// hash.rs:341-345
inline Digest blake3_256(const uint8_t* data, size_t n) {
    Hasher h;
    h.update(data, n);
    return h.finalize_consume();
}

// hash.rs:692-707
inline Digest blake3_256_small(const uint8_t* data, size_t n) {
    if (n == 0) {
        uint32_t zero[16] = {0};
        auto out = compress(B3_IV, zero, 0, 0, CHUNK_START | CHUNK_END | ROOT);
        return chaining_bytes(out.data());
    }
    if (n > CHUNK_LEN) return blake3_256(data, n);
    return chunk_output(data, n, 0, true);
}

// hash.rs:352-382 — BLAKE3(0x00 || data).
inline Digest leaf_digest(const uint8_t* data, size_t n) {
    std::vector<uint8_t> buf(n + 1);
    buf[0] = 0x00;
    std::memcpy(buf.data() + 1, data, n);
    if (n < CHUNK_LEN) return blake3_256_small(buf.data(), buf.size());
    return blake3_256(buf.data(), buf.size());
}

// hash.rs:388-397 — one full block, START|END|ROOT.
inline Digest leaf64_digest(const uint8_t* block) {
    uint32_t w[16];
    words_from_block(block, w);
    auto out = compress(B3_IV, w, 0, 64, CHUNK_START | CHUNK_END | ROOT);
    return chaining_bytes(out.data());
}

// hash.rs:417-443 — BLAKE3(0x01 || left || right) as two compressions.
inline Digest node_digest(const Digest& left, const Digest& right) {
    uint8_t first[64];
    first[0] = 0x01;
    std::memcpy(first + 1, left.data(), 32);
    std::memcpy(first + 33, right.data(), 31);
    uint32_t w[16];
    words_from_block(first, w);
    auto out = compress(B3_IV, w, 0, 64, CHUNK_START);
    uint32_t chaining[8];
    std::memcpy(chaining, out.data(), 32);
    uint8_t last[64] = {0};
    last[0] = right[31];
    words_from_block(last, w);
    auto out2 = compress(chaining, w, 0, 1, CHUNK_END | ROOT);
    return chaining_bytes(out2.data());
}

}  // namespace oracle
