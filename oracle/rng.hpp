// ORACLE — test infrastructure only (see bb.hpp).
//
// ChaCha20 stream (RFC 8439) used as the blinding CSPRNG, restated from
// src/rng.rs (z0p v0.3.0).
#pragma once
#include <algorithm>
#include <cstdint>
#include <cstring>

#include "bb.hpp"

namespace oracle {

// rng.rs:21-34 — quarter round by state index.
inline void chacha_quarter(uint32_t s[16], int a, int b, int c, int d) {
    auto rotl = [](uint32_t x, int r) { return (x << r) | (x >> (32 - r)); };
    s[a] += s[b];
    s[d] = rotl(s[d] ^ s[a], 16);
    s[c] += s[d];
    s[b] = rotl(s[b] ^ s[c], 12);
    s[a] += s[b];
    s[d] = rotl(s[d] ^ s[a], 8);
    s[c] += s[d];
    s[b] = rotl(s[b] ^ s[c], 7);
}

inline uint32_t le32(const uint8_t* p) {
    return uint32_t(p[0]) | (uint32_t(p[1]) << 8) | (uint32_t(p[2]) << 16) | (uint32_t(p[3]) << 24);
}

// rng.rs:38-72 — one 64-byte block for (key, counter, nonce).
inline void chacha_block(const uint8_t key[32], uint32_t counter, const uint8_t nonce[12], uint8_t out[64]) {
    uint32_t s[16];
    s[0] = 0x61707865u;
    s[1] = 0x3320646eu;
    s[2] = 0x79622d32u;
    s[3] = 0x6b206574u;
    for (int i = 0; i < 8; ++i) s[4 + i] = le32(key + 4 * i);
    s[12] = counter;
    for (int i = 0; i < 3; ++i) s[13 + i] = le32(nonce + 4 * i);
    uint32_t init[16];
    std::memcpy(init, s, sizeof(s));
    for (int r = 0; r < 10; ++r) {
        chacha_quarter(s, 0, 4, 8, 12);
        chacha_quarter(s, 1, 5, 9, 13);
        chacha_quarter(s, 2, 6, 10, 14);
        chacha_quarter(s, 3, 7, 11, 15);
        chacha_quarter(s, 0, 5, 10, 15);
        chacha_quarter(s, 1, 6, 11, 12);
        chacha_quarter(s, 2, 7, 8, 13);
        chacha_quarter(s, 3, 4, 9, 14);
    }
    for (int i = 0; i < 16; ++i) {
        uint32_t w = s[i] + init[i];
        for (int b = 0; b < 4; ++b) out[4 * i + b] = uint8_t(w >> (8 * b));
    }
}

// rng.rs:82-168 — seekable stream: `counter` names the next block.
struct ChaCha20 {
    uint8_t key[32];
    uint8_t nonce[12];
    uint32_t counter = 0;
    uint8_t block[64] = {0};
    size_t used = 64;

    ChaCha20(const uint8_t k[32], const uint8_t n[12]) {
        std::memcpy(key, k, 32);
        std::memcpy(nonce, n, 12);
    }

    // rng.rs:118-132
    void fill_bytes(uint8_t* out, size_t len) {
        size_t done = 0;
        while (done < len) {
            if (used == 64) {
                chacha_block(key, counter, nonce, block);
                counter += 1;  // wrapping u32
                used = 0;
            }
            size_t take = std::min<size_t>(64 - used, len - done);
            std::memcpy(out + done, block + used, take);
            used += take;
            done += take;
        }
    }
    // rng.rs:135-148
    uint32_t next_u32() {
        uint8_t w[4];
        fill_bytes(w, 4);
        return le32(w);
    }
    uint64_t next_u64() {
        uint8_t w[8];
        fill_bytes(w, 8);
        return uint64_t(le32(w)) | (uint64_t(le32(w + 4)) << 32);
    }
    // rng.rs:157-168 — rejection sampling: draws at or above
    // floor(2^32 / p) * p are dropped.
    Field next_field() {
        const uint64_t p = Field::MODULUS;
        const uint64_t limit = ((uint64_t(1) << 32) / p) * p;
        while (true) {
            uint64_t w = next_u32();
            if (w < limit) return Field::raw(uint32_t(w % p));
        }
    }
};

}  // namespace oracle
