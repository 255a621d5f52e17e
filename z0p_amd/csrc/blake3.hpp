// BLAKE3-256 compression for host and device, plus the host streaming hasher
// used by the Fiat-Shamir transcript.
//
// BLAKE3 is the reference's hash (src/hash.rs); it is a published function,
// so this is written from the BLAKE3 specification: 7 rounds, message
// schedule from the fixed permutation, flags CHUNK_START=1, CHUNK_END=2,
// PARENT=4, ROOT=8. The digest-producing paths on the proving path use only
// the first 8 output words (chaining value), so `compress8` returns those.
#pragma once
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>

#include "field.hpp"

namespace z0p {

namespace b3 {
constexpr uint32_t CHUNK_START = 1, CHUNK_END = 2, PARENT = 4, ROOT = 8;
constexpr size_t BLOCK_LEN = 64, CHUNK_LEN = 1024;
constexpr uint32_t IV0 = 0x6A09E667u, IV1 = 0xBB67AE85u, IV2 = 0x3C6EF372u, IV3 = 0xA54FF53Au,
                   IV4 = 0x510E527Fu, IV5 = 0x9B05688Cu, IV6 = 0x1F83D9ABu, IV7 = 0x5BE0CD19u;
}  // namespace b3

Z0P_HD uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

// Message word index used by round r at slot i: sigma^r(i), where sigma is the
// BLAKE3 permutation [2,6,3,10,7,0,4,13,1,11,12,5,9,14,15,8].
struct B3Schedule {
    uint8_t idx[7][16];
    constexpr B3Schedule() : idx{} {
        constexpr uint8_t perm[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
        for (int i = 0; i < 16; ++i) idx[0][i] = uint8_t(i);
        for (int r = 1; r < 7; ++r)
            for (int i = 0; i < 16; ++i) idx[r][i] = idx[r - 1][perm[i]];
    }
};

#define Z0P_B3_G(a, b, c, d, x, y)   \
    a = a + b + (x);                 \
    d = rotr(d ^ a, 16);             \
    c = c + d;                       \
    b = rotr(b ^ c, 12);             \
    a = a + b + (y);                 \
    d = rotr(d ^ a, 8);              \
    c = c + d;                       \
    b = rotr(b ^ c, 7);

// Compress one block; writes the 8-word chaining value to out (may alias cv).
Z0P_HD void compress8(const uint32_t cv[8], const uint32_t m[16], uint32_t counter_lo, uint32_t counter_hi,
                      uint32_t block_len, uint32_t flags, uint32_t out[8]) {
    constexpr B3Schedule S;
    uint32_t s0 = cv[0], s1 = cv[1], s2 = cv[2], s3 = cv[3], s4 = cv[4], s5 = cv[5], s6 = cv[6], s7 = cv[7];
    uint32_t s8 = b3::IV0, s9 = b3::IV1, s10 = b3::IV2, s11 = b3::IV3;
    uint32_t s12 = counter_lo, s13 = counter_hi, s14 = block_len, s15 = flags;
#pragma unroll
    for (int r = 0; r < 7; ++r) {
        Z0P_B3_G(s0, s4, s8, s12, m[S.idx[r][0]], m[S.idx[r][1]]);
        Z0P_B3_G(s1, s5, s9, s13, m[S.idx[r][2]], m[S.idx[r][3]]);
        Z0P_B3_G(s2, s6, s10, s14, m[S.idx[r][4]], m[S.idx[r][5]]);
        Z0P_B3_G(s3, s7, s11, s15, m[S.idx[r][6]], m[S.idx[r][7]]);
        Z0P_B3_G(s0, s5, s10, s15, m[S.idx[r][8]], m[S.idx[r][9]]);
        Z0P_B3_G(s1, s6, s11, s12, m[S.idx[r][10]], m[S.idx[r][11]]);
        Z0P_B3_G(s2, s7, s8, s13, m[S.idx[r][12]], m[S.idx[r][13]]);
        Z0P_B3_G(s3, s4, s9, s14, m[S.idx[r][14]], m[S.idx[r][15]]);
    }
    out[0] = s0 ^ s8;
    out[1] = s1 ^ s9;
    out[2] = s2 ^ s10;
    out[3] = s3 ^ s11;
    out[4] = s4 ^ s12;
    out[5] = s5 ^ s13;
    out[6] = s6 ^ s14;
    out[7] = s7 ^ s15;
}

Z0P_HD void b3_iv(uint32_t cv[8]) {
    cv[0] = b3::IV0; cv[1] = b3::IV1; cv[2] = b3::IV2; cv[3] = b3::IV3;
    cv[4] = b3::IV4; cv[5] = b3::IV5; cv[6] = b3::IV6; cv[7] = b3::IV7;
}

// Digest of one full 64-byte block hashed as a single-chunk BLAKE3 input
// (flags START|END|ROOT): the 64-byte Merkle leaf (hash.rs:388-397).
Z0P_HD void leaf64_words(const uint32_t m[16], uint32_t out[8]) {
    uint32_t iv[8];
    b3_iv(iv);
    compress8(iv, m, 0, 0, 64, b3::CHUNK_START | b3::CHUNK_END | b3::ROOT, out);
}

// Internal node BLAKE3(0x01 || left || right), 65 bytes = one full block with
// CHUNK_START then a one-byte block with CHUNK_END|ROOT (hash.rs:417-443).
// Digests are 8 little-endian words.
Z0P_HD void node_words(const uint32_t l[8], const uint32_t r[8], uint32_t out[8]) {
    uint32_t m[16];
    m[0] = 0x01u | (l[0] << 8);
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (int k = 1; k < 8; ++k) m[k] = (l[k - 1] >> 24) | (l[k] << 8);
    m[8] = (l[7] >> 24) | (r[0] << 8);
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (int k = 9; k < 16; ++k) m[k] = (r[k - 9] >> 24) | (r[k - 8] << 8);
    uint32_t iv[8], cv[8];
    b3_iv(iv);
    compress8(iv, m, 0, 0, 64, b3::CHUNK_START, cv);
    uint32_t m2[16] = {r[7] >> 24, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    compress8(cv, m2, 0, 0, 1, b3::CHUNK_END | b3::ROOT, out);
}

inline uint32_t load_le32(const uint8_t* p) {
    return uint32_t(p[0]) | (uint32_t(p[1]) << 8) | (uint32_t(p[2]) << 16) | (uint32_t(p[3]) << 24);
}
inline void store_le32(uint8_t* p, uint32_t x) {
    p[0] = uint8_t(x); p[1] = uint8_t(x >> 8); p[2] = uint8_t(x >> 16); p[3] = uint8_t(x >> 24);
}
inline void store_le64(uint8_t* p, uint64_t x) {
    for (int i = 0; i < 8; ++i) p[i] = uint8_t(x >> (8 * i));
}
inline uint64_t load_le64(const uint8_t* p) {
    uint64_t x = 0;
    for (int i = 7; i >= 0; --i) x = (x << 8) | p[i];
    return x;
}

// Host streaming BLAKE3 (unkeyed, 32-byte output): chunk state plus a stack
// of completed-subtree chaining values, merged whenever the chunk count
// allows. A full chunk is only folded into the stack once more input
// arrives, so the final chunk is always finalized with ROOT semantics.
class Blake3 {
   public:
    Blake3() { reset(); }

    void reset() {
        b3_iv(chunk_cv_);
        chunk_counter_ = 0;
        blocks_done_ = 0;
        buf_len_ = 0;
        stack_len_ = 0;
    }

    void update(const uint8_t* data, size_t n) {
        while (n > 0) {
            if (chunk_len() == b3::CHUNK_LEN) {
                uint32_t cv[8];
                chunk_finish(cv);
                push_cv(cv, chunk_counter_ + 1);
                ++chunk_counter_;
                b3_iv(chunk_cv_);
                blocks_done_ = 0;
                buf_len_ = 0;
            }
            if (buf_len_ == b3::BLOCK_LEN) {
                // Buffered block is not the chunk's last (more input follows).
                uint32_t m[16];
                words(buf_, m);
                uint32_t flags = blocks_done_ == 0 ? b3::CHUNK_START : 0;
                compress8(chunk_cv_, m, uint32_t(chunk_counter_), uint32_t(chunk_counter_ >> 32), 64, flags,
                          chunk_cv_);
                ++blocks_done_;
                buf_len_ = 0;
            }
            size_t take = b3::BLOCK_LEN - buf_len_;
            if (take > n) take = n;
            std::memcpy(buf_ + buf_len_, data, take);
            buf_len_ += take;
            data += take;
            n -= take;
        }
    }
    void update(const std::vector<uint8_t>& v) { update(v.data(), v.size()); }

    // Copies this state into `o` for a side computation, moving only the
    // live entries of the subtree stack.
    void fork_into(Blake3& o) const {
        std::memcpy(o.chunk_cv_, chunk_cv_, sizeof chunk_cv_);
        o.chunk_counter_ = chunk_counter_;
        o.blocks_done_ = blocks_done_;
        std::memcpy(o.buf_, buf_, buf_len_);
        o.buf_len_ = buf_len_;
        std::memcpy(o.stack_, stack_, 32 * stack_len_);
        o.stack_len_ = stack_len_;
    }

    // Digest of everything absorbed so far; the state stays usable.
    void digest(uint8_t out[32]) const {
        uint32_t m[16];
        uint8_t block[64] = {0};
        std::memcpy(block, buf_, buf_len_);
        words(block, m);
        uint32_t flags = b3::CHUNK_END | (blocks_done_ == 0 ? b3::CHUNK_START : 0);
        uint32_t cv[8];
        if (stack_len_ == 0) {
            compress8(chunk_cv_, m, uint32_t(chunk_counter_), uint32_t(chunk_counter_ >> 32), uint32_t(buf_len_),
                      flags | b3::ROOT, cv);
        } else {
            compress8(chunk_cv_, m, uint32_t(chunk_counter_), uint32_t(chunk_counter_ >> 32), uint32_t(buf_len_),
                      flags, cv);
            for (size_t i = stack_len_; i-- > 0;) {
                uint32_t pm[16];
                for (int k = 0; k < 8; ++k) {
                    pm[k] = stack_[i][k];
                    pm[8 + k] = cv[k];
                }
                uint32_t iv[8];
                b3_iv(iv);
                compress8(iv, pm, 0, 0, 64, b3::PARENT | (i == 0 ? b3::ROOT : 0), cv);
            }
        }
        for (int k = 0; k < 8; ++k) store_le32(out + 4 * k, cv[k]);
    }

   private:
    static void words(const uint8_t* b, uint32_t m[16]) {
        for (int k = 0; k < 16; ++k) m[k] = load_le32(b + 4 * k);
    }
    size_t chunk_len() const { return blocks_done_ * b3::BLOCK_LEN + buf_len_; }
    void chunk_finish(uint32_t cv[8]) {
        uint32_t m[16];
        words(buf_, m);
        uint32_t flags = b3::CHUNK_END | (blocks_done_ == 0 ? b3::CHUNK_START : 0);
        compress8(chunk_cv_, m, uint32_t(chunk_counter_), uint32_t(chunk_counter_ >> 32), uint32_t(buf_len_), flags,
                  cv);
    }
    // Merge completed subtrees: one merge per trailing zero bit of the new
    // total chunk count.
    void push_cv(const uint32_t cv_in[8], uint64_t total_chunks) {
        uint32_t cv[8];
        std::memcpy(cv, cv_in, 32);
        while ((total_chunks & 1) == 0) {
            --stack_len_;
            uint32_t pm[16];
            for (int k = 0; k < 8; ++k) {
                pm[k] = stack_[stack_len_][k];
                pm[8 + k] = cv[k];
            }
            uint32_t iv[8];
            b3_iv(iv);
            compress8(iv, pm, 0, 0, 64, b3::PARENT, cv);
            total_chunks >>= 1;
        }
        std::memcpy(stack_[stack_len_], cv, 32);
        ++stack_len_;
    }

    uint32_t chunk_cv_[8];
    uint64_t chunk_counter_;
    size_t blocks_done_;
    uint8_t buf_[64];
    size_t buf_len_;
    uint32_t stack_[54][8];
    size_t stack_len_;
};

inline void blake3_hash(const uint8_t* data, size_t n, uint8_t out[32]) {
    Blake3 h;
    h.update(data, n);
    h.digest(out);
}

}  // namespace z0p
