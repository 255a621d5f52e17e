// ChaCha20 keystream (RFC 8439) as the blinding CSPRNG: the reference's
// rng::ChaCha20 (src/rng.rs:81-177) and its field sampler. Host only: a
// proof draws 2 * n_rand field elements, a few kilobytes of keystream.
#pragma once
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "field.hpp"
#include "status.hpp"

namespace z0p {

class ChaCha20 {
public:
    ChaCha20(const uint8_t key[32], const uint8_t nonce[12]) {
        state_[0] = 0x61707865u;  // "expand 32-byte k"
        state_[1] = 0x3320646eu;
        state_[2] = 0x79622d32u;
        state_[3] = 0x6b206574u;
        for (int i = 0; i < 8; ++i) state_[4 + i] = load(key + 4 * i);
        state_[12] = 0;
        for (int i = 0; i < 3; ++i) state_[13 + i] = load(nonce + 4 * i);
    }
    ~ChaCha20() {
        // best-effort wipe of key material and buffered keystream
        volatile uint32_t* s = state_;
        for (int i = 0; i < 16; ++i) s[i] = 0;
        volatile uint8_t* b = buf_;
        for (int i = 0; i < 64; ++i) b[i] = 0;
    }
    ChaCha20(const ChaCha20&) = delete;  // a copy would replay the stream
    ChaCha20& operator=(const ChaCha20&) = delete;

    // rng.rs:118-132: bytes continue across calls; block `counter` is made
    // when the previous one is used up (the counter wraps as a u32).
    void fill(uint8_t* out, size_t len) {
        while (len) {
            if (used_ == 64) refill();
            const size_t take = std::min<size_t>(64 - used_, len);
            std::memcpy(out, buf_ + used_, take);
            used_ += take;
            out += take;
            len -= take;
        }
    }
    uint32_t next_u32() {
        uint8_t w[4];
        fill(w, 4);
        return load(w);
    }
    uint64_t next_u64() {
        uint8_t w[8];
        fill(w, 8);
        return uint64_t(load(w)) | uint64_t(load(w + 4)) << 32;
    }
    // rng.rs:157-168: uniform in [0, p) by rejecting words >= 2p, the largest
    // multiple of p below 2^32.
    uint32_t next_field() {
        constexpr uint32_t LIMIT = 2 * P;
        for (;;) {
            const uint32_t w = next_u32();
            if (w < LIMIT) return w >= P ? w - P : w;
        }
    }

    // rng.rs:185-209 (os_seed + fresh_stream): key from /dev/urandom, zero
    // nonce. Throws Status(RngUnavailable) when the device cannot be read.
    static ChaCha20* from_os() {
        uint8_t key[32];
        FILE* f = std::fopen("/dev/urandom", "rb");
        if (!f) throw Status(Z0P_ERR_RNG_UNAVAILABLE);
        const size_t got = std::fread(key, 1, 32, f);
        std::fclose(f);
        if (got != 32) throw Status(Z0P_ERR_RNG_UNAVAILABLE);
        const uint8_t nonce[12] = {0};
        ChaCha20* r = new ChaCha20(key, nonce);
        volatile uint8_t* k = key;
        for (int i = 0; i < 32; ++i) k[i] = 0;
        return r;
    }

private:
    static uint32_t load(const uint8_t* p) {
        return uint32_t(p[0]) | uint32_t(p[1]) << 8 | uint32_t(p[2]) << 16 | uint32_t(p[3]) << 24;
    }
    static uint32_t rotl(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }
    static void quarter(uint32_t& a, uint32_t& b, uint32_t& c, uint32_t& d) {
        a += b, d = rotl(d ^ a, 16);
        c += d, b = rotl(b ^ c, 12);
        a += b, d = rotl(d ^ a, 8);
        c += d, b = rotl(b ^ c, 7);
    }
    // RFC 8439 section 2.3: 10 double rounds, then add the input state.
    void refill() {
        uint32_t x[16];
        std::memcpy(x, state_, sizeof(x));
        for (int r = 0; r < 10; ++r) {
            quarter(x[0], x[4], x[8], x[12]);
            quarter(x[1], x[5], x[9], x[13]);
            quarter(x[2], x[6], x[10], x[14]);
            quarter(x[3], x[7], x[11], x[15]);
            quarter(x[0], x[5], x[10], x[15]);
            quarter(x[1], x[6], x[11], x[12]);
            quarter(x[2], x[7], x[8], x[13]);
            quarter(x[3], x[4], x[9], x[14]);
        }
        for (int i = 0; i < 16; ++i) {
            const uint32_t w = x[i] + state_[i];
            for (int b = 0; b < 4; ++b) buf_[4 * i + b] = uint8_t(w >> (8 * b));
        }
        state_[12] += 1;
        used_ = 0;
    }

    uint32_t state_[16];
    uint8_t buf_[64] = {0};
    size_t used_ = 64;
};

}  // namespace z0p
