// Host Fiat-Shamir transcript and wire codec.
//
// Transcript: the absorbed history is one streaming BLAKE3 state plus a
// squeeze counter, exactly the state of the reference transcript
// (src/transcript.rs:54-171): absorb frames `u64 len || label || u64 len ||
// data`; a squeeze hashes `history || u64 len || label || u64 counter`, bumps
// the counter and re-absorbs the 32-byte output under the same label.
//
// Codec: the VERSION = 6 wire format of src/codec.rs (u32 counts, 4-byte
// canonical fields, 32-byte digests, one direction byte per path step).
#pragma once
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "blake3.hpp"
#include "status.hpp"

namespace z0p {

class Transcript {
   public:
    Transcript() = default;
    explicit Transcript(const uint8_t* label, size_t n) { absorb("domain", label, n); }

    void absorb(const char* label, const uint8_t* data, size_t n) {
        size_t ll = std::strlen(label);
        uint8_t len8[8];
        store_le64(len8, ll);
        h_.update(len8, 8);
        h_.update(reinterpret_cast<const uint8_t*>(label), ll);
        store_le64(len8, n);
        h_.update(len8, 8);
        h_.update(data, n);
    }
    void absorb(const char* label, const std::vector<uint8_t>& v) { absorb(label, v.data(), v.size()); }

    void squeeze(const char* label, uint8_t out[32]) {
        Blake3 h;
        h_.fork_into(h);
        size_t ll = std::strlen(label);
        uint8_t len8[8];
        store_le64(len8, ll);
        h.update(len8, 8);
        h.update(reinterpret_cast<const uint8_t*>(label), ll);
        store_le64(len8, counter_);
        h.update(len8, 8);
        ++counter_;
        h.digest(out);
        absorb(label, out, 32);
    }

    // Four LE words of the squeeze, each reduced mod p (transcript.rs:136-141).
    E4 challenge_ext(const char* label) {
        uint8_t d[32];
        squeeze(label, d);
        E4 e;
        for (int i = 0; i < 4; ++i) e.v[i] = load_le32(d + 4 * i) % P;
        return e;
    }

    // Low u64 of the squeeze masked to a power-of-two domain (transcript.rs:148-160).
    uint64_t challenge_index(const char* label, uint64_t domain) {
        uint8_t d[32];
        squeeze(label, d);
        return load_le64(d) & (domain - 1);
    }

    // Proof-of-work head: BLAKE3(history) || u64 len || label (transcript.rs:182-189).
    std::vector<uint8_t> pow_head(const char* label) const {
        std::vector<uint8_t> head(32);
        h_.digest(head.data());
        size_t ll = std::strlen(label);
        uint8_t len8[8];
        store_le64(len8, ll);
        head.insert(head.end(), len8, len8 + 8);
        head.insert(head.end(), label, label + ll);
        return head;
    }

    void history_digest(uint8_t out[32]) const { h_.digest(out); }
    uint64_t counter() const { return counter_; }

   private:
    Blake3 h_;
    uint64_t counter_ = 0;
};

// Leading zero bits, MSB-first from byte 0 (transcript.rs:341-351).
inline uint32_t leading_zero_bits(const uint8_t d[32]) {
    uint32_t n = 0;
    for (int i = 0; i < 32; ++i) {
        if (d[i] == 0) {
            n += 8;
            continue;
        }
        uint32_t b = d[i];
        while ((b & 0x80u) == 0) {
            ++n;
            b <<= 1;
        }
        break;
    }
    return n;
}

// Digest of head || nonce_le for the proof-of-work check.
inline void pow_digest(const std::vector<uint8_t>& head, uint64_t nonce, uint8_t out[32]) {
    std::vector<uint8_t> in(head);
    uint8_t nb[8];
    store_le64(nb, nonce);
    in.insert(in.end(), nb, nb + 8);
    blake3_hash(in.data(), in.size(), out);
}

constexpr uint8_t WIRE_VERSION = 6;
constexpr uint32_t MAX_QUERIES = 1u << 20;
constexpr uint32_t MAX_FOLDS = 64;
constexpr uint32_t MAX_PATH_STEPS = 128;
constexpr uint32_t MAX_CAP_DIGESTS = 256;
constexpr uint32_t MAX_ELEMS = 1u << 27;

// Serialized bytes in a malloc'd buffer: the C boundary hands it to the
// caller (freed with z0p_free) without copying.
class Bytes {
   public:
    Bytes() = default;
    Bytes(const Bytes&) = delete;
    Bytes& operator=(const Bytes&) = delete;
    Bytes(Bytes&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr, o.n_ = 0; }
    Bytes& operator=(Bytes&& o) noexcept {
        if (this != &o) {
            std::free(p_);
            p_ = o.p_, n_ = o.n_;
            o.p_ = nullptr, o.n_ = 0;
        }
        return *this;
    }
    ~Bytes() { std::free(p_); }
    Bytes(uint8_t* p, size_t n) : p_(p), n_(n) {}
    const uint8_t* data() const { return p_; }
    size_t size() const { return n_; }
    bool empty() const { return n_ == 0; }
    bool operator==(const Bytes& o) const { return n_ == o.n_ && (n_ == 0 || std::memcmp(p_, o.p_, n_) == 0); }
    // Ownership to the caller (a 1-byte allocation for an empty buffer).
    uint8_t* release() {
        uint8_t* r = p_ ? p_ : static_cast<uint8_t*>(std::malloc(1));
        p_ = nullptr, n_ = 0;
        return r;
    }

   private:
    uint8_t* p_ = nullptr;
    size_t n_ = 0;
};

// Appends through a cursor into a malloc'd buffer sized up front (reserve()
// with the expected total); finish() trims it to the bytes written.
struct Writer {
    uint8_t* buf = nullptr;
    size_t cap = 0, pos = 0;
    Writer() = default;
    Writer(const Writer&) = delete;
    Writer& operator=(const Writer&) = delete;
    ~Writer() { std::free(buf); }
    void reserve(size_t bytes) {
        if (cap < bytes) grow(bytes);
    }
    uint8_t* room(size_t k) {
        if (pos + k > cap) grow(std::max(2 * cap, pos + k));
        uint8_t* p = buf + pos;
        pos += k;
        return p;
    }
    void byte(uint8_t b) { *room(1) = b; }
    void u32(uint32_t x) { store_le32(room(4), x); }
    void u64(uint64_t x) { store_le64(room(8), x); }
    void raw(const void* p, size_t n) { std::memcpy(room(n), p, n); }
    void e4(const E4& e) {
        uint8_t* p = room(16);
        for (int i = 0; i < 4; ++i) store_le32(p + 4 * i, e.v[i]);
    }
    // `count` u32 words as little-endian bytes (the host is little-endian).
    void words(const uint32_t* w, size_t count) {
        static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "little-endian host");
        std::memcpy(room(4 * count), w, 4 * count);
    }
    Bytes finish() {
        uint8_t* p = pos ? static_cast<uint8_t*>(std::realloc(buf, pos)) : buf;
        if (!p) p = buf;
        Bytes out(p, pos);
        buf = nullptr, cap = pos = 0;
        return out;
    }

   private:
    void grow(size_t bytes) {
        uint8_t* p = static_cast<uint8_t*>(std::realloc(buf, bytes));
        if (!p) throw Status(Z0P_ERR_OUT_OF_MEMORY);
        buf = p, cap = bytes;
    }
};

// Fail-closed reader: every violation is MalformedInput.
struct Reader {
    const uint8_t* p;
    size_t n, pos = 0;
    Reader(const uint8_t* d, size_t len) : p(d), n(len) {}
    const uint8_t* take(size_t k) {
        if (k > n - pos) throw Status(Z0P_ERR_MALFORMED_INPUT);
        const uint8_t* r = p + pos;
        pos += k;
        return r;
    }
    uint8_t byte() { return take(1)[0]; }
    uint32_t u32() { return load_le32(take(4)); }
    uint64_t u64() { return load_le64(take(8)); }
    uint32_t field() {
        uint32_t v = u32();
        if (v >= P) throw Status(Z0P_ERR_MALFORMED_INPUT);
        return v;
    }
    E4 e4() {
        E4 e;
        for (int i = 0; i < 4; ++i) e.v[i] = field();
        return e;
    }
    uint32_t count(uint32_t max) {
        uint32_t c = u32();
        if (c > max) throw Status(Z0P_ERR_MALFORMED_INPUT);
        return c;
    }
    bool direction() {
        uint8_t b = byte();
        if (b > 1) throw Status(Z0P_ERR_MALFORMED_INPUT);
        return b == 1;
    }
    void version() {
        if (byte() != WIRE_VERSION) throw Status(Z0P_ERR_MALFORMED_INPUT);
    }
    void end() const {
        if (pos != n) throw Status(Z0P_ERR_MALFORMED_INPUT);
    }
};

}  // namespace z0p
