// ORACLE — test infrastructure only (see bb.hpp).
//
// Canonical VERSION = 6 wire format, restated from src/codec.rs.
#pragma once
#include <vector>

#include "ext4.hpp"
#include "merkle.hpp"

namespace oracle {

// codec.rs:21-34
constexpr uint8_t VERSION = 6;
constexpr uint32_t MAX_QUERIES = 1u << 20;
constexpr uint32_t MAX_FOLDS = 64;
constexpr uint32_t MAX_PATH_STEPS = 128;
constexpr uint32_t MAX_CAP_DIGESTS = 256;
constexpr uint32_t MAX_ELEMS = 1u << 27;

// codec.rs:38-114
struct Writer {
    std::vector<uint8_t> out;
    void byte(uint8_t b) { out.push_back(b); }
    void u32_le(uint32_t x) { uint8_t b[4]; put_u32(b, x); out.insert(out.end(), b, b + 4); }
    void u64_le(uint64_t x) { uint8_t b[8]; put_u64(b, x); out.insert(out.end(), b, b + 8); }
    void count(size_t n) { u32_le(uint32_t(n)); }
    void bytes(const uint8_t* d, size_t n) { out.insert(out.end(), d, d + n); }
    void field(Field f) { u32_le(f.v); }
    void ext4(const Ext4& e) { for (int i = 0; i < 4; ++i) field(e.c[i]); }
    void direction(bool is_left) { byte(is_left ? 1 : 0); }
    void path(const Path& p) {
        count(p.size());
        for (auto& s : p) {
            bytes(s.first.data(), 32);
            direction(s.second);
        }
    }
    void cap(const std::vector<Digest>& c) {
        count(c.size());
        for (auto& d : c) bytes(d.data(), 32);
    }
};

// codec.rs:118-274 — fail-closed reader.
struct Reader {
    const uint8_t* in;
    size_t len;
    size_t pos = 0;
    Reader(const uint8_t* p, size_t n) : in(p), len(n) {}
    size_t remaining() const { return len - pos; }
    const uint8_t* take(size_t n) {
        if (n > len - pos) throw Z0pError(ERR_MALFORMED_INPUT, "truncated proof encoding");
        const uint8_t* p = in + pos;
        pos += n;
        return p;
    }
    uint8_t byte() { return take(1)[0]; }
    uint32_t u32_le() { return get_u32(take(4)); }
    uint64_t u64_le() { return get_u64(take(8)); }
    size_t dimension() { return size_t(u64_le()); }
    Digest digest() {
        Digest d;
        std::memcpy(d.data(), take(32), 32);
        return d;
    }
    Field field() {
        uint32_t v = u32_le();
        if (v >= Field::MODULUS) throw Z0pError(ERR_MALFORMED_INPUT, "non-canonical field encoding");
        return Field::raw(v);
    }
    Ext4 ext4() {
        Field a = field(), b = field(), c = field(), d = field();
        return Ext4::make(a, b, c, d);
    }
    bool direction() {
        uint8_t b = byte();
        if (b == 0) return false;
        if (b == 1) return true;
        throw Z0pError(ERR_MALFORMED_INPUT, "invalid Merkle direction byte");
    }
    size_t count(uint32_t max) {
        uint32_t n = u32_le();
        if (n > max) throw Z0pError(ERR_MALFORMED_INPUT, "proof count exceeds hard limit");
        return n;
    }
    Path path() {
        Path p;
        size_t n = count(MAX_PATH_STEPS);
        for (size_t i = 0; i < n; ++i) {
            Digest d = digest();
            bool dir = direction();
            p.emplace_back(d, dir);
        }
        return p;
    }
    std::vector<Digest> cap() {
        std::vector<Digest> c;
        size_t n = count(MAX_CAP_DIGESTS);
        for (size_t i = 0; i < n; ++i) c.push_back(digest());
        return c;
    }
    void version() {
        if (byte() != VERSION) throw Z0pError(ERR_MALFORMED_INPUT, "unsupported proof encoding version");
    }
    void end() const {
        if (remaining() != 0) throw Z0pError(ERR_MALFORMED_INPUT, "trailing bytes after proof encoding");
    }
};

}  // namespace oracle
