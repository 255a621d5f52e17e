// ORACLE — test infrastructure only (see bb.hpp).
//
// Fiat-Shamir transcript and proof-of-work grind, restated from
// src/transcript.rs.
#pragma once
#include <cstring>
#include <mutex>
#include <optional>
#include <string>
#include <thread>
#include <vector>

#include "blake3.hpp"
#include "ext4.hpp"
#include "par.hpp"

namespace oracle {

// transcript.rs:341-351 — MSB-first from byte 0.
inline uint32_t leading_zero_bits(const Digest& d) {
    uint32_t count = 0;
    for (uint8_t b : d) {
        uint32_t z = b == 0 ? 8 : uint32_t(__builtin_clz(uint32_t(b)) - 24);
        count += z;
        if (z < 8) break;
    }
    return count;
}

// transcript.rs:256-261 — head || nonce_le.
inline Digest pow_digest(const std::vector<uint8_t>& head, uint64_t nonce) {
    std::vector<uint8_t> input(head);
    uint8_t nb[8];
    put_u64(nb, nonce);
    input.insert(input.end(), nb, nb + 8);
    if (input.size() <= CHUNK_LEN) return blake3_256_small(input.data(), input.size());
    return blake3_256(input.data(), input.size());
}

// transcript.rs:284-303 — first passing nonce in [start, start + count).
inline std::optional<uint64_t> scan_range(const std::vector<uint8_t>& head, uint64_t start, uint64_t count,
                                          uint32_t bits) {
    std::vector<uint8_t> input(head);
    input.resize(head.size() + 8);
    size_t base = head.size();
    uint64_t nonce = start;
    for (uint64_t i = 0; i < count; ++i) {
        put_u64(input.data() + base, nonce);
        Digest d = input.size() <= CHUNK_LEN ? blake3_256_small(input.data(), input.size())
                                             : blake3_256(input.data(), input.size());
        if (leading_zero_bits(d) >= bits) return nonce;
        nonce += 1;
    }
    return std::nullopt;
}

// transcript.rs:263-281
inline uint64_t grind_sequential(const std::vector<uint8_t>& head, uint32_t bits) {
    uint64_t nonce = 0;
    for (;;) {
        auto r = scan_range(head, nonce, 1u << 16, bits);
        if (r) return *r;
        nonce += 1u << 16;
    }
}

// transcript.rs:310-338 — bands of workers * 2^min(d,16); band minimum wins.
inline uint64_t grind_parallel(const std::vector<uint8_t>& head, uint32_t bits, size_t workers) {
    uint64_t per_worker = 1ull << std::min<uint32_t>(bits, 16);
    uint64_t band = uint64_t(workers) * per_worker;
    uint64_t base = 0;
    for (;;) {
        std::vector<std::optional<uint64_t>> found(workers);
        std::vector<std::thread> threads;
        for (size_t w = 0; w < workers; ++w) {
            uint64_t start = base + uint64_t(w) * per_worker;
            threads.emplace_back([&, w, start]() { found[w] = scan_range(head, start, per_worker, bits); });
        }
        for (auto& t : threads) t.join();
        std::optional<uint64_t> best;
        for (auto& f : found)
            if (f && (!best || *f < *best)) best = f;
        if (best) return *best;
        base += band;
    }
}

// transcript.rs:54-171
struct Transcript {
    Hasher history;
    uint64_t counter = 0;

    explicit Transcript(const std::string& label) { absorb("domain", (const uint8_t*)label.data(), label.size()); }
    Transcript(const uint8_t* label, size_t n) { absorb("domain", label, n); }

    // transcript.rs:72-77
    void absorb(const std::string& label, const uint8_t* data, size_t n) {
        uint8_t lb[8];
        put_u64(lb, label.size());
        history.update(lb, 8);
        history.update((const uint8_t*)label.data(), label.size());
        put_u64(lb, n);
        history.update(lb, 8);
        history.update(data, n);
    }
    void absorb(const std::string& label, const std::vector<uint8_t>& data) { absorb(label, data.data(), data.size()); }

    // transcript.rs:162-171
    Digest squeeze_raw(const std::string& label) {
        Hasher h = history;
        uint8_t lb[8];
        put_u64(lb, label.size());
        h.update(lb, 8);
        h.update((const uint8_t*)label.data(), label.size());
        put_u64(lb, counter);
        h.update(lb, 8);
        counter += 1;
        Digest out = h.finalize_consume();
        absorb(label, out.data(), 32);
        return out;
    }

    // transcript.rs:136-141 — first 16 bytes, four LE words mod p.
    Ext4 challenge_ext(const std::string& label) {
        Digest d = squeeze_raw(label);
        return Ext4::from_le_bytes(d.data());
    }

    // transcript.rs:148-160
    std::optional<size_t> challenge_index(const std::string& label, size_t domain_size) {
        if (domain_size == 0 || !is_pow2(domain_size)) return std::nullopt;
        Digest d = squeeze_raw(label);
        return size_t(get_u64(d.data()) & (uint64_t(domain_size) - 1));
    }

    // transcript.rs:182-189
    std::vector<uint8_t> pow_head(const std::string& label) const {
        Digest hist = history.finalize();
        std::vector<uint8_t> head(hist.begin(), hist.end());
        uint8_t lb[8];
        put_u64(lb, label.size());
        head.insert(head.end(), lb, lb + 8);
        head.insert(head.end(), label.begin(), label.end());
        return head;
    }

    // transcript.rs:206-228
    uint64_t grind(const std::string& label, uint32_t bits) {
        uint8_t nb[8];
        if (bits > 32) {
            put_u64(nb, 0);
            absorb(label, nb, 8);
            return 0;
        }
        std::vector<uint8_t> head = pow_head(label);
        size_t units = size_t(1) << bits;
        size_t workers = worker_count(units, 4096);
        uint64_t nonce = workers > 1 ? grind_parallel(head, bits, workers) : grind_sequential(head, bits);
        put_u64(nb, nonce);
        absorb(label, nb, 8);
        return nonce;
    }

    // transcript.rs:236-246
    bool verify_grind(const std::string& label, uint64_t nonce, uint32_t bits) {
        if (bits > 256) return false;
        std::vector<uint8_t> head = pow_head(label);
        if (leading_zero_bits(pow_digest(head, nonce)) < bits) return false;
        uint8_t nb[8];
        put_u64(nb, nonce);
        absorb(label, nb, 8);
        return true;
    }
};

}  // namespace oracle
