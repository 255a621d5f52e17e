// ORACLE — test infrastructure only. Nothing under oracle/ is shipped or
// called by the product library; tests, smoke() and bench.py's cpu_baseline
// leg use it as the checker and as the host-CPU timing baseline.
//
// BabyBear prime field, restated from src/babybear.rs (z0p v0.3.0).
#pragma once
#include <cstdint>
#include <cstddef>
#include <vector>
#include <stdexcept>

#include "error.hpp"

namespace oracle {

// src/babybear.rs:22-38 — canonical u32 below p = 2^31 - 2^27 + 1.
struct Field {
    uint32_t v = 0;
    static constexpr uint32_t MODULUS = 2013265921u;
    static constexpr uint32_t TWO_ADICITY = 27;

    constexpr Field() = default;
    static constexpr Field raw(uint32_t x) { Field f; f.v = x; return f; }
    // babybear.rs:58 `new` reduces any u32.
    static constexpr Field make(uint32_t x) { return raw(x % MODULUS); }
    // babybear.rs:64-68 `from_u64`.
    static constexpr Field from_u64(uint64_t x) { return raw(uint32_t(x % MODULUS)); }

    static constexpr Field zero() { return raw(0); }
    static constexpr Field one() { return raw(1); }
    static constexpr Field generator() { return raw(31); }

    bool operator==(Field o) const { return v == o.v; }
    bool operator!=(Field o) const { return v != o.v; }

    // babybear.rs:74-83
    Field operator+(Field o) const {
        uint32_t s = v + o.v;
        return raw(s >= MODULUS ? s - MODULUS : s);
    }
    // babybear.rs:88-94
    Field operator-(Field o) const {
        return raw(v >= o.v ? v - o.v : MODULUS - (o.v - v));
    }
    // babybear.rs:99-106
    Field operator-() const { return v == 0 ? zero() : raw(MODULUS - v); }
    // babybear.rs:121-131 — Barrett with MU = floor(2^64 / p).
    Field operator*(Field o) const {
        constexpr uint64_t MU = 9162596893ull;
        uint64_t x = uint64_t(v) * uint64_t(o.v);
        uint64_t q = ((x >> 32) * MU) >> 32;
        uint64_t r = x - q * uint64_t(MODULUS);
        while (r >= MODULUS) r -= MODULUS;
        return raw(uint32_t(r));
    }
    Field& operator+=(Field o) { *this = *this + o; return *this; }
    Field& operator-=(Field o) { *this = *this - o; return *this; }
    Field& operator*=(Field o) { *this = *this * o; return *this; }

    Field square() const { return *this * *this; }
    // babybear.rs:147-158 — square-and-multiply over the exponent bits.
    Field pow(uint64_t e) const {
        Field acc = one(), base = *this;
        while (e > 0) {
            if (e & 1) acc = acc * base;
            base = base.square();
            e >>= 1;
        }
        return acc;
    }
    // babybear.rs:167-169 — Fermat, zero maps to zero.
    Field inv() const { return pow(uint64_t(MODULUS - 2)); }
    // babybear.rs:173-177
    Field halve() const {
        uint32_t t = v + (v & 1) * MODULUS;
        return raw(t >> 1);
    }
};

// babybear.rs:183-192
inline Field two_adic_generator(uint32_t bits) {
    if (bits > Field::TWO_ADICITY) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
    uint64_t e = uint64_t(Field::MODULUS - 1) >> bits;
    return Field::generator().pow(e);
}

inline bool is_pow2(size_t x) { return x != 0 && (x & (x - 1)) == 0; }
inline uint32_t log2_exact(size_t x) { return uint32_t(__builtin_ctzll(x)); }

// babybear.rs:199-210
inline Field primitive_root(size_t size) {
    if (!is_pow2(size)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
    uint32_t bits = log2_exact(size);
    if (bits > Field::TWO_ADICITY) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
    return two_adic_generator(bits);
}

// babybear.rs:307-331 — Montgomery batch inversion, zeros stay zero.
inline void batch_invert_in_place(Field* values, size_t n) {
    if (n == 0) return;
    std::vector<Field> prefix;
    prefix.reserve(n);
    Field acc = Field::one();
    for (size_t i = 0; i < n; ++i) {
        if (values[i] != Field::zero()) acc *= values[i];
        prefix.push_back(acc);
    }
    Field suffix = acc.inv();
    for (size_t i = n; i-- > 0;) {
        Field value = values[i];
        if (value == Field::zero()) continue;
        Field before = i == 0 ? Field::one() : prefix[i - 1];
        values[i] = before * suffix;
        suffix *= value;
    }
}

inline std::vector<Field> batch_invert(const std::vector<Field>& values) {
    std::vector<Field> out = values;
    batch_invert_in_place(out.data(), out.size());
    return out;
}

inline void put_u32(uint8_t* p, uint32_t x) {
    p[0] = uint8_t(x); p[1] = uint8_t(x >> 8); p[2] = uint8_t(x >> 16); p[3] = uint8_t(x >> 24);
}
inline void put_u64(uint8_t* p, uint64_t x) {
    for (int i = 0; i < 8; ++i) p[i] = uint8_t(x >> (8 * i));
}
inline uint32_t get_u32(const uint8_t* p) {
    return uint32_t(p[0]) | (uint32_t(p[1]) << 8) | (uint32_t(p[2]) << 16) | (uint32_t(p[3]) << 24);
}
inline uint64_t get_u64(const uint8_t* p) {
    uint64_t x = 0;
    for (int i = 7; i >= 0; --i) x = (x << 8) | p[i];
    return x;
}

// src/gf2.rs:425-431 — deterministic test PRNG step.
inline uint64_t splitmix64(uint64_t& state) {
    state += 0x9E3779B97F4A7C15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace oracle
