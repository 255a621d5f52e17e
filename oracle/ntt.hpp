// ORACLE — test infrastructure only (see bb.hpp).
//
// Radix-2 DIT NTT with chained twiddle tables, restated from src/ntt.rs.
#pragma once
#include <vector>

#include "ext4.hpp"
#include "par.hpp"

namespace oracle {

inline size_t reverse_bits_n(size_t x, uint32_t bits) {
    if (bits == 0) return 0;
    size_t r = 0;
    for (uint32_t i = 0; i < bits; ++i) {
        r = (r << 1) | (x & 1);
        x >>= 1;
    }
    return r;
}

// ntt.rs:430-435
inline void check_domain(size_t n) {
    if (n == 0 || !is_pow2(n) || n > (size_t(1) << Field::TWO_ADICITY)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
}

// ntt.rs:21-44, 50-68 (in place below 2^18, parallel gather above).
template <typename T>
inline void bit_reverse_permute_t(std::vector<T>& values, size_t threshold) {
    size_t n = values.size();
    if (n <= 1) return;
    uint32_t bits = log2_exact(n);
    if (n < threshold) {
        for (size_t i = 0; i < n; ++i) {
            size_t rev = reverse_bits_n(i, bits);
            if (i < rev) std::swap(values[i], values[rev]);
        }
        return;
    }
    std::vector<T> out(n);
    for_each_indexed(out.data(), n, size_t(1) << 14, [&](size_t base, T* piece, size_t len) {
        for (size_t k = 0; k < len; ++k) piece[k] = values[reverse_bits_n(base + k, bits)];
    });
    values.swap(out);
}
inline void bit_reverse_permute(std::vector<Field>& v) { bit_reverse_permute_t(v, size_t(1) << 18); }
// ntt.rs:76-99 — the Ext4 variant gathers in parallel only from 2^24.
inline void bit_reverse_permute_ext(std::vector<Ext4>& v) { bit_reverse_permute_t(v, size_t(1) << 24); }

// ntt.rs:180-187
inline std::vector<Field> twiddle_table(Field root, size_t n) {
    std::vector<Field> t;
    t.reserve(n / 2);
    t.push_back(Field::one());
    for (size_t i = 1; i < n / 2; ++i) t.push_back(t[i - 1] * root);
    return t;
}

constexpr size_t PAR_NTT = size_t(1) << 16;
constexpr size_t PAR_BUTTERFLIES_PER_THREAD = size_t(1) << 14;

// ntt.rs:293-310
inline void butterfly_block(Field* block, size_t half, const Field* tw, size_t stride, const Field* scale) {
    for (size_t j = 0; j < half; ++j) {
        Field u = block[j];
        Field v = block[j + half] * tw[j * stride];
        Field sum = u + v, diff = u - v;
        if (scale) {
            block[j] = sum * *scale;
            block[j + half] = diff * *scale;
        } else {
            block[j] = sum;
            block[j + half] = diff;
        }
    }
}

// ntt.rs:207-247, 254-287
inline void butterfly_stage(std::vector<Field>& values, size_t len, const std::vector<Field>& tw, size_t stride,
                            const Field* scale) {
    size_t half = len / 2;
    size_t blocks = values.size() / len;
    size_t butterflies = blocks * half;
    size_t workers = values.size() < PAR_NTT ? 1 : worker_count(butterflies, PAR_BUTTERFLIES_PER_THREAD);
    Field* data = values.data();
    const Field* t = tw.data();
    if (workers <= 1) {
        for (size_t b = 0; b < blocks; ++b) butterfly_block(data + b * len, half, t, stride, scale);
        return;
    }
    if (blocks >= workers) {
        size_t per = (blocks + workers - 1) / workers;
        std::vector<std::thread> th;
        for (size_t b0 = 0; b0 < blocks; b0 += per) {
            size_t b1 = std::min(blocks, b0 + per);
            th.emplace_back([=]() {
                for (size_t b = b0; b < b1; ++b) butterfly_block(data + b * len, half, t, stride, scale);
            });
        }
        for (auto& x : th) x.join();
        return;
    }
    for (size_t b = 0; b < blocks; ++b) {
        Field* block = data + b * len;
        size_t chunk = (half + workers - 1) / workers;
        std::vector<std::thread> th;
        for (size_t lo = 0; lo < half; lo += chunk) {
            size_t hi = std::min(half, lo + chunk);
            th.emplace_back([=]() {
                for (size_t j = lo; j < hi; ++j) {
                    Field product = block[j + half] * t[j * stride];
                    Field sum = block[j] + product, diff = block[j] - product;
                    if (scale) {
                        block[j] = sum * *scale;
                        block[j + half] = diff * *scale;
                    } else {
                        block[j] = sum;
                        block[j + half] = diff;
                    }
                }
            });
        }
        for (auto& x : th) x.join();
    }
}

// ntt.rs:142-149
inline void forward_stages(std::vector<Field>& values, size_t n, Field omega) {
    std::vector<Field> tw = twiddle_table(omega, n);
    for (size_t len = 2; len <= n; len *= 2) butterfly_stage(values, len, tw, n / len, nullptr);
}

// ntt.rs:130-140 (forward_owned) and 113-122 (forward).
inline void forward(std::vector<Field>& values) {
    size_t n = values.size();
    check_domain(n);
    if (n == 1) return;
    Field omega = primitive_root(n);
    bit_reverse_permute(values);
    forward_stages(values, n, omega);
}

// ntt.rs:158-177 — 1/n folded into the last stage.
inline void inverse(std::vector<Field>& values) {
    size_t n = values.size();
    check_domain(n);
    if (n == 1) return;
    Field omega_inv = primitive_root(n).inv();
    std::vector<Field> tw = twiddle_table(omega_inv, n);
    Field n_inv = Field::from_u64(n).inv();
    bit_reverse_permute(values);
    for (size_t len = 2; len < n; len *= 2) butterfly_stage(values, len, tw, n / len, nullptr);
    butterfly_stage(values, n, tw, 1, &n_inv);
}

// ntt.rs:316-321
inline std::vector<Field> interpolate(const std::vector<Field>& evals) {
    check_domain(evals.size());
    std::vector<Field> c = evals;
    inverse(c);
    return c;
}

// ntt.rs:327-333
inline Field evaluate(const std::vector<Field>& coeffs, Field point) {
    Field acc = Field::zero();
    for (size_t i = coeffs.size(); i-- > 0;) acc = acc * point + coeffs[i];
    return acc;
}

// ntt.rs:341-363
inline std::vector<Field> evaluate_on_coset(const std::vector<Field>& coeffs, size_t blowup, Field shift) {
    if (!is_pow2(blowup)) throw Z0pError(ERR_INVALID_BLOWUP);
    check_domain(coeffs.size());
    size_t n = std::max<size_t>(coeffs.size(), 1);
    size_t m = n * blowup;
    if (m > (size_t(1) << Field::TWO_ADICITY)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
    std::vector<Field> ext(m, Field::zero());
    Field power = Field::one();
    for (size_t i = 0; i < coeffs.size(); ++i) {
        ext[i] = coeffs[i] * power;
        power *= shift;
    }
    forward(ext);
    return ext;
}

// ntt.rs:394-420 — each coordinate extended independently over the base field.
inline std::vector<Ext4> evaluate_on_coset_ext(const std::vector<Ext4>& coeffs, size_t blowup, Field shift) {
    if (coeffs.empty()) throw Z0pError(ERR_EMPTY_INPUT);
    std::vector<Field> coord[4];
    for (int c = 0; c < 4; ++c) {
        std::vector<Field> v(coeffs.size());
        for (size_t i = 0; i < coeffs.size(); ++i) v[i] = coeffs[i].c[c];
        coord[c] = evaluate_on_coset(v, blowup, shift);
    }
    std::vector<Ext4> out(coord[0].size());
    for (size_t i = 0; i < out.size(); ++i) out[i] = Ext4::make(coord[0][i], coord[1][i], coord[2][i], coord[3][i]);
    return out;
}

}  // namespace oracle
