// ORACLE — test infrastructure only (see bb.hpp).
//
// Hiding layer, restated from src/blind.rs (z0p v0.3.0): M = T + Z_H * R with
// Z_H = x^n - 1 and n_rand random coefficients of R.
#pragma once
#include <vector>

#include "rng.hpp"

namespace oracle {

// blind.rs:38-79
inline std::vector<Field> mask_coeffs(const std::vector<Field>& coeffs, size_t n, size_t n_rand, ChaCha20& rng) {
    if (n == 0 || (n & (n - 1)) != 0) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
    if (coeffs.size() > n) throw Z0pError(ERR_MALFORMED_INPUT, "coefficients longer than base domain");
    if (n_rand > SIZE_MAX - n) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
    const size_t total = n + n_rand;
    size_t padded = 1;
    while (padded < total) {
        if (padded > SIZE_MAX / 2) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        padded <<= 1;
    }
    if (padded > (size_t(1) << Field::TWO_ADICITY)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
    std::vector<Field> random(n_rand);
    for (auto& r : random) r = rng.next_field();
    std::vector<Field> masked(total, Field::zero());
    for (size_t i = 0; i < coeffs.size(); ++i) masked[i] = masked[i] + coeffs[i];
    for (size_t i = 0; i < n_rand; ++i) {
        masked[i + n] = masked[i + n] + random[i];
        masked[i] = masked[i] - random[i];
    }
    masked.resize(padded, Field::zero());
    return masked;
}

}  // namespace oracle
