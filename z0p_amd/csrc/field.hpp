// BabyBear arithmetic for host and device (gfx950).
//
// Representation: data values are canonical u32 (< p) everywhere they touch
// memory, hashes or the wire. Constants (twiddles, coset scales, challenges)
// are pre-converted to Montgomery form (x * 2^32 mod p), so a data x constant
// product is a single Montgomery multiply whose output is again canonical:
// mont_mul(a, b*R) = a*b*R*R^{-1} = a*b.
//
// Field semantics follow src/babybear.rs and src/ext4.rs of the reference
// (same modulus 2^31 - 2^27 + 1, same Fp4 tower Fp[u]/(u^2-11), Fp2[v]/(v^2-u));
// exact modular arithmetic makes any correct multiplier bit-identical to the
// reference's Barrett form.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define Z0P_HD __host__ __device__ __forceinline__
#else
#define Z0P_HD inline
#endif

namespace z0p {

constexpr uint32_t P = 0x78000001u;           // 2013265921
constexpr uint32_t P_NEG_INV = 0x77FFFFFFu;   // -p^{-1} mod 2^32
constexpr uint32_t R1 = uint32_t((uint64_t(1) << 32) % P);  // 2^32 mod p
constexpr uint32_t R2 = uint32_t((uint64_t(R1) * R1) % P);  // 2^64 mod p
constexpr uint32_t GENERATOR = 31;
constexpr uint32_t TWO_ADICITY = 27;
constexpr uint32_t NONRESIDUE = 11;

// a, b < p. Both select by an unsigned minimum: the wrong candidate has
// wrapped past 2^32 - p and is the larger one (2p < 2^32).
Z0P_HD uint32_t add(uint32_t a, uint32_t b) {
    const uint32_t s = a + b, t = s - P;
    return s < t ? s : t;
}
Z0P_HD uint32_t sub(uint32_t a, uint32_t b) {
    const uint32_t d = a - b, e = d + P;
    return d < e ? d : e;
}
Z0P_HD uint32_t neg(uint32_t a) { return a == 0 ? 0 : P - a; }
// Unreduced forms for an operand that only a Montgomery product consumes:
// for a, b < p they lie in [0, 2p), and a Montgomery product of such a value
// with a canonical constant is still in range (below 2p^2 < p 2^32) and
// canonical. One and two instructions instead of three.
Z0P_HD uint32_t sub_lazy(uint32_t a, uint32_t b) { return a + (P - b); }
Z0P_HD uint32_t add_lazy(uint32_t a, uint32_t b) { return a + b; }
Z0P_HD uint32_t halve(uint32_t a) { return (a + ((a & 1u) ? P : 0u)) >> 1; }
Z0P_HD uint32_t dbl(uint32_t a) { return add(a, a); }

// Montgomery reduction T * 2^{-32} mod p for T < p * 2^32; output < p.
// m*p + T < 2 p 2^32 < 2^64, and the high word is below 2p.
// The final subtraction is an unsigned minimum (r < 2p < 2^32), which keeps
// the result in vector registers with no compare-mask hazard.
Z0P_HD uint32_t redc(uint64_t t) {
    uint32_t m = uint32_t(t) * P_NEG_INV;
    uint64_t u = uint64_t(m) * P + t;
    uint32_t r = uint32_t(u >> 32), s = r - P;
    return r < s ? r : s;
}

// Montgomery multiply: a*b*2^{-32} mod p for a < 2p, b < p; output < p.
Z0P_HD uint32_t mont_mul(uint32_t a, uint32_t b) { return redc(uint64_t(a) * b); }

// a*b + c*d reduced once (all < p: the sum is below 2p^2 < p 2^32).
Z0P_HD uint32_t mont_mul2(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    return redc(uint64_t(a) * b + uint64_t(c) * d);
}

Z0P_HD uint32_t to_mont(uint32_t a) { return mont_mul(a, R2); }
Z0P_HD uint32_t from_mont(uint32_t a) { return mont_mul(a, 1u); }
// Canonical x canonical -> canonical (two Montgomery steps).
Z0P_HD uint32_t mul(uint32_t a, uint32_t b) { return mont_mul(mont_mul(a, b), R2); }
// x * 11 by additions (the Fp2 non-residue).
Z0P_HD uint32_t mul11(uint32_t a) {
    uint32_t a2 = add(a, a), a4 = add(a2, a2), a8 = add(a4, a4);
    return add(add(a8, a2), a);
}

// Canonical power (host-side tables and the verifier).
Z0P_HD uint32_t pow_c(uint32_t base, uint64_t e) {
    uint32_t acc = R1, b = to_mont(base);  // Montgomery domain
    while (e) {
        if (e & 1) acc = mont_mul(acc, b);
        b = mont_mul(b, b);
        e >>= 1;
    }
    return from_mont(acc);
}
Z0P_HD uint32_t inv_c(uint32_t a) { return pow_c(a, P - 2); }

// Generator of the order-2^bits subgroup: 31^((p-1) >> bits), as in
// babybear.rs:183-192.
Z0P_HD uint32_t root_of_unity(uint32_t bits) { return pow_c(GENERATOR, uint64_t(P - 1) >> bits); }

// ---------------------------------------------------------------------------
// Fp4 = Fp2[v]/(v^2 - u), Fp2 = Fp[u]/(u^2 - 11); element a + b u + c v + d uv.
struct E4 {
    uint32_t v[4];
};

Z0P_HD E4 e4_zero() { return E4{{0, 0, 0, 0}}; }
Z0P_HD E4 e4_base(uint32_t x) { return E4{{x, 0, 0, 0}}; }
Z0P_HD E4 e4_add(const E4& a, const E4& b) {
    return E4{{add(a.v[0], b.v[0]), add(a.v[1], b.v[1]), add(a.v[2], b.v[2]), add(a.v[3], b.v[3])}};
}
Z0P_HD E4 e4_sub(const E4& a, const E4& b) {
    return E4{{sub(a.v[0], b.v[0]), sub(a.v[1], b.v[1]), sub(a.v[2], b.v[2]), sub(a.v[3], b.v[3])}};
}
Z0P_HD E4 e4_halve(const E4& a) { return E4{{halve(a.v[0]), halve(a.v[1]), halve(a.v[2]), halve(a.v[3])}}; }
Z0P_HD bool e4_eq(const E4& a, const E4& b) {
    return a.v[0] == b.v[0] && a.v[1] == b.v[1] && a.v[2] == b.v[2] && a.v[3] == b.v[3];
}
// Scale by a base-field constant given in Montgomery form.
Z0P_HD E4 e4_scale_m(const E4& a, uint32_t sm) {
    return E4{{mont_mul(a.v[0], sm), mont_mul(a.v[1], sm), mont_mul(a.v[2], sm), mont_mul(a.v[3], sm)}};
}

// Product of a canonical element x and an element y whose coordinates are
// in Montgomery form; output canonical. Karatsuba over the tower.
Z0P_HD E4 e4_mul_m(const E4& x, const E4& ym) {
    // Fp2 products (x0 + x1 u)(y0 + y1 u) = (x0y0 + 11 x1y1) + ((x0+x1)(y0+y1) - x0y0 - x1y1) u
    uint32_t a0 = x.v[0], a1 = x.v[1], c0 = x.v[2], c1 = x.v[3];
    uint32_t b0 = ym.v[0], b1 = ym.v[1], d0 = ym.v[2], d1 = ym.v[3];
    // ac = a*b
    uint32_t r = mont_mul(a0, b0), i = mont_mul(a1, b1);
    uint32_t ac0 = add(r, mul11(i));
    uint32_t ac1 = sub(sub(mont_mul(add(a0, a1), add(b0, b1)), r), i);
    // bd = c*d
    uint32_t r2 = mont_mul(c0, d0), i2 = mont_mul(c1, d1);
    uint32_t bd0 = add(r2, mul11(i2));
    uint32_t bd1 = sub(sub(mont_mul(add(c0, c1), add(d0, d1)), r2), i2);
    // cross = (a+c)(b+d) - ac - bd
    uint32_t s0 = add(a0, c0), s1 = add(a1, c1), t0 = add(b0, d0), t1 = add(b1, d1);
    uint32_t r3 = mont_mul(s0, t0), i3 = mont_mul(s1, t1);
    uint32_t x0 = add(r3, mul11(i3));
    uint32_t x1 = sub(sub(mont_mul(add(s0, s1), add(t0, t1)), r3), i3);
    uint32_t cr0 = sub(sub(x0, ac0), bd0), cr1 = sub(sub(x1, ac1), bd1);
    // bd * u = (11 bd1) + bd0 u
    return E4{{add(ac0, mul11(bd1)), add(ac1, bd0), cr0, cr1}};
}

// Multiplication by a fixed y as the 4x4 matrix it induces on the basis
// (1, u, v, uv) (u^2 = 11, v^2 = u):
//   out0 = x0 y0 + x1 11y1 + x2 11y3 + x3 11y2
//   out1 = x0 y1 + x1 y0   + x2 y2   + x3 11y3
//   out2 = x0 y2 + x1 11y3 + x2 y0   + x3 11y1
//   out3 = x0 y3 + x1 y2   + x2 y1   + x3 y0
// The seven distinct entries (y0, y1, y2, y3, 11y1, 11y2, 11y3) are given in
// Montgomery form; each output is two pair sums reduced once each.
struct E4Mat {
    uint32_t y0, y1, y2, y3, y1e, y2e, y3e, pad;
};
Z0P_HD E4Mat e4_mat_m(const E4& y) {
    return E4Mat{to_mont(y.v[0]), to_mont(y.v[1]), to_mont(y.v[2]), to_mont(y.v[3]),
                 to_mont(mul11(y.v[1])), to_mont(mul11(y.v[2])), to_mont(mul11(y.v[3])), 0};
}
Z0P_HD E4 e4_mul_mat(const E4& x, const E4Mat& M) {
    const uint32_t x0 = x.v[0], x1 = x.v[1], x2 = x.v[2], x3 = x.v[3];
    return E4{{add(mont_mul2(x0, M.y0, x1, M.y1e), mont_mul2(x2, M.y3e, x3, M.y2e)),
               add(mont_mul2(x0, M.y1, x1, M.y0), mont_mul2(x2, M.y2, x3, M.y3e)),
               add(mont_mul2(x0, M.y2, x1, M.y3e), mont_mul2(x2, M.y0, x3, M.y1e)),
               add(mont_mul2(x0, M.y3, x1, M.y2), mont_mul2(x2, M.y1, x3, M.y0))}};
}

Z0P_HD E4 e4_to_mont(const E4& a) {
    return E4{{to_mont(a.v[0]), to_mont(a.v[1]), to_mont(a.v[2]), to_mont(a.v[3])}};
}
// Canonical x canonical.
Z0P_HD E4 e4_mul(const E4& a, const E4& b) { return e4_mul_m(a, e4_to_mont(b)); }
Z0P_HD E4 e4_mul_base(const E4& a, uint32_t s) { return e4_scale_m(a, to_mont(s)); }

// Inverse via the tower norm (maps zero to zero), canonical in/out.
inline E4 e4_inv(const E4& x) {
    auto fp2_mul = [](uint32_t a0, uint32_t a1, uint32_t b0, uint32_t b1, uint32_t& o0, uint32_t& o1) {
        uint32_t r = mul(a0, b0), i = mul(a1, b1);
        o0 = add(r, mul11(i));
        o1 = sub(sub(mul(add(a0, a1), add(b0, b1)), r), i);
    };
    uint32_t cc0, cc1, dd0, dd1;
    fp2_mul(x.v[0], x.v[1], x.v[0], x.v[1], cc0, cc1);
    fp2_mul(x.v[2], x.v[3], x.v[2], x.v[3], dd0, dd1);
    // norm = c^2 - u d^2 ; u*(dd0 + dd1 u) = 11 dd1 + dd0 u
    uint32_t n0 = sub(cc0, mul11(dd1)), n1 = sub(cc1, dd0);
    // Fp2 inverse: (n0 - n1 u)/(n0^2 - 11 n1^2)
    uint32_t nn = sub(mul(n0, n0), mul11(mul(n1, n1)));
    uint32_t ni = inv_c(nn);
    uint32_t q0 = mul(n0, ni), q1 = neg(mul(n1, ni));
    uint32_t r0, r1, s0, s1;
    fp2_mul(x.v[0], x.v[1], q0, q1, r0, r1);
    fp2_mul(x.v[2], x.v[3], q0, q1, s0, s1);
    return E4{{r0, r1, neg(s0), neg(s1)}};
}

}  // namespace z0p
