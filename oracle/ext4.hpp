// ORACLE — test infrastructure only (see bb.hpp).
//
// Quartic extension Fp4 = Fp2[v]/(v^2 - u), Fp2 = Fp[u]/(u^2 - 11),
// restated from src/ext4.rs.
#pragma once
#include <array>
#include "bb.hpp"

namespace oracle {

using Fp2 = std::array<Field, 2>;

// ext4.rs:22 — the non-residue 11.
inline Field nonresidue() { return Field::raw(11); }

// ext4.rs:190-197
inline Fp2 fp2_mul(Fp2 x, Fp2 y) {
    Field real = x[0] * y[0];
    Field imag = x[1] * y[1];
    Field cross = (x[0] + x[1]) * (y[0] + y[1]) - real - imag;
    return {real + imag * nonresidue(), cross};
}
// ext4.rs:201-203
inline Fp2 fp2_mul_u(Fp2 x) { return {x[1] * nonresidue(), x[0]}; }
inline Fp2 fp2_add(Fp2 x, Fp2 y) { return {x[0] + y[0], x[1] + y[1]}; }
inline Fp2 fp2_sub(Fp2 x, Fp2 y) { return {x[0] - y[0], x[1] - y[1]}; }
// ext4.rs:219-224
inline Fp2 fp2_inv(Fp2 x) {
    Field norm = x[0] * x[0] - nonresidue() * x[1] * x[1];
    Field ni = norm.inv();
    return {x[0] * ni, -x[1] * ni};
}

// ext4.rs:28 — a + b u + c v + d u v.
struct Ext4 {
    std::array<Field, 4> c{};

    static Ext4 make(Field a, Field b, Field cc, Field d) { Ext4 e; e.c = {a, b, cc, d}; return e; }
    static Ext4 from_base(Field x) { return make(x, Field::zero(), Field::zero(), Field::zero()); }
    static Ext4 zero() { return Ext4(); }
    static Ext4 one() { return from_base(Field::one()); }

    bool operator==(const Ext4& o) const { return c == o.c; }
    bool operator!=(const Ext4& o) const { return !(c == o.c); }

    Ext4 operator+(const Ext4& o) const { Ext4 r; for (int i = 0; i < 4; ++i) r.c[i] = c[i] + o.c[i]; return r; }
    Ext4 operator-(const Ext4& o) const { Ext4 r; for (int i = 0; i < 4; ++i) r.c[i] = c[i] - o.c[i]; return r; }
    Ext4 operator-() const { Ext4 r; for (int i = 0; i < 4; ++i) r.c[i] = -c[i]; return r; }
    // ext4.rs:106-117 — Karatsuba over the Fp2 tower.
    Ext4 operator*(const Ext4& o) const {
        Fp2 a{c[0], c[1]}, b{o.c[0], o.c[1]}, cc{c[2], c[3]}, d{o.c[2], o.c[3]};
        Fp2 ac = fp2_mul(a, b);
        Fp2 bd = fp2_mul(cc, d);
        Fp2 cross = fp2_sub(fp2_sub(fp2_mul(fp2_add(a, cc), fp2_add(b, d)), ac), bd);
        Fp2 bd_u = fp2_mul_u(bd);
        return make(ac[0] + bd_u[0], ac[1] + bd_u[1], cross[0], cross[1]);
    }
    // ext4.rs:121-123
    Ext4 mul_base(Field s) const { Ext4 r; for (int i = 0; i < 4; ++i) r.c[i] = c[i] * s; return r; }
    // ext4.rs:68-70
    Ext4 halve() const { Ext4 r; for (int i = 0; i < 4; ++i) r.c[i] = c[i].halve(); return r; }
    bool is_zero() const { return c[0] == Field::zero() && c[1] == Field::zero() && c[2] == Field::zero() && c[3] == Field::zero(); }
    // ext4.rs:155-165
    Ext4 inv() const {
        Fp2 cc{c[0], c[1]}, d{c[2], c[3]};
        Fp2 norm = fp2_sub(fp2_mul(cc, cc), fp2_mul_u(fp2_mul(d, d)));
        Fp2 ni = fp2_inv(norm);
        Fp2 real = fp2_mul(cc, ni);
        Fp2 imag = fp2_mul(d, ni);
        return make(real[0], real[1], -imag[0], -imag[1]);
    }
    // ext4.rs:134-145 — exponent up to 128 bits as (hi, lo).
    Ext4 pow(unsigned __int128 e) const {
        Ext4 acc = one(), base = *this;
        while (e > 0) {
            if (e & 1) acc = acc * base;
            base = base * base;
            e >>= 1;
        }
        return acc;
    }
    // ext4.rs:169-175
    void to_le_bytes(uint8_t out[16]) const {
        for (int i = 0; i < 4; ++i) put_u32(out + 4 * i, c[i].v);
    }
    // ext4.rs:178-187 — each coordinate reduced mod p.
    static Ext4 from_le_bytes(const uint8_t in[16]) {
        Ext4 r;
        for (int i = 0; i < 4; ++i) r.c[i] = Field::make(get_u32(in + 4 * i));
        return r;
    }
};

}  // namespace oracle
