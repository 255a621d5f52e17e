// ORACLE — test infrastructure only (see bb.hpp).
//
// Range-check AIR prover/verifier (plain and zero-knowledge), restated from
// src/range.rs: a value column V and n_bits bit columns C_j, one joint row
// tree over the LDE (one leaf per LDE row, BLAKE3(0x00 || V || C_0 || ..)),
// and the composition
//   a_d (V - sum 2^j C_j) / Z_H + sum_j a_j C_j (C_j - 1) / Z_H + a_b (V - v) / (x - 1)
// committed by fri::commit_ext_evals.
#pragma once
#include <vector>

#include "blind.hpp"
#include "fri.hpp"

namespace oracle {
namespace range {

// range.rs:29-52
struct Params {
    size_t n_rows = 0, n_bits = 0, blowup = 0, num_queries = 0, final_size = 0;
    uint32_t grinding_bits = 0;
    size_t cap_height = 0;
    Field value;

    // range.rs:59-98
    static Params make(size_t n_rows, size_t n_bits, size_t blowup, size_t num_queries, size_t final_size,
                       Field value) {
        if (n_rows < 2 || !is_pow2(n_rows)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        if (n_bits == 0 || n_bits > 24) throw Z0pError(ERR_MALFORMED_INPUT, "n_bits must be within 1..=24");
        if (blowup < 2 || !is_pow2(blowup)) throw Z0pError(ERR_INVALID_BLOWUP);
        if (num_queries == 0) throw Z0pError(ERR_NO_QUERIES);
        if (final_size < 2 || !is_pow2(final_size)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        unsigned __int128 lde = (unsigned __int128)n_rows * blowup;
        if (lde > ((unsigned __int128)1 << Field::TWO_ADICITY)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        Params p;
        p.n_rows = n_rows;
        p.n_bits = n_bits;
        p.blowup = blowup;
        p.num_queries = num_queries;
        p.final_size = final_size;
        p.value = value;
        return p;
    }

    // range.rs:124-158 — blowup 16, strength from the composition's
    // millibits, shape from one column opening of 8 + 4 (n_bits + 1) bytes.
    static Params for_security(size_t n_rows, size_t n_bits, Field value, uint32_t target_bits) {
        if (n_rows < 2 || !is_pow2(n_rows)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        if (n_bits == 0 || n_bits > 24) throw Z0pError(ERR_MALFORMED_INPUT, "n_bits must be within 1..=24");
        size_t blowup = 16;
        if (n_rows > (size_t(1) << Field::TWO_ADICITY) / blowup) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        size_t lde_len = n_rows * blowup;
        uint32_t mb = fri::composition_millibits(lde_len, n_rows - 1);
        auto [nq, gb] = fri::tune_strength(target_bits, mb);
        size_t value_bytes = 8 + 4 * (n_bits + 1);
        auto [fs, ch] = fri::tune_shape(lde_len, nq, {{lde_len, value_bytes}});
        Params p;
        p.n_rows = n_rows;
        p.n_bits = n_bits;
        p.blowup = blowup;
        p.num_queries = nq;
        p.final_size = fs;
        p.grinding_bits = gb;
        p.cap_height = ch;
        p.value = value;
        return p;
    }

    size_t n_columns() const { return n_bits + 1; }

    // range.rs:172-179
    fri::Params fri_params() const {
        fri::Params p = fri::Params::make(blowup, num_queries, final_size);
        p.grinding_bits = grinding_bits;
        p.cap_height = cap_height;
        return p;
    }

    // range.rs:182-215
    void validate() const {
        if (n_rows < 2 || !is_pow2(n_rows)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        if (n_bits == 0 || n_bits > 24) throw Z0pError(ERR_MALFORMED_INPUT, "n_bits must be within 1..=24");
        if (blowup < 2 || !is_pow2(blowup)) throw Z0pError(ERR_INVALID_BLOWUP);
        if (num_queries == 0) throw Z0pError(ERR_NO_QUERIES);
        if (final_size < 2 || !is_pow2(final_size)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        if (cap_height > 8) throw Z0pError(ERR_MALFORMED_INPUT, "cap_height above 8");
        unsigned __int128 lde = (unsigned __int128)n_rows * blowup;
        if (lde > ((unsigned __int128)1 << Field::TWO_ADICITY)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
    }
};

// range.rs:219-230
struct Opening {
    size_t index = 0;
    Field v;
    std::vector<Field> bits;
    Path path;
};

// range.rs:233-306
struct Proof {
    std::vector<Digest> column_cap;
    fri::Commitment fri_commitment;
    fri::Proof fri_proof;
    std::vector<Opening> openings;

    std::vector<uint8_t> to_bytes() const {
        Writer w;
        w.byte(VERSION);
        w.cap(column_cap);
        fri_commitment.write_body(w);
        fri_proof.write_body(w);
        w.count(openings.size());
        for (auto& o : openings) {
            w.u64_le(o.index);
            w.field(o.v);
            w.count(o.bits.size());
            for (auto& b : o.bits) w.field(b);
            w.path(o.path);
        }
        return w.out;
    }

    static Proof from_bytes(const uint8_t* p, size_t n) {
        Reader r(p, n);
        r.version();
        Proof pr;
        pr.column_cap = r.cap();
        pr.fri_commitment = fri::Commitment::read_body(r);
        pr.fri_proof = fri::Proof::read_body(r);
        size_t no = r.count(MAX_QUERIES);
        for (size_t i = 0; i < no; ++i) {
            Opening o;
            o.index = r.dimension();
            o.v = r.field();
            size_t nb = r.count(MAX_ELEMS);
            for (size_t k = 0; k < nb; ++k) o.bits.push_back(r.field());
            o.path = r.path();
            pr.openings.push_back(std::move(o));
        }
        r.end();
        return pr;
    }
};

// range.rs:312-327 — V = value, C_j = bit j of value (LSB first).
inline std::pair<std::vector<Field>, std::vector<std::vector<Field>>> build_columns(Field value, size_t n_bits,
                                                                                     size_t n_rows) {
    std::vector<Field> v(n_rows, value);
    std::vector<std::vector<Field>> bits;
    for (size_t j = 0; j < n_bits; ++j) {
        Field b = j < 32 ? Field::make((value.v >> j) & 1) : Field::zero();
        bits.emplace_back(n_rows, b);
    }
    return {std::move(v), std::move(bits)};
}

// range.rs:723-735 — 56 bytes, grinding_bits as u32.
inline void absorb_params(const Params& p, Transcript& t) {
    std::vector<uint8_t> buf(56);
    uint8_t* q = buf.data();
    put_u64(q, p.n_rows);
    put_u64(q + 8, p.n_bits);
    put_u64(q + 16, p.blowup);
    put_u64(q + 24, p.num_queries);
    put_u64(q + 32, p.final_size);
    put_u32(q + 40, p.grinding_bits);
    put_u64(q + 44, p.cap_height);
    put_u32(q + 52, p.value.v);
    t.absorb("range/params", buf);
}

// range.rs:739-748
inline std::vector<Field> lde_points(size_t m) {
    Field omega = primitive_root(m);
    std::vector<Field> pts;
    pts.reserve(m);
    Field cur = fri::lde_shift();
    for (size_t i = 0; i < m; ++i) {
        pts.push_back(cur);
        cur *= omega;
    }
    return pts;
}

// range.rs:758-761
inline Field lde_point_at(size_t m, size_t index) { return fri::lde_shift() * primitive_root(m).pow(index); }

// range.rs:752-774
inline Ext4 composition_at(Field v, const std::vector<Field>& bits, Field point, Field value, size_t n,
                           const Ext4& alpha_decomp, const std::vector<Ext4>& alpha_bits, const Ext4& alpha_bound) {
    Field z_inv = (point.pow(n) - Field::one()).inv();
    Field rec = Field::zero(), power = Field::one();
    for (auto& b : bits) {
        rec += power * b;
        power += power;
    }
    Ext4 acc = alpha_decomp.mul_base((v - rec) * z_inv);
    for (size_t j = 0; j < bits.size(); ++j) acc = acc + alpha_bits[j].mul_base(bits[j] * (bits[j] - Field::one()) * z_inv);
    acc = acc + alpha_bound.mul_base((v - value) * (point - Field::one()).inv());
    return acc;
}

// range.rs:767-823 — batched: Z_H by the geometric x^n chain, both
// denominators inverted per 2048-row chunk.
inline std::vector<Ext4> compose(const std::vector<Field>& lde_v, const std::vector<std::vector<Field>>& lde_bits,
                                 const std::vector<Field>& points, Field value, size_t n, const Ext4& alpha_decomp,
                                 const std::vector<Ext4>& alpha_bits, const Ext4& alpha_bound) {
    std::vector<Ext4> out(points.size());
    if (points.size() < 2) {
        for (size_t i = 0; i < points.size(); ++i) {
            std::vector<Field> bits;
            for (auto& col : lde_bits) bits.push_back(col[i]);
            out[i] = composition_at(lde_v[i], bits, points[i], value, n, alpha_decomp, alpha_bits, alpha_bound);
        }
        return out;
    }
    Field p0 = points[0].pow(n);
    Field ratio = (points[1] * points[0].inv()).pow(n);
    for_each_indexed(out.data(), out.size(), 2048, [&](size_t base, Ext4* piece, size_t len) {
        std::vector<Field> den_z(len), den_b(len);
        Field cur = p0 * ratio.pow(base);
        for (size_t j = 0; j < len; ++j) {
            den_z[j] = cur - Field::one();
            den_b[j] = points[base + j] - Field::one();
            cur *= ratio;
        }
        batch_invert_in_place(den_z.data(), len);
        batch_invert_in_place(den_b.data(), len);
        for (size_t j = 0; j < len; ++j) {
            size_t i = base + j;
            Field rec = Field::zero(), power = Field::one();
            for (auto& col : lde_bits) {
                rec += power * col[i];
                power += power;
            }
            Ext4 acc = alpha_decomp.mul_base((lde_v[i] - rec) * den_z[j]);
            for (size_t k = 0; k < lde_bits.size(); ++k) {
                Field b = lde_bits[k][i];
                acc = acc + alpha_bits[k].mul_base(b * (b - Field::one()) * den_z[j]);
            }
            acc = acc + alpha_bound.mul_base((lde_v[i] - value) * den_b[j]);
            piece[j] = acc;
        }
    });
    return out;
}

// range.rs:436-508 — coefficient columns (V first) of any power-of-two length.
inline Proof prove_from_columns(const Params& params, const std::vector<std::vector<Field>>& columns,
                                size_t degree_bound, Transcript& transcript) {
    size_t n = params.n_rows;
    if (columns.size() != params.n_columns())
        throw Z0pError(ERR_MALFORMED_INPUT, "column count must match n_columns");
    absorb_params(params, transcript);
    std::vector<std::vector<Field>> ldes;
    for (auto& col : columns) ldes.push_back(evaluate_on_coset(col, params.blowup, fri::lde_shift()));
    size_t m = ldes[0].size();
    size_t row_len = 4 * params.n_columns();
    std::vector<uint8_t> flat(m * row_len);
    for (size_t i = 0; i < m; ++i)
        for (size_t c = 0; c < ldes.size(); ++c) put_u32(flat.data() + i * row_len + 4 * c, ldes[c][i].v);
    MerkleTree tree = MerkleTree::from_flat(flat.data(), flat.size(), row_len);
    size_t drop = capped_shape(m, params.cap_height).first;
    std::vector<Digest> column_cap = tree.cap(drop);
    transcript.absorb("range/column", cap_bytes(column_cap));
    Ext4 alpha_decomp = transcript.challenge_ext("range/alpha");
    std::vector<Ext4> alpha_bits;
    for (size_t j = 0; j < params.n_bits; ++j) alpha_bits.push_back(transcript.challenge_ext("range/alpha"));
    Ext4 alpha_bound = transcript.challenge_ext("range/alpha");

    std::vector<Field> points = lde_points(m);
    std::vector<std::vector<Field>> bit_ldes(ldes.begin() + 1, ldes.end());
    std::vector<Ext4> composition =
        compose(ldes[0], bit_ldes, points, params.value, n, alpha_decomp, alpha_bits, alpha_bound);
    fri::Params fp = params.fri_params();
    fri::ProverState st = fri::commit_ext_evals_owned(std::move(composition), degree_bound, fp, transcript);
    fri::Commitment fc = fri::public_commitment(st);
    auto opened = fri::open(st, transcript);

    uint32_t log_m = log2_exact(m);
    Proof proof;
    for (size_t leaf : opened.second) {
        size_t index = reverse_bits_n(leaf, log_m);
        Opening o;
        o.index = index;
        o.v = ldes[0][index];
        for (size_t c = 1; c < ldes.size(); ++c) o.bits.push_back(ldes[c][index]);
        o.path = tree.prove_capped(index, drop);
        proof.openings.push_back(std::move(o));
    }
    proof.column_cap = std::move(column_cap);
    proof.fri_commitment = std::move(fc);
    proof.fri_proof = std::move(opened.first);
    return proof;
}

// range.rs:357-382 — explicit evaluation columns, interpolated.
inline Proof prove_with_columns(const Params& params, const std::vector<Field>& v_col,
                                const std::vector<std::vector<Field>>& bit_cols, Transcript& transcript) {
    params.validate();
    size_t n = params.n_rows;
    bool bad = v_col.size() != n || bit_cols.size() != params.n_bits;
    for (auto& c : bit_cols) bad = bad || c.size() != n;
    if (bad) throw Z0pError(ERR_MALFORMED_INPUT, "range columns must match n_rows/n_bits");
    std::vector<std::vector<Field>> coeffs;
    coeffs.push_back(interpolate(v_col));
    for (auto& c : bit_cols) coeffs.push_back(interpolate(c));
    return prove_from_columns(params, coeffs, n - 1, transcript);
}

// range.rs:334-342
inline Proof prove(const Params& params, Transcript& transcript) {
    params.validate();
    auto cols = build_columns(params.value, params.n_bits, params.n_rows);
    return prove_with_columns(params, cols.first, cols.second, transcript);
}

// range.rs:401-428 — every column masked with n_rand maskers (V first);
// bound n + 2 n_rand - 2.
inline Proof prove_zk(const Params& params, size_t n_rand, ChaCha20& rng, Transcript& transcript) {
    params.validate();
    if (n_rand < params.num_queries) throw Z0pError(ERR_MALFORMED_INPUT, "n_rand must cover num_queries for hiding");
    size_t n = params.n_rows;
    auto cols = build_columns(params.value, params.n_bits, n);
    std::vector<std::vector<Field>> coeffs;
    coeffs.push_back(interpolate(cols.first));
    for (auto& c : cols.second) coeffs.push_back(interpolate(c));
    std::vector<std::vector<Field>> masked;
    for (auto& c : coeffs) masked.push_back(mask_coeffs(c, n, n_rand, rng));
    return prove_from_columns(params, masked, n + 2 * n_rand - 2, transcript);
}

// range.rs:521-721
inline void verify(const Params& params, const Proof& proof, Transcript& transcript) {
    params.validate();
    size_t n = params.n_rows;
    size_t m = proof.fri_commitment.lde_len;
    size_t expected_m = n * params.blowup;
    if (m < expected_m || m % expected_m != 0) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    size_t degree_bound = proof.fri_commitment.degree_bound;
    if (m == expected_m) {
        if (degree_bound != n - 1) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    } else {
        size_t padded = m / params.blowup;
        if (2 * padded < n + 2) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
        if (degree_bound > 2 * padded - (n + 2)) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    }
    if (proof.fri_commitment.poly_len != m) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    absorb_params(params, transcript);
    if (proof.openings.size() != params.num_queries) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    transcript.absorb("range/column", cap_bytes(proof.column_cap));
    Ext4 alpha_decomp = transcript.challenge_ext("range/alpha");
    std::vector<Ext4> alpha_bits;
    for (size_t j = 0; j < params.n_bits; ++j) alpha_bits.push_back(transcript.challenge_ext("range/alpha"));
    Ext4 alpha_bound = transcript.challenge_ext("range/alpha");
    fri::Params fp = params.fri_params();
    std::vector<size_t> idx = fri::verify(proof.fri_commitment, proof.fri_proof, fp, transcript);
    uint32_t log_m = log2_exact(m);
    size_t count = std::min(std::min(proof.fri_proof.queries.size(), proof.openings.size()), idx.size());
    for (size_t q = 0; q < count; ++q) {
        const Opening& o = proof.openings[q];
        size_t leaf = idx[q];
        size_t index = reverse_bits_n(leaf, log_m);
        if (o.index != index || o.bits.size() != params.n_bits) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
        std::vector<uint8_t> row(4 * params.n_columns());
        put_u32(row.data(), o.v.v);
        for (size_t j = 0; j < o.bits.size(); ++j) put_u32(row.data() + 4 * (j + 1), o.bits[j].v);
        auto [drop, depth] = capped_shape(m, params.cap_height);
        if (proof.column_cap.size() != (size_t(1) << drop) || o.path.size() != depth - drop)
            throw Z0pError(ERR_BAD_MERKLE_PATH);
        size_t e = index >> (depth - drop);
        if (e >= proof.column_cap.size()) throw Z0pError(ERR_BAD_MERKLE_PATH);
        if (!verify_capped(proof.column_cap[e], row.data(), row.size(), index, depth, o.path))
            throw Z0pError(ERR_BAD_MERKLE_PATH);
        Ext4 expected = composition_at(o.v, o.bits, lde_point_at(m, index), params.value, n, alpha_decomp, alpha_bits,
                                       alpha_bound);
        const auto& query = proof.fri_proof.queries[q];
        if (query.folds.empty()) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
        if (query.folds[0].values[leaf % 4] != expected) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    }
}

}  // namespace range
}  // namespace oracle
