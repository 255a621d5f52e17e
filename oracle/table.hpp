// ORACLE — test infrastructure only (see bb.hpp).
//
// Table-membership AIR prover/verifier (plain and zero-knowledge), restated
// from src/table.rs: one value column V, one tree leaf per LDE value
// (BLAKE3(0x00 || V LE)), and the composition
//   a_m prod_{t in T} (V - t) / Z_H + a_b (V - v) / (x - 1)
// committed by fri::commit_ext_evals under the bound |T| (n + n_rand - 1) - n.
#pragma once
#include <vector>

#include "blind.hpp"
#include "fri.hpp"

namespace oracle {
namespace table {

// table.rs:27-46
struct Params {
    size_t n_rows = 0, blowup = 0, num_queries = 0, final_size = 0;
    uint32_t grinding_bits = 0;
    size_t cap_height = 0;
    Field value;

    // table.rs:53-87
    static Params make(size_t n_rows, size_t blowup, size_t num_queries, size_t final_size, Field value) {
        if (n_rows < 2 || !is_pow2(n_rows)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        if (blowup < 2 || !is_pow2(blowup)) throw Z0pError(ERR_INVALID_BLOWUP);
        if (num_queries == 0) throw Z0pError(ERR_NO_QUERIES);
        if (final_size < 2 || !is_pow2(final_size)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        unsigned __int128 lde = (unsigned __int128)n_rows * blowup;
        if (lde > ((unsigned __int128)1 << Field::TWO_ADICITY)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        Params p;
        p.n_rows = n_rows;
        p.blowup = blowup;
        p.num_queries = num_queries;
        p.final_size = final_size;
        p.value = value;
        return p;
    }

    // table.rs:114-168 — the smallest blowup >= 4 with (bound + 1) <= lde / 4.
    static Params for_security(size_t n_rows, size_t table_len, Field value, uint32_t target_bits) {
        if (n_rows < 2 || !is_pow2(n_rows)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        if (table_len < 2) throw Z0pError(ERR_MALFORMED_INPUT, "table needs at least two entries");
        unsigned __int128 prod = (unsigned __int128)table_len * (n_rows - 1);
        if (prod > SIZE_MAX || prod < n_rows) throw Z0pError(ERR_MALFORMED_INPUT, "table length overflows bound");
        size_t bound = size_t(prod) - n_rows;
        size_t need = bound == SIZE_MAX ? SIZE_MAX : bound + 1;
        size_t blowup = 4, lde_len = 0;
        while (true) {
            unsigned __int128 lde = (unsigned __int128)n_rows * blowup;
            if (lde > SIZE_MAX) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
            lde_len = size_t(lde);
            if (need <= lde_len / 4 || blowup >= (size_t(1) << 27)) break;
            blowup *= 2;
        }
        if (lde_len > (size_t(1) << Field::TWO_ADICITY) || need > lde_len / 4) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        uint32_t mb = fri::composition_millibits(lde_len, bound);
        auto [nq, gb] = fri::tune_strength(target_bits, mb);
        auto [fs, ch] = fri::tune_shape(lde_len, nq, {{lde_len, 12}});
        Params p;
        p.n_rows = n_rows;
        p.blowup = blowup;
        p.num_queries = nq;
        p.final_size = fs;
        p.grinding_bits = gb;
        p.cap_height = ch;
        p.value = value;
        return p;
    }

    // table.rs:175-181
    fri::Params fri_params() const {
        fri::Params p = fri::Params::make(blowup, num_queries, final_size);
        p.grinding_bits = grinding_bits;
        p.cap_height = cap_height;
        return p;
    }

    // table.rs:184-212
    void validate() const {
        if (n_rows < 2 || !is_pow2(n_rows)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        if (blowup < 2 || !is_pow2(blowup)) throw Z0pError(ERR_INVALID_BLOWUP);
        if (num_queries == 0) throw Z0pError(ERR_NO_QUERIES);
        if (final_size < 2 || !is_pow2(final_size)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        if (cap_height > 8) throw Z0pError(ERR_MALFORMED_INPUT, "cap_height above 8");
        unsigned __int128 lde = (unsigned __int128)n_rows * blowup;
        if (lde > ((unsigned __int128)1 << Field::TWO_ADICITY)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
    }
};

// table.rs:216-224
struct Opening {
    size_t index = 0;
    Field v;
    Path path;
};

// table.rs:227-288
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
            o.path = r.path();
            pr.openings.push_back(std::move(o));
        }
        r.end();
        return pr;
    }
};

// table.rs:291-298
inline void check_table(const std::vector<Field>& table) {
    if (table.size() < 2) throw Z0pError(ERR_MALFORMED_INPUT, "table needs at least two entries");
}

// table.rs:357-363 — saturating |T| (n + n_rand - 1) - n.
inline size_t membership_bound(size_t table_len, size_t n, size_t n_rand) {
    auto sat_add = [](size_t a, size_t b) { return a > SIZE_MAX - b ? SIZE_MAX : a + b; };
    auto sat_sub = [](size_t a, size_t b) { return a < b ? 0 : a - b; };
    auto sat_mul = [](size_t a, size_t b) {
        unsigned __int128 p = (unsigned __int128)a * b;
        return p > SIZE_MAX ? SIZE_MAX : size_t(p);
    };
    return sat_sub(sat_mul(table_len, sat_sub(sat_add(n, n_rand), 1)), n);
}

// table.rs:540-555 — 56 bytes of parameters, then |T| and the entries.
inline void absorb_params(const Params& p, const std::vector<Field>& table, Transcript& t) {
    std::vector<uint8_t> buf(56 + 4 * table.size());
    uint8_t* q = buf.data();
    put_u64(q, p.n_rows);
    put_u64(q + 8, p.blowup);
    put_u64(q + 16, p.num_queries);
    put_u64(q + 24, p.final_size);
    put_u32(q + 32, p.grinding_bits);
    put_u64(q + 36, p.cap_height);
    put_u32(q + 44, p.value.v);
    put_u64(q + 48, table.size());
    for (size_t k = 0; k < table.size(); ++k) put_u32(q + 56 + 4 * k, table[k].v);
    t.absorb("table/params", buf);
}

// table.rs:557-579
inline Field lde_point_at(size_t m, size_t index) { return fri::lde_shift() * primitive_root(m).pow(index); }

// table.rs:646-663
inline Ext4 composition_at(Field v, Field point, const std::vector<Field>& table, Field value, size_t n,
                           const Ext4& alpha_member, const Ext4& alpha_bound) {
    Field z_inv = (point.pow(n) - Field::one()).inv();
    Field product = Field::one();
    for (auto& t : table) product *= v - t;
    return alpha_member.mul_base(product * z_inv) + alpha_bound.mul_base((v - value) * (point - Field::one()).inv());
}

// table.rs:597-644 — batched per 2048-point chunk, as composition_at.
inline std::vector<Ext4> compose(const std::vector<Field>& lde, const std::vector<Field>& points,
                                 const std::vector<Field>& table, Field value, size_t n, const Ext4& alpha_member,
                                 const Ext4& alpha_bound) {
    std::vector<Ext4> out(points.size());
    if (points.size() < 2) {
        for (size_t i = 0; i < points.size(); ++i)
            out[i] = composition_at(lde[i], points[i], table, value, n, alpha_member, alpha_bound);
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
            Field product = Field::one();
            for (auto& t : table) product *= lde[i] - t;
            piece[j] = alpha_member.mul_base(product * den_z[j]) + alpha_bound.mul_base((lde[i] - value) * den_b[j]);
        }
    });
    return out;
}

// table.rs:370-428
inline Proof prove_from_coeffs(const Params& params, const std::vector<Field>& table, const std::vector<Field>& coeffs,
                               size_t degree_bound, Transcript& transcript) {
    params.validate();
    absorb_params(params, table, transcript);
    std::vector<Field> lde = evaluate_on_coset(coeffs, params.blowup, fri::lde_shift());
    std::vector<uint8_t> flat(lde.size() * 4);
    for (size_t i = 0; i < lde.size(); ++i) put_u32(flat.data() + 4 * i, lde[i].v);
    MerkleTree tree = MerkleTree::from_flat(flat.data(), flat.size(), 4);
    size_t m = lde.size();
    size_t drop = capped_shape(m, params.cap_height).first;
    std::vector<Digest> column_cap = tree.cap(drop);
    transcript.absorb("table/column", cap_bytes(column_cap));
    Ext4 alpha_member = transcript.challenge_ext("table/alpha");
    Ext4 alpha_bound = transcript.challenge_ext("table/alpha");
    std::vector<Field> points(m);
    {
        Field omega = primitive_root(m), cur = fri::lde_shift();
        for (size_t i = 0; i < m; ++i, cur *= omega) points[i] = cur;
    }
    std::vector<Ext4> composition = compose(lde, points, table, params.value, params.n_rows, alpha_member, alpha_bound);
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
        o.v = lde[index];
        o.path = tree.prove_capped(index, drop);
        proof.openings.push_back(std::move(o));
    }
    proof.column_cap = std::move(column_cap);
    proof.fri_commitment = std::move(fc);
    proof.fri_proof = std::move(opened.first);
    return proof;
}

// table.rs:305-318
inline Proof prove(const Params& params, const std::vector<Field>& table, Transcript& transcript) {
    params.validate();
    check_table(table);
    std::vector<Field> coeffs = interpolate(std::vector<Field>(params.n_rows, params.value));
    return prove_from_coeffs(params, table, coeffs, membership_bound(table.size(), params.n_rows, 0), transcript);
}

// table.rs:329-355
inline Proof prove_zk(const Params& params, const std::vector<Field>& table, size_t n_rand, ChaCha20& rng,
                      Transcript& transcript) {
    params.validate();
    check_table(table);
    if (n_rand < params.num_queries) throw Z0pError(ERR_MALFORMED_INPUT, "n_rand must cover num_queries for hiding");
    size_t n = params.n_rows;
    std::vector<Field> base = interpolate(std::vector<Field>(n, params.value));
    std::vector<Field> masked = mask_coeffs(base, n, n_rand, rng);
    return prove_from_coeffs(params, table, masked, membership_bound(table.size(), n, n_rand), transcript);
}

// table.rs:430-538
inline void verify(const Params& params, const std::vector<Field>& table, const Proof& proof, Transcript& transcript) {
    params.validate();
    check_table(table);
    size_t n = params.n_rows;
    size_t m = proof.fri_commitment.lde_len;
    size_t expected_m = n * params.blowup;
    if (m < expected_m || m % expected_m != 0) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    size_t degree_bound = proof.fri_commitment.degree_bound;
    if (m == expected_m) {
        if (degree_bound != membership_bound(table.size(), n, 0)) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    } else {
        size_t padded = m / params.blowup;
        unsigned __int128 prod = (unsigned __int128)table.size() * (padded == 0 ? 0 : padded - 1);
        if (prod > SIZE_MAX || prod < n) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
        if (degree_bound > size_t(prod) - n) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    }
    if (proof.fri_commitment.poly_len != m) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    absorb_params(params, table, transcript);
    if (proof.openings.size() != params.num_queries) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    transcript.absorb("table/column", cap_bytes(proof.column_cap));
    Ext4 alpha_member = transcript.challenge_ext("table/alpha");
    Ext4 alpha_bound = transcript.challenge_ext("table/alpha");
    fri::Params fp = params.fri_params();
    std::vector<size_t> idx = fri::verify(proof.fri_commitment, proof.fri_proof, fp, transcript);
    uint32_t log_m = log2_exact(m);
    size_t count = std::min(std::min(proof.fri_proof.queries.size(), proof.openings.size()), idx.size());
    for (size_t q = 0; q < count; ++q) {
        const Opening& o = proof.openings[q];
        size_t leaf = idx[q];
        size_t index = reverse_bits_n(leaf, log_m);
        if (o.index != index) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
        auto [drop, depth] = capped_shape(m, params.cap_height);
        if (proof.column_cap.size() != (size_t(1) << drop) || o.path.size() != depth - drop)
            throw Z0pError(ERR_BAD_MERKLE_PATH);
        size_t e = index >> (depth - drop);
        if (e >= proof.column_cap.size()) throw Z0pError(ERR_BAD_MERKLE_PATH);
        uint8_t vb[4];
        put_u32(vb, o.v.v);
        if (!verify_capped(proof.column_cap[e], vb, 4, index, depth, o.path)) throw Z0pError(ERR_BAD_MERKLE_PATH);
        Ext4 expected =
            composition_at(o.v, lde_point_at(m, index), table, params.value, n, alpha_member, alpha_bound);
        const auto& query = proof.fri_proof.queries[q];
        if (query.folds.empty()) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
        if (query.folds[0].values[leaf % 4] != expected) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    }
}

}  // namespace table
}  // namespace oracle
