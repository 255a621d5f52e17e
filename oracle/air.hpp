// ORACLE — test infrastructure only (see bb.hpp).
//
// Fibonacci AIR prover/verifier (plain and zero-knowledge), restated from src/air.rs.
#pragma once
#include <vector>

#include "blind.hpp"
#include "fri.hpp"

namespace oracle {
namespace air {

// air.rs:31-54
struct Params {
    size_t n_steps = 0, blowup = 0, num_queries = 0, final_size = 0;
    uint32_t grinding_bits = 0;
    size_t cap_height = 0;
    Field a0, b0, b_last;

    // air.rs:63-101
    static Params make(size_t n_steps, size_t blowup, size_t num_queries, size_t final_size, Field a0, Field b0,
                       Field b_last) {
        if (n_steps < 2 || !is_pow2(n_steps)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        if (blowup < 4 || !is_pow2(blowup)) throw Z0pError(ERR_INVALID_BLOWUP);
        if (num_queries == 0) throw Z0pError(ERR_NO_QUERIES);
        if (final_size < 2 || !is_pow2(final_size)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        if (n_steps * blowup > (size_t(1) << Field::TWO_ADICITY)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        Params p;
        p.n_steps = n_steps;
        p.blowup = blowup;
        p.num_queries = num_queries;
        p.final_size = final_size;
        p.a0 = a0;
        p.b0 = b0;
        p.b_last = b_last;
        return p;
    }

    // air.rs:129-162
    static Params for_security(size_t n_steps, Field a0, Field b0, Field b_last, uint32_t target_bits) {
        if (n_steps < 2 || !is_pow2(n_steps)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        size_t blowup = 16;
        size_t lde_len = n_steps * blowup;
        if (lde_len > (size_t(1) << Field::TWO_ADICITY)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        uint32_t mb = fri::composition_millibits(lde_len, n_steps - 1);
        auto [nq, gb] = fri::tune_strength(target_bits, mb);
        auto [fs, ch] = fri::tune_shape(lde_len, nq, {{lde_len, 24}});
        Params p;
        p.n_steps = n_steps;
        p.blowup = blowup;
        p.num_queries = nq;
        p.final_size = fs;
        p.grinding_bits = gb;
        p.cap_height = ch;
        p.a0 = a0;
        p.b0 = b0;
        p.b_last = b_last;
        return p;
    }

    size_t lde_len() const { return n_steps * blowup; }

    // air.rs:170-176
    fri::Params fri_params() const {
        fri::Params p = fri::Params::make(blowup, num_queries, final_size);
        p.grinding_bits = grinding_bits;
        p.cap_height = cap_height;
        return p;
    }

    // air.rs:181-211
    void validate() const {
        if (n_steps < 2 || !is_pow2(n_steps)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        if (blowup < 4 || !is_pow2(blowup)) throw Z0pError(ERR_INVALID_BLOWUP);
        if (num_queries == 0) throw Z0pError(ERR_NO_QUERIES);
        if (final_size < 2 || !is_pow2(final_size)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        if (cap_height > 8) throw Z0pError(ERR_MALFORMED_INPUT, "cap_height above 8");
        unsigned __int128 lde = (unsigned __int128)n_steps * blowup;
        if (lde > ((unsigned __int128)1 << Field::TWO_ADICITY)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
    }
};

// air.rs:223-246
struct TraceOpening {
    size_t index = 0;
    Field a, b, a_next, b_next;
    Path path, shift_path;
    uint8_t leaf[64];
    uint8_t shift_leaf[64];
};

// air.rs:250-334
struct Proof {
    std::vector<Digest> column_cap;
    fri::Commitment fri_commitment;
    fri::Proof fri_proof;
    std::vector<TraceOpening> openings;

    std::vector<uint8_t> to_bytes() const {
        Writer w;
        w.byte(VERSION);
        w.cap(column_cap);
        fri_commitment.write_body(w);
        fri_proof.write_body(w);
        w.count(openings.size());
        for (auto& o : openings) {
            w.u64_le(o.index);
            w.field(o.a);
            w.field(o.b);
            w.field(o.a_next);
            w.field(o.b_next);
            w.path(o.path);
            w.path(o.shift_path);
            w.bytes(o.leaf, 64);
            w.bytes(o.shift_leaf, 64);
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
            TraceOpening o;
            o.index = r.dimension();
            o.a = r.field();
            o.b = r.field();
            o.a_next = r.field();
            o.b_next = r.field();
            o.path = r.path();
            o.shift_path = r.path();
            std::memcpy(o.leaf, r.take(64), 64);
            std::memcpy(o.shift_leaf, r.take(64), 64);
            pr.openings.push_back(std::move(o));
        }
        r.end();
        return pr;
    }
};

// air.rs:341-354
inline std::pair<std::vector<Field>, std::vector<Field>> build_trace(Field a0, Field b0, size_t n) {
    if (n == 0) return {};
    std::vector<Field> a(n), b(n);
    a[0] = a0;
    b[0] = b0;
    for (size_t i = 0; i + 1 < n; ++i) {
        a[i + 1] = b[i];
        b[i + 1] = a[i] + b[i];
    }
    return {std::move(a), std::move(b)};
}

// air.rs:754-766 — 56 bytes, grinding_bits as u32.
inline void absorb_params(const Params& p, Transcript& t) {
    std::vector<uint8_t> buf(56);
    uint8_t* q = buf.data();
    put_u64(q, p.n_steps);
    put_u64(q + 8, p.blowup);
    put_u64(q + 16, p.num_queries);
    put_u64(q + 24, p.final_size);
    put_u32(q + 32, p.grinding_bits);
    put_u64(q + 36, p.cap_height);
    put_u32(q + 44, p.a0.v);
    put_u32(q + 48, p.b0.v);
    put_u32(q + 52, p.b_last.v);
    t.absorb("fib/params", buf);
}

// air.rs:770-779 — serial chain SHIFT * omega^i.
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

// air.rs:789-792
inline Field lde_point_at(size_t m, size_t index) { return fri::lde_shift() * primitive_root(m).pow(index); }

// air.rs:806-844
inline std::vector<Field> linked_lde(const std::vector<Field>& lde_a, size_t n, Field actual_b_last, Field a0) {
    size_t m = lde_a.size();
    if (n == 0 || !is_pow2(n) || m % n != 0) throw Z0pError(ERR_MALFORMED_INPUT, "linked LDE dimensions must be powers of two");
    size_t offset = m / n;
    Field omega_m = primitive_root(m);
    Field omega_n = primitive_root(n);
    Field w_last = omega_n.pow(n - 1);
    Field ratio = omega_m.pow(n);
    Field p0 = fri::lde_shift().pow(n);
    Field n_inv = Field::from_u64(n).inv();
    Field coeff = (actual_b_last - a0) * w_last * n_inv;
    std::vector<Field> out(m);
    for_each_indexed(out.data(), m, 2048, [&](size_t base, Field* piece, size_t len) {
        Field x = fri::lde_shift() * omega_m.pow(base);
        Field xn = p0 * ratio.pow(base);
        std::vector<Field> den(len), van(len);
        for (size_t j = 0; j < len; ++j) {
            den[j] = x - w_last;
            van[j] = xn - Field::one();
            x *= omega_m;
            xn *= ratio;
        }
        batch_invert_in_place(den.data(), len);
        for (size_t j = 0; j < len; ++j) {
            size_t i = base + j;
            size_t sh = i + offset;
            if (sh >= m) sh -= m;
            piece[j] = lde_a[sh] + coeff * van[j] * den[j];
        }
    });
    return out;
}

// air.rs:938-962
inline Ext4 composition_at(const Field cols[2], const Field next[2], Field point, Field w_last, Field a0, Field b0,
                           Field b_last, size_t n, const std::array<Ext4, 5>& al) {
    Field vanishing = point.pow(n) - Field::one();
    Field z_trans = vanishing * (point - w_last).inv();
    Field trans1 = (next[0] - cols[1]) * z_trans.inv();
    Field trans2 = (next[1] - (cols[0] + cols[1])) * z_trans.inv();
    Field bound_a = (cols[0] - a0) * (point - Field::one()).inv();
    Field bound_b = (cols[1] - b0) * (point - Field::one()).inv();
    Field bound_out = (cols[1] - b_last) * (point - w_last).inv();
    return al[0].mul_base(trans1) + al[1].mul_base(trans2) + al[2].mul_base(bound_a) + al[3].mul_base(bound_b) +
           al[4].mul_base(bound_out);
}

// air.rs:847-929
inline std::vector<Ext4> compose(const std::vector<Field>* cols, const std::vector<Field>& points, Field w_last, Field a0,
                                 Field b0, Field b_last, size_t n, size_t offset, const std::array<Ext4, 5>& al) {
    size_t m = cols[0].size();
    std::vector<Ext4> out(points.size());
    if (points.size() < 2) {
        for (size_t i = 0; i < points.size(); ++i) {
            size_t nx = i + offset;
            if (nx >= m) nx -= m;
            Field c[2] = {cols[0][i], cols[1][i]};
            Field nn[2] = {cols[0][nx], cols[1][nx]};
            out[i] = composition_at(c, nn, points[i], w_last, a0, b0, b_last, n, al);
        }
        return out;
    }
    Field p0 = points[0].pow(n);
    Field ratio = (points[1] * points[0].inv()).pow(n);
    for_each_indexed(out.data(), out.size(), 2048, [&](size_t base, Ext4* piece, size_t len) {
        std::vector<Field> den_a(len), den_b(len), inv_z(len);
        Field cur = p0 * ratio.pow(base);
        for (size_t j = 0; j < len; ++j) {
            size_t i = base + j;
            den_a[j] = points[i] - w_last;
            den_b[j] = points[i] - Field::one();
            inv_z[j] = cur - Field::one();
            cur *= ratio;
        }
        batch_invert_in_place(den_a.data(), len);
        batch_invert_in_place(den_b.data(), len);
        for (size_t j = 0; j < len; ++j) inv_z[j] *= den_a[j];
        batch_invert_in_place(inv_z.data(), len);
        for (size_t j = 0; j < len; ++j) {
            size_t i = base + j;
            size_t nx = i + offset;
            if (nx >= m) nx -= m;
            Field trans1 = (cols[0][nx] - cols[1][i]) * inv_z[j];
            Field trans2 = (cols[1][nx] - (cols[0][i] + cols[1][i])) * inv_z[j];
            Field bound_a = (cols[0][i] - a0) * den_b[j];
            Field bound_b = (cols[1][i] - b0) * den_b[j];
            Field bound_out = (cols[1][i] - b_last) * den_a[j];
            piece[j] = al[0].mul_base(trans1) + al[1].mul_base(trans2) + al[2].mul_base(bound_a) +
                       al[3].mul_base(bound_b) + al[4].mul_base(bound_out);
        }
    });
    return out;
}

// air.rs:964-971
inline size_t bit_reverse_index(size_t x, uint32_t bits) { return reverse_bits_n(x, bits); }

// air.rs:485-599 (coeff_b == nullptr selects the linked path, air.rs:408-422).
inline Proof prove_inner(const Params& params, const std::vector<Field>& coeff_a, const std::vector<Field>* coeff_b,
                         size_t degree_bound, const Field* linked_b_last, Transcript& transcript) {
    params.validate();
    size_t n = params.n_steps;
    absorb_params(params, transcript);
    Field omega_base = primitive_root(n);
    std::vector<Field> lde_cols[2];
    {
        ScopedStage s("lde");
        lde_cols[0] = evaluate_on_coset(coeff_a, params.blowup, fri::lde_shift());
    }
    size_t m = lde_cols[0].size();
    {
        ScopedStage s("linked_b");
        if (coeff_b) {
            lde_cols[1] = evaluate_on_coset(*coeff_b, params.blowup, fri::lde_shift());
        } else {
            if (!linked_b_last) throw Z0pError(ERR_MALFORMED_INPUT, "linked trace needs its final column value");
            lde_cols[1] = linked_lde(lde_cols[0], n, *linked_b_last, params.a0);
        }
    }
    if (m % n != 0) throw Z0pError(ERR_MALFORMED_INPUT, "LDE length must be a multiple of n_steps");
    size_t offset = m / n;
    size_t row_len = 8;
    std::vector<uint8_t> flat(m * row_len);
    MerkleTree tree;
    {
        ScopedStage s("trace_tree");
        for (size_t i = 0; i < m; ++i) {
            put_u32(flat.data() + i * row_len, lde_cols[0][i].v);
            put_u32(flat.data() + i * row_len + 4, lde_cols[1][i].v);
        }
        tree = MerkleTree::from_flat64(flat.data(), flat.size());
    }
    size_t leaves = m / 8;
    size_t drop = capped_shape(leaves, params.cap_height).first;
    std::vector<Digest> column_cap = tree.cap(drop);
    transcript.absorb("fib/trace", cap_bytes(column_cap));
    std::array<Ext4, 5> alphas;
    for (auto& a : alphas) a = transcript.challenge_ext("fib/alpha");

    std::vector<Ext4> composition;
    {
        ScopedStage s("compose");
        std::vector<Field> points = lde_points(m);
        Field w_last = omega_base.pow(n - 1);
        composition = compose(lde_cols, points, w_last, params.a0, params.b0, params.b_last, n, offset, alphas);
    }
    fri::Params fp = params.fri_params();
    fri::ProverState st = fri::commit_ext_evals_owned(std::move(composition), degree_bound, fp, transcript);
    fri::Commitment fc = fri::public_commitment(st);
    auto opened = fri::open(st, transcript);

    ScopedStage s("trace_open");
    uint32_t log_m = log2_exact(m);
    Proof proof;
    for (size_t leaf : opened.second) {
        size_t index = bit_reverse_index(leaf, log_m);
        size_t shifted = (index + offset) % m;
        size_t base = index & ~size_t(7);
        size_t shift_base = shifted & ~size_t(7);
        TraceOpening o;
        o.index = index;
        o.a = lde_cols[0][index];
        o.b = lde_cols[1][index];
        o.a_next = lde_cols[0][shifted];
        o.b_next = lde_cols[1][shifted];
        o.path = tree.prove_capped(base / 8, drop);
        o.shift_path = tree.prove_capped(shift_base / 8, drop);
        std::memcpy(o.leaf, flat.data() + base * row_len, 64);
        std::memcpy(o.shift_leaf, flat.data() + shift_base * row_len, 64);
        proof.openings.push_back(std::move(o));
    }
    proof.column_cap = std::move(column_cap);
    proof.fri_commitment = std::move(fc);
    proof.fri_proof = std::move(opened.first);
    return proof;
}

// air.rs:365-372
inline Proof prove(const Params& params, Transcript& transcript) {
    params.validate();
    std::vector<Field> coeff_a;
    Field b_last_actual;
    {
        ScopedStage s("trace_interpolate");
        auto cols = build_trace(params.a0, params.b0, params.n_steps);
        coeff_a = interpolate(cols.first);
        b_last_actual = cols.second[params.n_steps - 1];
    }
    return prove_inner(params, coeff_a, nullptr, params.n_steps - 1, &b_last_actual, transcript);
}

// air.rs:385-406 — general path with two independent NTTs.
inline Proof prove_with_trace(const Params& params, const std::vector<Field>& col_a, const std::vector<Field>& col_b,
                              Transcript& transcript) {
    params.validate();
    size_t n = params.n_steps;
    if (col_a.size() != n || col_b.size() != n) throw Z0pError(ERR_MALFORMED_INPUT, "trace columns must match n_steps");
    std::vector<Field> ca = interpolate(col_a), cb = interpolate(col_b);
    return prove_inner(params, ca, &cb, n - 1, nullptr, transcript);
}

// air.rs:439-464 — both columns masked with n_rand maskers each (A first).
inline Proof prove_zk(const Params& params, size_t n_rand, ChaCha20& rng, Transcript& transcript) {
    params.validate();
    if (params.num_queries > SIZE_MAX / 16) throw Z0pError(ERR_MALFORMED_INPUT, "query budget overflows the hiding requirement");
    if (n_rand < params.num_queries * 16)
        throw Z0pError(ERR_MALFORMED_INPUT, "n_rand must cover every revealed leaf row for hiding");
    size_t n = params.n_steps;
    auto cols = build_trace(params.a0, params.b0, n);
    std::vector<Field> ca = interpolate(cols.first), cb = interpolate(cols.second);
    std::vector<Field> ma = mask_coeffs(ca, n, n_rand, rng);
    std::vector<Field> mb = mask_coeffs(cb, n, n_rand, rng);
    return prove_inner(params, ma, &mb, n + n_rand - 1, nullptr, transcript);
}

// air.rs:607-752
inline void verify(const Params& params, const Proof& proof, Transcript& transcript) {
    params.validate();
    size_t n = params.n_steps;
    size_t m = proof.fri_commitment.lde_len;
    size_t expected_m = n * params.blowup;
    if (m < expected_m || m % expected_m != 0) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    size_t degree_bound = proof.fri_commitment.degree_bound;
    if (m == expected_m) {
        if (degree_bound != n - 1) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    } else if (degree_bound >= m / params.blowup) {
        throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    }
    if (proof.fri_commitment.poly_len != m) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    size_t offset = m / n;
    absorb_params(params, transcript);
    if (proof.openings.size() != params.num_queries) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    transcript.absorb("fib/trace", cap_bytes(proof.column_cap));
    std::array<Ext4, 5> alphas;
    for (auto& a : alphas) a = transcript.challenge_ext("fib/alpha");
    fri::Params fp = params.fri_params();
    std::vector<size_t> idx = fri::verify(proof.fri_commitment, proof.fri_proof, fp, transcript);
    Field omega_base = primitive_root(n);
    Field w_last = omega_base.pow(n - 1);
    uint32_t log_m = log2_exact(m);
    size_t count = std::min(std::min(proof.fri_proof.queries.size(), proof.openings.size()), idx.size());
    for (size_t q = 0; q < count; ++q) {
        const TraceOpening& o = proof.openings[q];
        size_t leaf = idx[q];
        size_t index = bit_reverse_index(leaf, log_m);
        if (o.index != index) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
        if (m % 8 != 0) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
        size_t leaves = m / 8;
        auto [drop, depth] = capped_shape(leaves, params.cap_height);
        if (proof.column_cap.size() != (size_t(1) << drop) || o.path.size() != depth - drop ||
            o.shift_path.size() != depth - drop)
            throw Z0pError(ERR_BAD_MERKLE_PATH);
        size_t ro = (index & 7) * 8;
        if (get_u32(o.leaf + ro) != o.a.v || get_u32(o.leaf + ro + 4) != o.b.v) throw Z0pError(ERR_BAD_MERKLE_PATH);
        size_t e = (index / 8) >> (depth - drop);
        if (e >= proof.column_cap.size()) throw Z0pError(ERR_BAD_MERKLE_PATH);
        if (!verify_capped64(proof.column_cap[e], o.leaf, index / 8, depth, o.path)) throw Z0pError(ERR_BAD_MERKLE_PATH);
        size_t shifted = (index + offset) % m;
        size_t so = (shifted & 7) * 8;
        if (get_u32(o.shift_leaf + so) != o.a_next.v || get_u32(o.shift_leaf + so + 4) != o.b_next.v)
            throw Z0pError(ERR_BAD_MERKLE_PATH);
        size_t se = (shifted / 8) >> (depth - drop);
        if (se >= proof.column_cap.size()) throw Z0pError(ERR_BAD_MERKLE_PATH);
        if (!verify_capped64(proof.column_cap[se], o.shift_leaf, shifted / 8, depth, o.shift_path))
            throw Z0pError(ERR_BAD_MERKLE_PATH);
        Field cols[2] = {o.a, o.b};
        Field next[2] = {o.a_next, o.b_next};
        Ext4 expected = composition_at(cols, next, lde_point_at(m, index), w_last, params.a0, params.b0, params.b_last, n,
                                       alphas);
        const auto& query = proof.fri_proof.queries[q];
        if (query.folds.empty()) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
        if (query.folds[0].values[leaf % 4] != expected) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    }
}

}  // namespace air
}  // namespace oracle
