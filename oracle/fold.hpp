// ORACLE — test infrastructure only (see bb.hpp).
//
// Hash-based folding / IVC accumulation over Fibonacci segments, restated
// from src/fold.rs: every segment of seg_steps rows is committed with its own
// (A, B) row tree (one leaf per LDE row, BLAKE3(0x00 || A || B)) and six
// composition alphas; finalize draws one gamma per segment and commits
// sum gamma_k C_k through one fri::commit_ext_evals.
#pragma once
#include <array>
#include <vector>

#include "fri.hpp"

namespace oracle {
namespace fold {

constexpr size_t ALPHAS = 6;         // fold.rs:38
constexpr uint32_t MAX_STEPS = 1u << 20;  // codec.rs:34

// fold.rs:41-65
struct Params {
    size_t seg_steps = 0, blowup = 0, num_queries = 0, final_size = 0;
    uint32_t grinding_bits = 0;
    size_t cap_height = 0;
    Field in_a, in_b, out_b;

    // fold.rs:72-111
    static Params make(size_t seg_steps, size_t blowup, size_t num_queries, size_t final_size, Field in_a, Field in_b,
                       Field out_b) {
        if (seg_steps < 2 || !is_pow2(seg_steps)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        if (blowup < 4 || !is_pow2(blowup)) throw Z0pError(ERR_INVALID_BLOWUP);
        if (num_queries == 0) throw Z0pError(ERR_NO_QUERIES);
        if (final_size < 2 || !is_pow2(final_size)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        unsigned __int128 lde = (unsigned __int128)seg_steps * blowup;
        if (lde > ((unsigned __int128)1 << Field::TWO_ADICITY)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        Params p;
        p.seg_steps = seg_steps;
        p.blowup = blowup;
        p.num_queries = num_queries;
        p.final_size = final_size;
        p.in_a = in_a;
        p.in_b = in_b;
        p.out_b = out_b;
        return p;
    }

    // fold.rs:136-170 — blowup 16; step openings carry 24 bytes.
    static Params for_security(size_t seg_steps, Field in_a, Field in_b, Field out_b, uint32_t target_bits) {
        if (seg_steps < 2 || !is_pow2(seg_steps)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        size_t blowup = 16;
        if (seg_steps > (size_t(1) << Field::TWO_ADICITY) / blowup) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        size_t lde_len = seg_steps * blowup;
        uint32_t mb = fri::composition_millibits(lde_len, seg_steps - 1);
        auto [nq, gb] = fri::tune_strength(target_bits, mb);
        auto [fs, ch] = fri::tune_shape(lde_len, nq, {{lde_len, 24}});
        Params p;
        p.seg_steps = seg_steps;
        p.blowup = blowup;
        p.num_queries = nq;
        p.final_size = fs;
        p.grinding_bits = gb;
        p.cap_height = ch;
        p.in_a = in_a;
        p.in_b = in_b;
        p.out_b = out_b;
        return p;
    }

    size_t lde_len() const { return seg_steps * blowup; }

    // fold.rs:177-183
    fri::Params fri_params() const {
        fri::Params p = fri::Params::make(blowup, num_queries, final_size);
        p.grinding_bits = grinding_bits;
        p.cap_height = cap_height;
        return p;
    }

    // fold.rs:186-218
    void validate() const {
        if (seg_steps < 2 || !is_pow2(seg_steps)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        if (blowup < 4 || !is_pow2(blowup)) throw Z0pError(ERR_INVALID_BLOWUP);
        if (num_queries == 0) throw Z0pError(ERR_NO_QUERIES);
        if (final_size < 2 || !is_pow2(final_size)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        if (cap_height > 8) throw Z0pError(ERR_MALFORMED_INPUT, "cap_height above 8");
        unsigned __int128 lde = (unsigned __int128)seg_steps * blowup;
        if (lde > ((unsigned __int128)1 << Field::TWO_ADICITY)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
    }
};

// fold.rs:224-265
struct StepClaims {
    Field in_a, in_b, out_a, out_b;
};
struct StepOpening {
    size_t index = 0;
    std::array<Field, 4> cols;  // A, B at index; A, B at the shifted index
    Path path, shift_path;
};
struct StepProof {
    StepClaims claims;
    std::vector<Digest> column_cap;
    std::vector<StepOpening> openings;
};

// fold.rs:267-359
struct IvcProof {
    std::vector<StepProof> steps;
    fri::Commitment fri_commitment;
    fri::Proof fri_proof;

    std::vector<uint8_t> to_bytes() const {
        Writer w;
        w.byte(VERSION);
        w.count(steps.size());
        for (auto& s : steps) {
            w.field(s.claims.in_a);
            w.field(s.claims.in_b);
            w.field(s.claims.out_a);
            w.field(s.claims.out_b);
            w.cap(s.column_cap);
            w.count(s.openings.size());
            for (auto& o : s.openings) {
                w.u64_le(o.index);
                for (auto& v : o.cols) w.field(v);
                w.path(o.path);
                w.path(o.shift_path);
            }
        }
        fri_commitment.write_body(w);
        fri_proof.write_body(w);
        return w.out;
    }

    static IvcProof from_bytes(const uint8_t* p, size_t n) {
        Reader r(p, n);
        r.version();
        IvcProof pr;
        size_t ns = r.count(MAX_STEPS);
        for (size_t i = 0; i < ns; ++i) {
            StepProof s;
            s.claims.in_a = r.field();
            s.claims.in_b = r.field();
            s.claims.out_a = r.field();
            s.claims.out_b = r.field();
            s.column_cap = r.cap();
            size_t no = r.count(MAX_QUERIES);
            for (size_t k = 0; k < no; ++k) {
                StepOpening o;
                o.index = r.dimension();
                for (auto& v : o.cols) v = r.field();
                o.path = r.path();
                o.shift_path = r.path();
                s.openings.push_back(std::move(o));
            }
            pr.steps.push_back(std::move(s));
        }
        pr.fri_commitment = fri::Commitment::read_body(r);
        pr.fri_proof = fri::Proof::read_body(r);
        r.end();
        return pr;
    }
};

// fold.rs:729-741 — 60 bytes, grinding_bits as u32.
inline void absorb_params(const Params& p, Transcript& t) {
    std::vector<uint8_t> buf(56);
    uint8_t* q = buf.data();
    put_u64(q, p.seg_steps);
    put_u64(q + 8, p.blowup);
    put_u64(q + 16, p.num_queries);
    put_u64(q + 24, p.final_size);
    put_u32(q + 32, p.grinding_bits);
    put_u64(q + 36, p.cap_height);
    put_u32(q + 44, p.in_a.v);
    put_u32(q + 48, p.in_b.v);
    put_u32(q + 52, p.out_b.v);
    t.absorb("fold/params", buf);
}

// fold.rs:743-751
inline void absorb_step(const StepClaims& c, Transcript& t) {
    std::vector<uint8_t> buf(16);
    put_u32(buf.data(), c.in_a.v);
    put_u32(buf.data() + 4, c.in_b.v);
    put_u32(buf.data() + 8, c.out_a.v);
    put_u32(buf.data() + 12, c.out_b.v);
    t.absorb("fold/step", buf);
}

inline Field lde_point_at(size_t m, size_t index) { return fri::lde_shift() * primitive_root(m).pow(index); }

// fold.rs:856-879
inline Ext4 composition_at(const Field cols[2], const Field next[2], Field point, Field w_last, const StepClaims& c,
                           size_t s, const std::array<Ext4, ALPHAS>& al) {
    Field vanishing = point.pow(s) - Field::one();
    Field z_trans = vanishing * (point - w_last).inv();
    Field trans1 = (next[0] - cols[1]) * z_trans.inv();
    Field trans2 = (next[1] - (cols[0] + cols[1])) * z_trans.inv();
    Field in_a = (cols[0] - c.in_a) * (point - Field::one()).inv();
    Field in_b = (cols[1] - c.in_b) * (point - Field::one()).inv();
    Field out_a = (cols[0] - c.out_a) * (point - w_last).inv();
    Field out_b = (cols[1] - c.out_b) * (point - w_last).inv();
    return al[0].mul_base(trans1) + al[1].mul_base(trans2) + al[2].mul_base(in_a) + al[3].mul_base(in_b) +
           al[4].mul_base(out_a) + al[5].mul_base(out_b);
}

// fold.rs:785-854 — batched per 2048-point chunk, as composition_at.
inline std::vector<Ext4> compose_step(const std::vector<Field>* cols, Field w_last, const StepClaims& c, size_t s,
                                      size_t offset, const std::array<Ext4, ALPHAS>& al) {
    size_t m = cols[0].size();
    std::vector<Ext4> out(m);
    Field omega = primitive_root(m);
    Field p0 = fri::lde_shift().pow(s);
    Field ratio = omega.pow(s);
    for_each_indexed(out.data(), m, 2048, [&](size_t base, Ext4* piece, size_t len) {
        std::vector<Field> den_a(len), den_b(len), z(len);
        Field x = fri::lde_shift() * omega.pow(base);
        Field cur = p0 * ratio.pow(base);
        for (size_t j = 0; j < len; ++j) {
            den_a[j] = x - w_last;
            den_b[j] = x - Field::one();
            z[j] = cur - Field::one();
            x *= omega;
            cur *= ratio;
        }
        batch_invert_in_place(den_a.data(), len);
        batch_invert_in_place(den_b.data(), len);
        for (size_t j = 0; j < len; ++j) z[j] *= den_a[j];
        batch_invert_in_place(z.data(), len);
        for (size_t j = 0; j < len; ++j) {
            size_t i = base + j, nx = (i + offset) % m;
            Field a = cols[0][i], b = cols[1][i];
            Field t1 = (cols[0][nx] - b) * z[j];
            Field t2 = (cols[1][nx] - (a + b)) * z[j];
            piece[j] = al[0].mul_base(t1) + al[1].mul_base(t2) + al[2].mul_base((a - c.in_a) * den_b[j]) +
                       al[3].mul_base((b - c.in_b) * den_b[j]) + al[4].mul_base((a - c.out_a) * den_a[j]) +
                       al[5].mul_base((b - c.out_b) * den_a[j]);
        }
    });
    return out;
}

// fold.rs:361-570
struct Accumulator {
    struct StepData {
        StepClaims claims;
        std::vector<Field> ldes[2];
        MerkleTree tree;
        std::vector<Ext4> composition;
    };
    Params params;
    std::vector<StepData> steps;

    explicit Accumulator(const Params& p) : params(p) {}
    size_t len() const { return steps.size(); }

    // fold.rs:407-492
    void add_segment(const std::vector<Field>& col_a, const std::vector<Field>& col_b, Transcript& transcript) {
        params.validate();
        size_t s = params.seg_steps;
        if (col_a.size() != s || col_b.size() != s)
            throw Z0pError(ERR_MALFORMED_INPUT, "segment columns must match seg_steps");
        StepClaims claims{col_a[0], col_b[0], col_a[s - 1], col_b[s - 1]};
        if (steps.empty()) absorb_params(params, transcript);
        absorb_step(claims, transcript);
        Field omega_base = primitive_root(s);
        StepData d;
        d.claims = claims;
        d.ldes[0] = evaluate_on_coset(interpolate(col_a), params.blowup, fri::lde_shift());
        d.ldes[1] = evaluate_on_coset(interpolate(col_b), params.blowup, fri::lde_shift());
        size_t m = d.ldes[0].size();
        size_t offset = m / s;
        std::vector<uint8_t> flat(m * 8);
        for (size_t i = 0; i < m; ++i) {
            put_u32(flat.data() + 8 * i, d.ldes[0][i].v);
            put_u32(flat.data() + 8 * i + 4, d.ldes[1][i].v);
        }
        d.tree = MerkleTree::from_flat(flat.data(), flat.size(), 8);
        size_t drop = capped_shape(m, params.cap_height).first;
        transcript.absorb("fold/trace", cap_bytes(d.tree.cap(drop)));
        std::array<Ext4, ALPHAS> al;
        for (auto& a : al) a = transcript.challenge_ext("fold/alpha");
        Field w_last = omega_base.pow(s - 1);
        d.composition = compose_step(d.ldes, w_last, claims, s, offset, al);
        steps.push_back(std::move(d));
    }

    // fold.rs:494-570
    IvcProof finalize(Transcript& transcript) const {
        if (steps.empty()) throw Z0pError(ERR_EMPTY_INPUT);
        size_t m = params.lde_len();
        std::vector<Ext4> gammas;
        for (size_t k = 0; k < steps.size(); ++k) gammas.push_back(transcript.challenge_ext("fold/gamma"));
        std::vector<Ext4> combined(m);
        for_each_indexed(combined.data(), m, 2048, [&](size_t base, Ext4* piece, size_t len) {
            for (size_t j = 0; j < len; ++j) {
                Ext4 acc;
                for (size_t k = 0; k < steps.size(); ++k) acc = acc + gammas[k] * steps[k].composition[base + j];
                piece[j] = acc;
            }
        });
        fri::Params fp = params.fri_params();
        fri::ProverState st = fri::commit_ext_evals_owned(std::move(combined), params.seg_steps - 1, fp, transcript);
        fri::Commitment fc = fri::public_commitment(st);
        auto opened = fri::open(st, transcript);
        uint32_t log_m = log2_exact(m);
        size_t offset = m / params.seg_steps;
        size_t drop = capped_shape(m, params.cap_height).first;
        IvcProof proof;
        for (auto& d : steps) {
            StepProof sp;
            sp.claims = d.claims;
            sp.column_cap = d.tree.cap(drop);
            for (size_t leaf : opened.second) {
                size_t index = reverse_bits_n(leaf, log_m);
                size_t shifted = (index + offset) % m;
                StepOpening o;
                o.index = index;
                o.cols = {d.ldes[0][index], d.ldes[1][index], d.ldes[0][shifted], d.ldes[1][shifted]};
                o.path = d.tree.prove_capped(index, drop);
                o.shift_path = d.tree.prove_capped(shifted, drop);
                sp.openings.push_back(std::move(o));
            }
            proof.steps.push_back(std::move(sp));
        }
        proof.fri_commitment = std::move(fc);
        proof.fri_proof = std::move(opened.first);
        return proof;
    }
};

// fold.rs:710-727
inline void check_chain(const Params& params, const IvcProof& proof) {
    const auto& first = proof.steps[0];
    if (first.claims.in_a != params.in_a || first.claims.in_b != params.in_b)
        throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    for (size_t k = 0; k + 1 < proof.steps.size(); ++k)
        if (proof.steps[k].claims.out_a != proof.steps[k + 1].claims.in_a ||
            proof.steps[k].claims.out_b != proof.steps[k + 1].claims.in_b)
            throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    if (proof.steps.back().claims.out_b != params.out_b) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
}

// fold.rs:669-708
inline void check_step_openings(const StepProof& step, const StepOpening& o, size_t index, size_t m, size_t offset,
                                size_t cap_height) {
    auto [drop, depth] = capped_shape(m, cap_height);
    if (step.column_cap.size() != (size_t(1) << drop) || o.path.size() != depth - drop ||
        o.shift_path.size() != depth - drop)
        throw Z0pError(ERR_BAD_MERKLE_PATH);
    uint8_t row[8];
    put_u32(row, o.cols[0].v);
    put_u32(row + 4, o.cols[1].v);
    size_t e = index >> (depth - drop);
    if (e >= step.column_cap.size() || !verify_capped(step.column_cap[e], row, 8, index, depth, o.path))
        throw Z0pError(ERR_BAD_MERKLE_PATH);
    size_t shifted = (index + offset) % m;
    put_u32(row, o.cols[2].v);
    put_u32(row + 4, o.cols[3].v);
    size_t se = shifted >> (depth - drop);
    if (se >= step.column_cap.size() || !verify_capped(step.column_cap[se], row, 8, shifted, depth, o.shift_path))
        throw Z0pError(ERR_BAD_MERKLE_PATH);
}

// fold.rs:572-667
inline void verify(const Params& params, const IvcProof& proof, Transcript& transcript) {
    params.validate();
    if (proof.steps.empty()) throw Z0pError(ERR_EMPTY_INPUT);
    size_t s = params.seg_steps, m = params.lde_len();
    if (m % s != 0) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    if (proof.fri_commitment.lde_len != m || proof.fri_commitment.poly_len != m ||
        proof.fri_commitment.degree_bound != s - 1)
        throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    size_t offset = m / s;
    absorb_params(params, transcript);
    check_chain(params, proof);
    Field w_last = primitive_root(s).pow(s - 1);
    std::vector<std::array<Ext4, ALPHAS>> all_alphas;
    for (auto& step : proof.steps) {
        absorb_step(step.claims, transcript);
        transcript.absorb("fold/trace", cap_bytes(step.column_cap));
        std::array<Ext4, ALPHAS> al;
        for (auto& a : al) a = transcript.challenge_ext("fold/alpha");
        all_alphas.push_back(al);
        if (step.openings.size() != params.num_queries) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    }
    std::vector<Ext4> gammas;
    for (size_t k = 0; k < proof.steps.size(); ++k) gammas.push_back(transcript.challenge_ext("fold/gamma"));
    fri::Params fp = params.fri_params();
    std::vector<size_t> idx = fri::verify(proof.fri_commitment, proof.fri_proof, fp, transcript);
    uint32_t log_m = log2_exact(m);
    size_t count = std::min(proof.fri_proof.queries.size(), idx.size());
    for (size_t q = 0; q < count; ++q) {
        size_t leaf = idx[q];
        size_t index = reverse_bits_n(leaf, log_m);
        Field point = lde_point_at(m, index);
        Ext4 combined;
        for (size_t k = 0; k < proof.steps.size(); ++k) {
            const StepOpening& o = proof.steps[k].openings[q];
            if (o.index != index) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
            check_step_openings(proof.steps[k], o, index, m, offset, params.cap_height);
            Field cols[2] = {o.cols[0], o.cols[1]};
            Field next[2] = {o.cols[2], o.cols[3]};
            combined = combined + gammas[k] * composition_at(cols, next, point, w_last, proof.steps[k].claims, s,
                                                              all_alphas[k]);
        }
        const auto& query = proof.fri_proof.queries[q];
        if (query.folds.empty()) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
        if (query.folds[0].values[leaf % 4] != combined) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    }
}

}  // namespace fold
}  // namespace oracle
