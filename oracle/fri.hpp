// ORACLE — test infrastructure only (see bb.hpp).
//
// FRI polynomial commitment (arity-4 Fp4 folds, quad-leaf Merkle layers,
// caps, proof-of-work), restated from src/fri.rs.
#pragma once
#include <vector>

#include "codec.hpp"
#include "ntt.hpp"
#include "timers.hpp"
#include "transcript.hpp"

namespace oracle {
namespace fri {

// fri.rs:94
inline Field lde_shift() { return Field::generator(); }
// fri.rs:101
constexpr size_t MAX_FINAL_SIZE = size_t(1) << 10;

// fri.rs:105-133
struct Params {
    size_t blowup = 0;
    size_t num_queries = 0;
    size_t final_size = 0;
    uint32_t grinding_bits = 0;
    size_t cap_height = 0;

    // fri.rs:141-158
    static Params make(size_t blowup, size_t num_queries, size_t final_size) {
        if (blowup < 2 || !is_pow2(blowup)) throw Z0pError(ERR_INVALID_BLOWUP);
        if (num_queries == 0) throw Z0pError(ERR_NO_QUERIES);
        if (final_size < 2 || !is_pow2(final_size) || final_size > MAX_FINAL_SIZE) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        Params p;
        p.blowup = blowup;
        p.num_queries = num_queries;
        p.final_size = final_size;
        return p;
    }

    // fri.rs:163-196
    void validate() const {
        if (blowup < 2 || !is_pow2(blowup)) throw Z0pError(ERR_INVALID_BLOWUP);
        if (num_queries == 0) throw Z0pError(ERR_NO_QUERIES);
        if (num_queries > MAX_QUERIES) throw Z0pError(ERR_MALFORMED_INPUT, "num_queries above protocol cap");
        if (final_size < 2 || !is_pow2(final_size) || final_size > MAX_FINAL_SIZE) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
        if (grinding_bits > 32) throw Z0pError(ERR_MALFORMED_INPUT, "grinding_bits above 32");
        if (cap_height > 8) throw Z0pError(ERR_MALFORMED_INPUT, "cap_height above 8");
    }

    static Params for_security(size_t poly_len, uint32_t target_bits);
};

// fri.rs:963-975
inline std::vector<size_t> fri_layer_sizes(size_t len, size_t final_size) {
    std::vector<size_t> out;
    for (;;) {
        out.push_back(len);
        if (len <= final_size) break;
        len /= 4;
    }
    return out;
}

// fri.rs:983-999
inline size_t estimate_fri_bytes(const std::vector<size_t>& layers, size_t num_queries, size_t cap_height) {
    size_t total = 24 + 8;
    for (size_t i = 0; i < layers.size(); ++i) {
        size_t values = layers[i];
        size_t tree_leaves = values % 4 == 0 ? values / 4 : values;
        // usize::trailing_zeros(0) = 64: the final-size enumeration of
        // tune_shape can reach an empty layer on odd log2 sizes.
        size_t depth = tree_leaves == 0 ? 64 : size_t(__builtin_ctzll(tree_leaves));
        size_t drop = std::min(cap_height, depth);
        total += 32 * (size_t(1) << std::min<size_t>(drop, 16));
        if (i + 1 < layers.size()) {
            size_t steps = depth > drop ? depth - drop : 0;
            total += num_queries * (64 + 33 * steps);
        }
    }
    return total;
}

// fri.rs:1007-1021
inline size_t estimate_col_bytes(size_t lde_len, size_t num_queries, size_t value_bytes, size_t cap_height) {
    size_t depth = size_t(__builtin_ctzll(lde_len));
    size_t drop = std::min(cap_height, depth);
    return 32 * (size_t(1) << std::min<size_t>(drop, 16)) + num_queries * (value_bytes + 33 * (depth - drop));
}

// fri.rs:1009-1044 (tune_shape) — ties keep the larger final, smaller cap.
inline std::pair<size_t, size_t> tune_shape(size_t lde_len, size_t num_queries,
                                            const std::vector<std::pair<size_t, size_t>>& columns) {
    size_t shifted = std::max<size_t>(lde_len / 4, 2);
    while (shifted > 32) shifted /= 4;
    std::pair<size_t, size_t> best{shifted, 0};
    size_t best_bytes = SIZE_MAX;
    for (;;) {
        std::vector<size_t> layers = fri_layer_sizes(lde_len, shifted);
        for (size_t h = 0; h <= 4; ++h) {
            size_t total = estimate_fri_bytes(layers, num_queries, h);
            for (auto& c : columns) total += estimate_col_bytes(c.first, num_queries, c.second, h);
            if (total < best_bytes) {
                best_bytes = total;
                best = {shifted, h};
            }
        }
        if (shifted <= 1) break;
        shifted /= 4;
        if (shifted < 1) break;
    }
    return best;
}

// fri.rs:1051-1057
inline uint32_t composition_millibits(size_t lde_len, size_t degree_bound) {
    size_t den = 1;
    while (den < degree_bound + 1) den <<= 1;
    if (den == 0 || lde_len < den) return 500;
    return 500 * (uint32_t(__builtin_ctzll(lde_len)) - uint32_t(__builtin_ctzll(den)));
}

// fri.rs:1066-1074
inline std::pair<size_t, uint32_t> tune_strength(uint32_t target_bits, uint32_t millibits_per_query) {
    uint32_t grinding = std::min<uint32_t>(20, target_bits / 2);
    uint64_t remaining = uint64_t(target_bits > grinding ? target_bits - grinding : 0) * 1000;
    uint64_t per_query = std::max<uint32_t>(millibits_per_query, 1);
    uint64_t q = (remaining + per_query - 1) / per_query;
    q = std::min<uint64_t>(std::max<uint64_t>(q, 1), 1000000);
    return {size_t(q), grinding};
}

// fri.rs:232-255
inline Params Params::for_security(size_t poly_len, uint32_t target_bits) {
    if (poly_len == 0) throw Z0pError(ERR_EMPTY_INPUT);
    if (!is_pow2(poly_len)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
    size_t blowup = 16;
    size_t lde_len = poly_len * blowup;
    if (lde_len > (size_t(1) << Field::TWO_ADICITY)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
    auto [nq, gb] = tune_strength(target_bits, 2000);
    auto [fs, ch] = tune_shape(lde_len, nq, {});
    Params p;
    p.blowup = blowup;
    p.num_queries = nq;
    p.final_size = fs;
    p.grinding_bits = gb;
    p.cap_height = ch;
    return p;
}

// fri.rs:258-275
struct Commitment {
    std::vector<std::vector<Digest>> caps;
    size_t lde_len = 0, poly_len = 0, degree_bound = 0;

    // fri.rs:896-904
    void write_body(Writer& w) const {
        w.u64_le(lde_len);
        w.u64_le(poly_len);
        w.u64_le(degree_bound);
        w.count(caps.size());
        for (auto& c : caps) w.cap(c);
    }
    // fri.rs:906-921
    static Commitment read_body(Reader& r) {
        Commitment c;
        c.lde_len = r.dimension();
        c.poly_len = r.dimension();
        c.degree_bound = r.dimension();
        size_t n = r.count(MAX_FOLDS);
        for (size_t i = 0; i < n; ++i) c.caps.push_back(r.cap());
        return c;
    }
    std::vector<uint8_t> to_bytes() const {
        Writer w;
        w.byte(VERSION);
        write_body(w);
        return w.out;
    }
    static Commitment from_bytes(const uint8_t* p, size_t n) {
        Reader r(p, n);
        r.version();
        Commitment c = read_body(r);
        r.end();
        return c;
    }
};

// fri.rs:300-306
struct FoldWitness {
    Ext4 values[4];
    Path path;
};
struct QueryProof {
    std::vector<FoldWitness> folds;
};

// fri.rs:316-398
struct Proof {
    std::vector<Ext4> final_poly;
    std::vector<QueryProof> queries;
    uint64_t pow_nonce = 0;

    void write_body(Writer& w) const {
        w.count(final_poly.size());
        for (auto& v : final_poly) w.ext4(v);
        w.u64_le(pow_nonce);
        w.count(queries.size());
        for (auto& q : queries) {
            w.count(q.folds.size());
            for (auto& f : q.folds) {
                for (auto& v : f.values) w.ext4(v);
                w.path(f.path);
            }
        }
    }
    static Proof read_body(Reader& r) {
        Proof p;
        size_t nf = r.count(MAX_ELEMS);
        for (size_t i = 0; i < nf; ++i) p.final_poly.push_back(r.ext4());
        p.pow_nonce = r.u64_le();
        size_t nq = r.count(MAX_QUERIES);
        for (size_t i = 0; i < nq; ++i) {
            QueryProof q;
            size_t nfo = r.count(MAX_FOLDS);
            for (size_t j = 0; j < nfo; ++j) {
                FoldWitness fw;
                for (int k = 0; k < 4; ++k) fw.values[k] = r.ext4();
                fw.path = r.path();
                q.folds.push_back(std::move(fw));
            }
            p.queries.push_back(std::move(q));
        }
        return p;
    }
    std::vector<uint8_t> to_bytes() const {
        Writer w;
        w.byte(VERSION);
        write_body(w);
        return w.out;
    }
    static Proof from_bytes(const uint8_t* b, size_t n) {
        Reader r(b, n);
        r.version();
        Proof p = read_body(r);
        r.end();
        return p;
    }
};

// fri.rs:277-294
struct ProverState {
    std::vector<std::vector<Ext4>> layers;
    std::vector<MerkleTree> trees;
    std::vector<std::vector<Digest>> caps;
    std::vector<std::array<Ext4, 3>> fold_challenges;
    size_t lde_len = 0, poly_len = 0, degree_bound = 0;
    Params params;
};

// fri.rs:940-943
inline size_t fri_drop_top(size_t layer_leaves, size_t cap_height) {
    return std::min(cap_height, size_t(__builtin_ctzll(layer_leaves)));
}
// fri.rs:928-930
inline bool fold_shift_fits(size_t folds) { return folds <= 30 && folds < 32; }

// fri.rs:1086-1094
inline std::vector<Field> coset_points_with_step(size_t len, Field step) {
    std::vector<Field> pts;
    pts.reserve(len);
    Field cur = lde_shift();
    for (size_t i = 0; i < len; ++i) {
        pts.push_back(cur);
        cur *= step;
    }
    return pts;
}

// fri.rs:1108-1115
inline Field point_at(size_t lde_len, size_t depth, size_t j) {
    size_t log_m = size_t(__builtin_ctzll(lde_len));
    size_t bits = log_m - 2 * depth;
    Field omega = primitive_root(lde_len);
    size_t e = reverse_bits_n(j, uint32_t(bits));
    Field base = lde_shift() * omega.pow(e);
    return base.pow(uint64_t(1) << (2 * depth));
}

// fri.rs:1125-1140
inline void check_layers(size_t lde_len, size_t final_size) {
    if (final_size >= lde_len || lde_len % final_size != 0) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
    size_t ratio = lde_len / final_size;
    if (!is_pow2(ratio) || __builtin_ctzll(ratio) % 2 != 0) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
}

// fri.rs:1311-1331 — A + r0 B + r1 C + r2 D by Horner in x^{-1}.
inline Ext4 fold_values(const Ext4 quad[4], Field inv_point, const std::array<Ext4, 3>& ch, Field unit) {
    Ext4 even_sum = (quad[0] + quad[1]) + (quad[2] + quad[3]);
    Ext4 odd_diff = quad[0] - quad[1];
    Ext4 skew_diff = quad[2] - quad[3];
    Ext4 mean = even_sum.halve().halve();
    Ext4 odd = (odd_diff - skew_diff.mul_base(unit)).halve().halve();
    Ext4 even = ((quad[0] + quad[1]) - (quad[2] + quad[3])).halve().halve();
    Ext4 skew = (odd_diff + skew_diff.mul_base(unit)).halve().halve();
    Ext4 acc = ((ch[2] * skew).mul_base(inv_point) + ch[1] * even).mul_base(inv_point) + ch[0] * odd;
    return mean + acc.mul_base(inv_point);
}

// fri.rs:1273-1302
inline std::pair<std::vector<Ext4>, std::vector<Field>> fold_layer(const std::vector<Ext4>& values,
                                                                   const std::vector<Field>& first_points,
                                                                   const std::array<Ext4, 3>& ch, Field unit) {
    size_t quarter = values.size() / 4;
    std::vector<Field> inv_points = batch_invert(first_points);
    std::vector<Ext4> next(quarter);
    for_each_indexed(next.data(), quarter, 512, [&](size_t base, Ext4* piece, size_t n) {
        for (size_t j = 0; j < n; ++j) {
            size_t i = base + j;
            Ext4 quad[4] = {values[4 * i], values[4 * i + 1], values[4 * i + 2], values[4 * i + 3]};
            piece[j] = fold_values(quad, inv_points[i], ch, unit);
        }
    });
    std::vector<Field> next_first;
    next_first.reserve(quarter / 4);
    for (size_t i = 0; i + 3 < first_points.size(); i += 4) next_first.push_back(first_points[i].square().square());
    return {std::move(next), std::move(next_first)};
}

// fri.rs:1339-1364
inline bool check_degree_ext(const std::vector<Field>& points, const std::vector<Ext4>& values, size_t bound) {
    if (points.size() != values.size() || values.empty()) return false;
    size_t n = values.size();
    if (bound >= n - 1) return true;
    std::vector<Ext4> table = values;
    size_t last_order = std::min(bound + 1, n - 1);
    for (size_t order = 1; order <= last_order; ++order) {
        for (size_t i = 0; i < n - order; ++i) {
            Field denom = points[i + order] - points[i];
            if (denom == Field::zero()) return false;
            table[i] = (table[i + 1] - table[i]) * Ext4::from_base(denom.inv());
        }
        if (order > bound)
            for (size_t i = 0; i < n - order; ++i)
                if (table[i] != Ext4::zero()) return false;
    }
    return true;
}

// fri.rs:1401-1407, 1413-1420, 1422-1428
inline MerkleTree commit_layer(const std::vector<Ext4>& layer) {
    std::vector<uint8_t> flat(16 * layer.size());
    for (size_t i = 0; i < layer.size(); ++i) layer[i].to_le_bytes(flat.data() + 16 * i);
    if (layer.size() % 4 == 0) return MerkleTree::from_flat64(flat.data(), flat.size());
    return MerkleTree::from_flat(flat.data(), flat.size(), 16);
}

// fri.rs:1431-1437
inline void quad_leaf(const Ext4 v[4], uint8_t out[64]) {
    for (int i = 0; i < 4; ++i) v[i].to_le_bytes(out + 16 * i);
}
// fri.rs:1439-1445
inline std::vector<uint8_t> encode_fields(const std::vector<Ext4>& f) {
    std::vector<uint8_t> out(16 * f.size());
    for (size_t i = 0; i < f.size(); ++i) f[i].to_le_bytes(out.data() + 16 * i);
    return out;
}
// fri.rs:1450-1456
inline std::vector<uint8_t> encode_dims(size_t lde_len, size_t poly_len, size_t degree_bound) {
    std::vector<uint8_t> out(24);
    put_u64(out.data(), lde_len);
    put_u64(out.data() + 8, poly_len);
    put_u64(out.data() + 16, degree_bound);
    return out;
}

// fri.rs:551-633
inline ProverState commit_inner(std::vector<Ext4> current, size_t degree_bound, size_t poly_len, const Params& params,
                                Transcript& transcript) {
    params.validate();
    size_t lde_len = current.size();
    check_layers(lde_len, params.final_size);
    if (lde_len < 2 || degree_bound + 1 > lde_len / 2) throw Z0pError(ERR_MALFORMED_INPUT, "degree bound implies rate >= 1/2");
    size_t folds = size_t(__builtin_ctzll(lde_len / params.final_size)) / 2;
    size_t residual = degree_bound >> (2 * folds);
    if (residual + 1 >= params.final_size)
        throw Z0pError(ERR_MALFORMED_INPUT, "final layer too small for a non-vacuous degree check");
    transcript.absorb("fri-dims", encode_dims(lde_len, poly_len, degree_bound));

    std::vector<Field> first_points;
    {
        ScopedStage s("fri_layer0_prep");
        first_points = coset_points_with_step(lde_len / 4, primitive_root(lde_len));
        bit_reverse_permute_ext(current);
        bit_reverse_permute(first_points);
    }
    Field unit = two_adic_generator(2);

    ProverState st;
    st.layers.push_back(std::move(current));
    for (;;) {
        size_t idx = st.layers.size() - 1;
        MerkleTree tree;
        {
            ScopedStage s("fri_commit");
            tree = commit_layer(st.layers[idx]);
        }
        std::vector<Digest> cap = tree.cap(fri_drop_top(tree.num_leaves(), params.cap_height));
        for (auto& d : cap) transcript.absorb("fri-layer-root", d.data(), 32);
        st.trees.push_back(std::move(tree));
        st.caps.push_back(cap);
        if (st.layers[idx].size() == params.final_size) break;
        std::array<Ext4, 3> ch;
        for (auto& c : ch) c = transcript.challenge_ext("fri-fold");
        st.fold_challenges.push_back(ch);
        ScopedStage s("fri_fold");
        auto folded = fold_layer(st.layers[idx], first_points, ch, unit);
        st.layers.push_back(std::move(folded.first));
        first_points = std::move(folded.second);
    }
    transcript.absorb("fri-final", encode_fields(st.layers.back()));
    st.lde_len = lde_len;
    st.poly_len = poly_len;
    st.degree_bound = degree_bound;
    st.params = params;
    return st;
}

// fri.rs:408-433
inline ProverState commit(const std::vector<Field>& coeffs, const Params& params, Transcript& transcript) {
    params.validate();
    if (coeffs.empty()) throw Z0pError(ERR_EMPTY_INPUT);
    if (!is_pow2(coeffs.size())) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
    std::vector<Field> lde;
    {
        ScopedStage s("lde");
        lde = evaluate_on_coset(coeffs, params.blowup, lde_shift());
    }
    std::vector<Ext4> ext(lde.size());
    for (size_t i = 0; i < lde.size(); ++i) ext[i] = Ext4::from_base(lde[i]);
    return commit_inner(std::move(ext), coeffs.size() - 1, coeffs.size(), params, transcript);
}

// fri.rs:444-461
inline ProverState commit_ext(const std::vector<Ext4>& coeffs, const Params& params, Transcript& transcript) {
    params.validate();
    if (coeffs.empty()) throw Z0pError(ERR_EMPTY_INPUT);
    if (!is_pow2(coeffs.size())) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
    std::vector<Ext4> lde = evaluate_on_coset_ext(coeffs, params.blowup, lde_shift());
    return commit_inner(std::move(lde), coeffs.size() - 1, coeffs.size(), params, transcript);
}

// fri.rs:473-500
inline ProverState commit_evals(const std::vector<Field>& evals, size_t degree_bound, const Params& params,
                                Transcript& transcript) {
    params.validate();
    if (evals.empty()) throw Z0pError(ERR_EMPTY_INPUT);
    if (!is_pow2(evals.size()) || evals.size() > (size_t(1) << Field::TWO_ADICITY)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
    if (degree_bound >= evals.size()) throw Z0pError(ERR_MALFORMED_INPUT, "degree bound must be below LDE size (rate < 1)");
    std::vector<Ext4> ext(evals.size());
    for (size_t i = 0; i < evals.size(); ++i) ext[i] = Ext4::from_base(evals[i]);
    return commit_inner(std::move(ext), degree_bound, evals.size(), params, transcript);
}

// fri.rs:527-549
inline ProverState commit_ext_evals_owned(std::vector<Ext4> evals, size_t degree_bound, const Params& params,
                                          Transcript& transcript) {
    params.validate();
    if (evals.empty()) throw Z0pError(ERR_EMPTY_INPUT);
    if (!is_pow2(evals.size()) || evals.size() > (size_t(1) << Field::TWO_ADICITY)) throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
    size_t lde_len = evals.size();
    if (degree_bound >= lde_len) throw Z0pError(ERR_MALFORMED_INPUT, "degree bound must be below LDE size (rate < 1)");
    return commit_inner(std::move(evals), degree_bound, lde_len, params, transcript);
}

// fri.rs:863-870
inline Commitment public_commitment(const ProverState& st) {
    Commitment c;
    c.caps = st.caps;
    c.lde_len = st.lde_len;
    c.poly_len = st.poly_len;
    c.degree_bound = st.degree_bound;
    return c;
}

// fri.rs:1229-1260
inline QueryProof open_one(const ProverState& st, size_t leaf_index) {
    if (leaf_index >= st.lde_len) throw Z0pError(ERR_INDEX_OUT_OF_BOUNDS);
    QueryProof q;
    size_t index = leaf_index;
    for (size_t depth = 0; depth + 1 < st.layers.size(); ++depth) {
        const auto& layer = st.layers[depth];
        size_t base = index / 4 * 4;
        FoldWitness fw;
        for (int k = 0; k < 4; ++k) fw.values[k] = layer[base + k];
        Path full = st.trees[depth].prove(base / 4);
        size_t drop = fri_drop_top(st.trees[depth].num_leaves(), st.params.cap_height);
        size_t stored = full.size() > drop ? full.size() - drop : 0;
        full.resize(stored);
        fw.path = std::move(full);
        q.folds.push_back(std::move(fw));
        index /= 4;
    }
    return q;
}

// fri.rs:645-666
inline std::pair<Proof, std::vector<size_t>> open(const ProverState& st, Transcript& transcript) {
    uint64_t nonce = 0;
    if (st.params.grinding_bits > 0) {
        ScopedStage s("grind");
        nonce = transcript.grind("fri-pow", st.params.grinding_bits);
    }
    ScopedStage s("fri_open");
    Proof proof;
    std::vector<size_t> indices;
    for (size_t i = 0; i < st.params.num_queries; ++i) {
        auto li = transcript.challenge_index("fri-query", st.lde_len);
        if (!li) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
        proof.queries.push_back(open_one(st, *li));
        indices.push_back(*li);
    }
    proof.final_poly = st.layers.back();
    proof.pow_nonce = nonce;
    return {std::move(proof), std::move(indices)};
}

// fri.rs:1153-1227
inline void verify_one(const Commitment& cm, const Proof& proof, const std::vector<std::array<Ext4, 3>>& fc,
                       const QueryProof& query, size_t leaf_index, Field unit, size_t cap_height) {
    size_t index = leaf_index;
    for (size_t depth = 0; depth < query.folds.size(); ++depth) {
        const FoldWitness& w = query.folds[depth];
        const auto& cap = cm.caps[depth];
        size_t layer_len = cm.lde_len >> (2 * depth);
        size_t base = index / 4 * 4;
        if (base + 3 >= layer_len) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
        size_t quad_len = layer_len / 4;
        size_t depth_bits = size_t(__builtin_ctzll(quad_len));
        size_t drop = std::min(cap_height, depth_bits);
        if (w.path.size() != depth_bits - drop) throw Z0pError(ERR_BAD_MERKLE_PATH);
        size_t quad_index = base / 4;
        size_t e = quad_index >> (depth_bits - drop);
        if (e >= cap.size()) throw Z0pError(ERR_BAD_MERKLE_PATH);
        Field x = point_at(cm.lde_len, depth, base);
        uint8_t leaf[64];
        quad_leaf(w.values, leaf);
        Digest node = leaf64_digest(leaf);
        size_t position = quad_index;
        for (auto& step : w.path) {
            if (step.second != ((position & 1) == 0)) throw Z0pError(ERR_BAD_MERKLE_PATH);
            node = step.second ? hash_node(node, step.first) : hash_node(step.first, node);
            position >>= 1;
        }
        if (node != cap[e]) throw Z0pError(ERR_BAD_MERKLE_PATH);
        Ext4 parent = fold_values(w.values, x.inv(), fc[depth], unit);
        if (depth + 1 == query.folds.size()) {
            size_t fi = index / 4;
            if (fi >= proof.final_poly.size() || proof.final_poly[fi] != parent) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
        } else {
            size_t parent_pos = index / 4;
            if (query.folds[depth + 1].values[parent_pos % 4] != parent) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
        }
        index /= 4;
    }
}

// fri.rs:686-859
inline std::vector<size_t> verify(const Commitment& cm, const Proof& proof, const Params& params, Transcript& transcript) {
    params.validate();
    if (proof.queries.size() != params.num_queries) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    if (proof.final_poly.size() != params.final_size) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    if (cm.caps.empty()) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    if (cm.lde_len == 0 || !is_pow2(cm.lde_len) || cm.lde_len > (size_t(1) << Field::TWO_ADICITY))
        throw Z0pError(ERR_INVALID_DOMAIN_SIZE);
    if (cm.degree_bound >= cm.lde_len) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    if (cm.poly_len == 0 || !is_pow2(cm.poly_len)) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    if (cm.poly_len != cm.lde_len) {
        unsigned __int128 expected = (unsigned __int128)cm.poly_len * params.blowup;
        if (expected != cm.lde_len) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    }
    if (cm.lde_len < 2 || cm.degree_bound + 1 > cm.lde_len / 2) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    size_t folds = cm.caps.size() - 1;
    if (folds == 0 || !fold_shift_fits(folds) || (cm.lde_len >> (2 * folds)) != params.final_size)
        throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    for (size_t i = 0; i < cm.caps.size(); ++i) {
        size_t layer_len = cm.lde_len >> (2 * i);
        size_t tree_leaves = layer_len % 4 == 0 ? layer_len / 4 : layer_len;
        size_t depth = size_t(__builtin_ctzll(tree_leaves));
        size_t drop = std::min(params.cap_height, depth);
        if (cm.caps[i].size() != (size_t(1) << drop)) throw Z0pError(ERR_BAD_MERKLE_PATH);
    }
    transcript.absorb("fri-dims", encode_dims(cm.lde_len, cm.poly_len, cm.degree_bound));
    std::vector<std::array<Ext4, 3>> fc;
    for (size_t i = 0; i < cm.caps.size(); ++i) {
        for (auto& d : cm.caps[i]) transcript.absorb("fri-layer-root", d.data(), 32);
        if (i + 1 < cm.caps.size()) {
            std::array<Ext4, 3> ch;
            for (auto& c : ch) c = transcript.challenge_ext("fri-fold");
            fc.push_back(ch);
        }
    }
    transcript.absorb("fri-final", encode_fields(proof.final_poly));
    if (params.grinding_bits == 0) {
        if (proof.pow_nonce != 0) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    } else if (!transcript.verify_grind("fri-pow", proof.pow_nonce, params.grinding_bits)) {
        throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    }
    Field unit = two_adic_generator(2);
    size_t residual = cm.degree_bound >> (2 * folds);
    if (residual + 1 >= params.final_size) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    std::vector<Field> final_points;
    for (size_t j = 0; j < params.final_size; ++j) final_points.push_back(point_at(cm.lde_len, folds, j));
    if (!check_degree_ext(final_points, proof.final_poly, residual)) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    MerkleTree ft = commit_layer(proof.final_poly);
    if (cm.caps.back() != ft.cap(fri_drop_top(ft.num_leaves(), params.cap_height)))
        throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
    std::vector<size_t> indices;
    for (size_t i = 0; i < proof.queries.size(); ++i) {
        auto li = transcript.challenge_index("fri-query", cm.lde_len);
        if (!li) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
        indices.push_back(*li);
    }
    for (size_t i = 0; i < proof.queries.size(); ++i) {
        if (proof.queries[i].folds.size() != folds) throw Z0pError(ERR_FRI_VERIFICATION_FAILED);
        verify_one(cm, proof, fc, proof.queries[i], indices[i], unit, params.cap_height);
    }
    return indices;
}

}  // namespace fri
}  // namespace oracle
