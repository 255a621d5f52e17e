// Host verifier for the AIR, range-check, table-membership, IVC and FRI
// proofs (the reference verifies on the CPU: air.rs:607-752,
// range.rs:521-721, table.rs:430-538, fold.rs:572-727, fri.rs:686-859,
// 1153-1227, 1339-1364). It parses the
// VERSION = 6 bytes fail-closed (codec.rs:118-274) and replays the
// transcript exactly as the prover advanced it.
#include <algorithm>
#include <array>
#include <cstring>

#include "blake3.hpp"
#include "host.hpp"
#include "params.hpp"

namespace z0p {

namespace {

using Digest = std::array<uint8_t, 32>;
using Path = std::vector<std::pair<Digest, bool>>;

struct CommitmentV {
    std::vector<std::vector<Digest>> caps;
    uint64_t lde_len = 0, poly_len = 0, degree_bound = 0;
};
struct FoldW {
    E4 values[4];
    Path path;
};
struct FriProofV {
    std::vector<E4> final_poly;
    uint64_t nonce = 0;
    std::vector<std::vector<FoldW>> queries;
};
struct TraceOpeningV {
    uint64_t index = 0;
    uint32_t a = 0, b = 0, an = 0, bn = 0;
    Path path, shift_path;
    uint8_t leaf[64], shift_leaf[64];
};

Digest read_digest(Reader& r) {
    Digest d;
    std::memcpy(d.data(), r.take(32), 32);
    return d;
}
std::vector<Digest> read_cap(Reader& r) {
    std::vector<Digest> c;
    uint32_t n = r.count(MAX_CAP_DIGESTS);
    for (uint32_t i = 0; i < n; ++i) c.push_back(read_digest(r));
    return c;
}
Path read_path(Reader& r) {
    Path p;
    uint32_t n = r.count(MAX_PATH_STEPS);
    for (uint32_t i = 0; i < n; ++i) {
        Digest d = read_digest(r);
        bool dir = r.direction();
        p.emplace_back(d, dir);
    }
    return p;
}
CommitmentV read_commitment_body(Reader& r) {
    CommitmentV c;
    c.lde_len = r.u64();
    c.poly_len = r.u64();
    c.degree_bound = r.u64();
    uint32_t n = r.count(MAX_FOLDS);
    for (uint32_t i = 0; i < n; ++i) c.caps.push_back(read_cap(r));
    return c;
}
FriProofV read_fri_proof_body(Reader& r) {
    FriProofV p;
    uint32_t nf = r.count(MAX_ELEMS);
    for (uint32_t i = 0; i < nf; ++i) p.final_poly.push_back(r.e4());
    p.nonce = r.u64();
    uint32_t nq = r.count(MAX_QUERIES);
    for (uint32_t q = 0; q < nq; ++q) {
        std::vector<FoldW> folds;
        uint32_t nfo = r.count(MAX_FOLDS);
        for (uint32_t k = 0; k < nfo; ++k) {
            FoldW fw;
            for (int v = 0; v < 4; ++v) fw.values[v] = r.e4();
            fw.path = read_path(r);
            folds.push_back(std::move(fw));
        }
        p.queries.push_back(std::move(folds));
    }
    return p;
}

void digest_of_words(const uint32_t w[8], Digest& d) {
    for (int k = 0; k < 8; ++k) store_le32(d.data() + 4 * k, w[k]);
}
void words_of_digest(const Digest& d, uint32_t w[8]) {
    for (int k = 0; k < 8; ++k) w[k] = load_le32(d.data() + 4 * k);
}
Digest hash_node(const Digest& l, const Digest& r) {
    uint32_t a[8], b[8], o[8];
    words_of_digest(l, a);
    words_of_digest(r, b);
    node_words(a, b, o);
    Digest d;
    digest_of_words(o, d);
    return d;
}
Digest leaf64(const uint8_t block[64]) {
    uint32_t m[16], o[8];
    for (int k = 0; k < 16; ++k) m[k] = load_le32(block + 4 * k);
    leaf64_words(m, o);
    Digest d;
    digest_of_words(o, d);
    return d;
}
Digest quad_leaf_digest(const E4 v[4]) {
    uint8_t b[64];
    for (int k = 0; k < 4; ++k)
        for (int c = 0; c < 4; ++c) store_le32(b + 16 * k + 4 * c, v[k].v[c]);
    return leaf64(b);
}

// merkle.rs:254-277 and 320-324.
bool directions_match(uint64_t index, const Path& path) {
    for (size_t l = 0; l < path.size(); ++l)
        if (path[l].second != (((index >> l) & 1) == 0)) return false;
    return true;
}
bool verify_capped64(const Digest& entry, const uint8_t leaf[64], uint64_t index, uint64_t total_depth,
                     const Path& path) {
    if (total_depth >= 64 || path.size() > total_depth || (index >> total_depth) != 0 || !directions_match(index, path))
        return false;
    Digest cur = leaf64(leaf);
    for (auto& st : path) cur = st.second ? hash_node(cur, st.first) : hash_node(st.first, cur);
    return cur == entry;
}

// Small host Merkle cap of a final layer (fri.rs:1401-1428).
std::vector<Digest> layer_cap(const std::vector<E4>& vals, uint64_t cap_height) {
    std::vector<Digest> level;
    if (vals.size() % 4 == 0) {
        for (size_t i = 0; i < vals.size(); i += 4) level.push_back(quad_leaf_digest(&vals[i]));
    } else {
        for (const E4& v : vals) {
            uint8_t buf[17];
            buf[0] = 0;
            for (int c = 0; c < 4; ++c) store_le32(buf + 1 + 4 * c, v.v[c]);
            Digest d;
            blake3_hash(buf, 17, d.data());
            level.push_back(d);
        }
    }
    size_t target = 1;
    while (target < level.size()) target <<= 1;
    while (level.size() < target) level.push_back(level.back());
    uint64_t depth = ilog2(level.size());
    uint64_t drop = std::min(cap_height, depth);
    for (uint64_t l = 0; l < depth - drop; ++l) {
        std::vector<Digest> next(level.size() / 2);
        for (size_t i = 0; i < next.size(); ++i) next[i] = hash_node(level[2 * i], level[2 * i + 1]);
        level.swap(next);
    }
    return level;
}

uint64_t brev64(uint64_t x, uint64_t bits) {
    uint64_t r = 0;
    for (uint64_t i = 0; i < bits; ++i) {
        r = (r << 1) | (x & 1);
        x >>= 1;
    }
    return r;
}

// The points of point_at (fri.rs:1108-1115) without a power chain per call: for layer
// depth d of an LDE of 2^log_m points, point_at(m, d, j) =
// g^(4^d) * omega_m^(brev(j, log_m - 2d) * 4^d), with omega_m^e read from
// three 512-entry tables (e < 2^27) and g^(4^d) kept per depth.
struct DomainPoints {
    uint64_t log_m;
    uint32_t t0[512], t1[512], t2[512];
    std::vector<uint32_t> g4;  // g^(4^d)
    explicit DomainPoints(uint64_t lm) : log_m(lm) {
        const uint32_t w = root_of_unity(uint32_t(lm));
        const uint32_t w9 = pow_c(w, uint64_t(1) << 9), w18 = pow_c(w, uint64_t(1) << 18);
        t0[0] = t1[0] = t2[0] = 1;
        for (int i = 1; i < 512; ++i) {
            t0[i] = mul(t0[i - 1], w);
            t1[i] = mul(t1[i - 1], w9);
            t2[i] = mul(t2[i - 1], w18);
        }
        uint32_t g = GENERATOR;
        for (uint64_t d = 0; 2 * d <= lm; ++d) {
            g4.push_back(g);
            g = mul(mul(g, g), mul(g, g));
        }
    }
    uint32_t at(uint64_t depth, uint64_t j) const {
        const uint64_t e = brev64(j, log_m - 2 * depth) << (2 * depth);
        return mul(g4[depth], mul(t0[e & 511], mul(t1[(e >> 9) & 511], t2[e >> 18])));
    }
};

// In-place inverses of nonzero values (one field inversion for the batch).
void batch_inverse(std::vector<uint32_t>& v) {
    if (v.empty()) return;
    std::vector<uint32_t> pre(v.size());
    uint32_t acc = 1;
    for (size_t i = 0; i < v.size(); ++i) {
        pre[i] = acc;
        acc = mul(acc, v[i]);
    }
    uint32_t inv = inv_c(acc);
    for (size_t i = v.size(); i-- > 0;) {
        const uint32_t vi = v[i];
        v[i] = mul(inv, pre[i]);
        inv = mul(inv, vi);
    }
}

// fri.rs:1311-1331 with canonical arithmetic.
E4 fold_values(const E4 q[4], uint32_t inv_point, const E4 ch[3], uint32_t unit) {
    E4 even_sum = e4_add(e4_add(q[0], q[1]), e4_add(q[2], q[3]));
    E4 odd_diff = e4_sub(q[0], q[1]);
    E4 skew_diff = e4_sub(q[2], q[3]);
    E4 su = e4_mul_base(skew_diff, unit);
    E4 mean = e4_halve(e4_halve(even_sum));
    E4 odd = e4_halve(e4_halve(e4_sub(odd_diff, su)));
    E4 even = e4_halve(e4_halve(e4_sub(e4_add(q[0], q[1]), e4_add(q[2], q[3]))));
    E4 skew = e4_halve(e4_halve(e4_add(odd_diff, su)));
    E4 acc = e4_mul_base(e4_add(e4_mul_base(e4_mul(ch[2], skew), inv_point), e4_mul(ch[1], even)), inv_point);
    acc = e4_add(acc, e4_mul(ch[0], odd));
    return e4_add(mean, e4_mul_base(acc, inv_point));
}

// fri.rs:1339-1364
bool check_degree_ext(const std::vector<uint32_t>& pts, const std::vector<E4>& vals, uint64_t bound) {
    if (pts.size() != vals.size() || vals.empty()) return false;
    uint64_t n = vals.size();
    if (bound >= n - 1) return true;
    std::vector<E4> t = vals;
    uint64_t last = std::min(bound + 1, n - 1);
    for (uint64_t order = 1; order <= last; ++order) {
        for (uint64_t i = 0; i < n - order; ++i) {
            uint32_t den = sub(pts[i + order], pts[i]);
            if (den == 0) return false;
            t[i] = e4_mul_base(e4_sub(t[i + 1], t[i]), inv_c(den));
        }
        if (order > bound)
            for (uint64_t i = 0; i < n - order; ++i)
                if (!e4_eq(t[i], e4_zero())) return false;
    }
    return true;
}

void absorb_fri_dims(Transcript& tr, uint64_t m, uint64_t poly_len, uint64_t bound) {
    uint8_t d[24];
    store_le64(d, m);
    store_le64(d + 8, poly_len);
    store_le64(d + 16, bound);
    tr.absorb("fri-dims", d, 24);
}

// fri::verify (fri.rs:686-859); returns the query indices.
// Records a 64-byte-leaf path check for the device (directions already
// checked against the position by the caller).
void defer_path(PathChecks& d, const uint8_t leaf[64], const Path& path, const Digest& expected) {
    PathCheck c{};
    std::memcpy(c.leaf, leaf, 64);
    std::memcpy(c.expected, expected.data(), 32);
    c.steps = uint32_t(path.size());
    c.sib = uint32_t(d.siblings.size() / 8);
    for (size_t l = 0; l < path.size(); ++l) {
        if (path[l].second) c.dirs |= uint64_t(1) << l;
        uint32_t w[8];
        words_of_digest(path[l].first, w);
        d.siblings.insert(d.siblings.end(), w, w + 8);
    }
    d.checks.push_back(c);
}

std::vector<uint64_t> fri_verify_impl(const CommitmentV& cm, const FriProofV& pf, const z0p_fri_params& fp,
                                      Transcript& tr, PathChecks* defer = nullptr) {
    validate_fri(fp);
    require(pf.queries.size() == fp.num_queries, Z0P_ERR_FRI_VERIFICATION_FAILED);
    require(pf.final_poly.size() == fp.final_size, Z0P_ERR_FRI_VERIFICATION_FAILED);
    require(!cm.caps.empty(), Z0P_ERR_FRI_VERIFICATION_FAILED);
    require(pow2(cm.lde_len) && cm.lde_len <= MAX_DOMAIN, Z0P_ERR_INVALID_DOMAIN_SIZE);
    require(cm.degree_bound < cm.lde_len, Z0P_ERR_FRI_VERIFICATION_FAILED);
    require(pow2(cm.poly_len), Z0P_ERR_FRI_VERIFICATION_FAILED);
    if (cm.poly_len != cm.lde_len) {
        unsigned __int128 e = (unsigned __int128)cm.poly_len * fp.blowup;
        require(e == cm.lde_len, Z0P_ERR_FRI_VERIFICATION_FAILED);
    }
    require(cm.lde_len >= 2 && cm.degree_bound + 1 <= cm.lde_len / 2, Z0P_ERR_FRI_VERIFICATION_FAILED);
    const uint64_t folds = cm.caps.size() - 1;
    require(folds != 0 && folds <= 30 && (cm.lde_len >> (2 * folds)) == fp.final_size,
            Z0P_ERR_FRI_VERIFICATION_FAILED);
    for (size_t i = 0; i < cm.caps.size(); ++i) {
        uint64_t len = cm.lde_len >> (2 * i);
        uint64_t drop = std::min<uint64_t>(fp.cap_height, ilog2(fri_tree_leaves(len)));
        require(cm.caps[i].size() == (uint64_t(1) << drop), Z0P_ERR_BAD_MERKLE_PATH);
    }
    absorb_fri_dims(tr, cm.lde_len, cm.poly_len, cm.degree_bound);
    std::vector<std::array<E4, 3>> fc;
    for (size_t i = 0; i < cm.caps.size(); ++i) {
        for (auto& d : cm.caps[i]) tr.absorb("fri-layer-root", d.data(), 32);
        if (i + 1 < cm.caps.size()) {
            std::array<E4, 3> ch;
            for (auto& c : ch) c = tr.challenge_ext("fri-fold");
            fc.push_back(ch);
        }
    }
    {
        std::vector<uint8_t> fin(pf.final_poly.size() * 16);
        for (size_t i = 0; i < pf.final_poly.size(); ++i)
            for (int c = 0; c < 4; ++c) store_le32(fin.data() + 16 * i + 4 * c, pf.final_poly[i].v[c]);
        tr.absorb("fri-final", fin);
    }
    if (fp.grinding_bits == 0) {
        require(pf.nonce == 0, Z0P_ERR_FRI_VERIFICATION_FAILED);
    } else {
        std::vector<uint8_t> head = tr.pow_head("fri-pow");
        uint8_t d[32];
        pow_digest(head, pf.nonce, d);
        require(leading_zero_bits(d) >= fp.grinding_bits, Z0P_ERR_FRI_VERIFICATION_FAILED);
        uint8_t nb[8];
        store_le64(nb, pf.nonce);
        tr.absorb("fri-pow", nb, 8);
    }
    const uint32_t unit = root_of_unity(2);
    const uint64_t residual = cm.degree_bound >> (2 * folds);
    require(residual + 1 < fp.final_size, Z0P_ERR_FRI_VERIFICATION_FAILED);
    const DomainPoints dom(ilog2(cm.lde_len));
    std::vector<uint32_t> fpts;
    for (uint64_t j = 0; j < fp.final_size; ++j) fpts.push_back(dom.at(folds, j));
    require(check_degree_ext(fpts, pf.final_poly, residual), Z0P_ERR_FRI_VERIFICATION_FAILED);
    require(layer_cap(pf.final_poly, fp.cap_height) == cm.caps.back(), Z0P_ERR_FRI_VERIFICATION_FAILED);
    std::vector<uint64_t> idx;
    for (size_t q = 0; q < pf.queries.size(); ++q) idx.push_back(tr.challenge_index("fri-query", cm.lde_len));
    // 1 / x for every (query, depth) whose base index is in range, inverted
    // as one batch; the loop below reads them in the same order.
    std::vector<uint32_t> xinv;
    for (size_t q = 0; q < pf.queries.size(); ++q) {
        uint64_t index = idx[q];
        for (uint64_t d = 0; d < std::min<uint64_t>(pf.queries[q].size(), folds); ++d) {
            const uint64_t base = index / 4 * 4;
            if (base + 3 >= (cm.lde_len >> (2 * d))) break;
            xinv.push_back(dom.at(d, base));
            index /= 4;
        }
    }
    batch_inverse(xinv);
    size_t xi = 0;
    for (size_t q = 0; q < pf.queries.size(); ++q) {
        const auto& folds_w = pf.queries[q];
        require(folds_w.size() == folds, Z0P_ERR_FRI_VERIFICATION_FAILED);
        uint64_t index = idx[q];
        for (uint64_t d = 0; d < folds_w.size(); ++d) {
            const FoldW& w = folds_w[d];
            const auto& cap = cm.caps[d];
            uint64_t len = cm.lde_len >> (2 * d);
            uint64_t base = index / 4 * 4;
            require(base + 3 < len, Z0P_ERR_FRI_VERIFICATION_FAILED);
            uint64_t depth_bits = ilog2(len / 4);
            uint64_t drop = std::min<uint64_t>(fp.cap_height, depth_bits);
            require(w.path.size() == depth_bits - drop, Z0P_ERR_BAD_MERKLE_PATH);
            uint64_t qi = base / 4;
            uint64_t e = qi >> (depth_bits - drop);
            require(e < cap.size(), Z0P_ERR_BAD_MERKLE_PATH);
            const uint32_t x_inv = xinv[xi++];
            if (defer) {
                uint64_t pos = qi;
                for (auto& st : w.path) {
                    require(st.second == ((pos & 1) == 0), Z0P_ERR_BAD_MERKLE_PATH);
                    pos >>= 1;
                }
                uint8_t leaf[64];
                for (int k = 0; k < 4; ++k)
                    for (int c = 0; c < 4; ++c) store_le32(leaf + 16 * k + 4 * c, w.values[k].v[c]);
                defer_path(*defer, leaf, w.path, cap[e]);
            } else {
                Digest node = quad_leaf_digest(w.values);
                uint64_t pos = qi;
                for (auto& st : w.path) {
                    require(st.second == ((pos & 1) == 0), Z0P_ERR_BAD_MERKLE_PATH);
                    node = st.second ? hash_node(node, st.first) : hash_node(st.first, node);
                    pos >>= 1;
                }
                require(node == cap[e], Z0P_ERR_BAD_MERKLE_PATH);
            }
            E4 parent = fold_values(w.values, x_inv, fc[d].data(), unit);
            if (d + 1 == folds_w.size()) {
                uint64_t fi = index / 4;
                require(fi < pf.final_poly.size() && e4_eq(pf.final_poly[fi], parent), Z0P_ERR_FRI_VERIFICATION_FAILED);
            } else {
                require(e4_eq(folds_w[d + 1].values[(index / 4) % 4], parent), Z0P_ERR_FRI_VERIFICATION_FAILED);
            }
            index /= 4;
        }
    }
    return idx;
}

// air.rs:938-962
E4 composition_at(uint32_t a, uint32_t b, uint32_t an, uint32_t bn, uint32_t x, uint32_t w_last, uint32_t a0,
                  uint32_t b0, uint32_t b_last, uint64_t n, const E4 al[5]) {
    uint32_t van = sub(pow_c(x, n), 1);
    uint32_t z = mul(van, inv_c(sub(x, w_last)));
    uint32_t zi = inv_c(z);
    uint32_t t1 = mul(sub(an, b), zi);
    uint32_t t2 = mul(sub(bn, add(a, b)), zi);
    uint32_t i1 = inv_c(sub(x, 1));
    uint32_t ba = mul(sub(a, a0), i1);
    uint32_t bb = mul(sub(b, b0), i1);
    uint32_t bo = mul(sub(b, b_last), inv_c(sub(x, w_last)));
    E4 r = e4_mul_base(al[0], t1);
    r = e4_add(r, e4_mul_base(al[1], t2));
    r = e4_add(r, e4_mul_base(al[2], ba));
    r = e4_add(r, e4_mul_base(al[3], bb));
    r = e4_add(r, e4_mul_base(al[4], bo));
    return r;
}

// range.rs:752-774
E4 range_composition_at(uint32_t v, const std::vector<uint32_t>& bits, uint32_t x, uint32_t value, uint64_t n,
                        const std::vector<E4>& al) {
    const uint32_t zi = inv_c(sub(pow_c(x, n), 1));
    uint32_t rec = 0, power = 1;
    for (uint32_t b : bits) {
        rec = add(rec, mul(power, b));
        power = add(power, power);
    }
    E4 r = e4_mul_base(al[0], mul(sub(v, rec), zi));
    for (size_t j = 0; j < bits.size(); ++j)
        r = e4_add(r, e4_mul_base(al[1 + j], mul(mul(bits[j], sub(bits[j], 1)), zi)));
    return e4_add(r, e4_mul_base(al[bits.size() + 1], mul(sub(v, value), inv_c(sub(x, 1)))));
}

// merkle.rs:286-309: a variable-length leaf, BLAKE3(0x00 || data).
bool verify_capped_row(const Digest& entry, const std::vector<uint8_t>& row, uint64_t index, uint64_t total_depth,
                       const Path& path) {
    if (total_depth >= 64 || path.size() > total_depth || (index >> total_depth) != 0 ||
        !directions_match(index, path))
        return false;
    std::vector<uint8_t> buf(row.size() + 1, 0);
    std::memcpy(buf.data() + 1, row.data(), row.size());
    Digest cur;
    blake3_hash(buf.data(), buf.size(), cur.data());
    for (auto& st : path) cur = st.second ? hash_node(cur, st.first) : hash_node(st.first, cur);
    return cur == entry;
}

}  // namespace

std::vector<uint64_t> fri_verify(const uint8_t* cmb, size_t cm_len, const uint8_t* pfb, size_t pf_len,
                                 const z0p_fri_params& fp, Transcript& tr) {
    Reader rc(cmb, cm_len);
    rc.version();
    CommitmentV cm = read_commitment_body(rc);
    rc.end();
    Reader rp(pfb, pf_len);
    rp.version();
    FriProofV pf = read_fri_proof_body(rp);
    rp.end();
    return fri_verify_impl(cm, pf, fp, tr);
}

namespace {

// The checks of verify_capped64 that come before the hashing.
bool capped64_shape_ok(uint64_t index, uint64_t total_depth, const Path& path) {
    return total_depth < 64 && path.size() <= total_depth && (index >> total_depth) == 0 &&
           directions_match(index, path);
}

void air_verify_impl(const z0p_air_params& p, Transcript& tr, const uint8_t* bytes, size_t len, PathChecks* defer) {
    // Proof::from_bytes (air.rs:289-327)
    Reader r(bytes, len);
    r.version();
    std::vector<Digest> column_cap = read_cap(r);
    CommitmentV cm = read_commitment_body(r);
    FriProofV pf = read_fri_proof_body(r);
    std::vector<TraceOpeningV> ops;
    uint32_t no = r.count(MAX_QUERIES);
    for (uint32_t i = 0; i < no; ++i) {
        TraceOpeningV o;
        o.index = r.u64();
        o.a = r.field();
        o.b = r.field();
        o.an = r.field();
        o.bn = r.field();
        o.path = read_path(r);
        o.shift_path = read_path(r);
        std::memcpy(o.leaf, r.take(64), 64);
        std::memcpy(o.shift_leaf, r.take(64), 64);
        ops.push_back(std::move(o));
    }
    r.end();

    validate_air(p);
    const uint64_t n = p.n_steps;
    const uint64_t m = cm.lde_len;
    const uint64_t expected_m = n * p.blowup;  // validated: <= 2^27
    require(m >= expected_m && m % expected_m == 0, Z0P_ERR_FRI_VERIFICATION_FAILED);
    if (m == expected_m) {
        require(cm.degree_bound == n - 1, Z0P_ERR_FRI_VERIFICATION_FAILED);
    } else {
        require(cm.degree_bound < m / p.blowup, Z0P_ERR_FRI_VERIFICATION_FAILED);
    }
    require(cm.poly_len == m, Z0P_ERR_FRI_VERIFICATION_FAILED);
    const uint64_t offset = m / n;
    const uint32_t a0 = p.a0 % P, b0 = p.b0 % P, b_last = p.b_last % P;
    uint8_t pb[56];
    store_le64(pb, p.n_steps);
    store_le64(pb + 8, p.blowup);
    store_le64(pb + 16, p.num_queries);
    store_le64(pb + 24, p.final_size);
    store_le32(pb + 32, p.grinding_bits);
    store_le64(pb + 36, p.cap_height);
    store_le32(pb + 44, a0);
    store_le32(pb + 48, b0);
    store_le32(pb + 52, b_last);
    tr.absorb("fib/params", pb, 56);
    require(ops.size() == p.num_queries, Z0P_ERR_FRI_VERIFICATION_FAILED);
    {
        std::vector<uint8_t> cb;
        for (auto& d : column_cap) cb.insert(cb.end(), d.begin(), d.end());
        tr.absorb("fib/trace", cb);
    }
    E4 al[5];
    for (auto& a : al) a = tr.challenge_ext("fib/alpha");
    const z0p_fri_params fp = air_fri_params(p);
    std::vector<uint64_t> idx = fri_verify_impl(cm, pf, fp, tr, defer);
    require(pow2(m), Z0P_ERR_FRI_VERIFICATION_FAILED);
    const uint32_t w_last = pow_c(root_of_unity(ilog2(n)), n - 1);
    const uint64_t log_m = ilog2(m);
    const uint32_t omega_m = root_of_unity(uint32_t(log_m));
    size_t count = std::min(std::min(pf.queries.size(), ops.size()), idx.size());
    for (size_t q = 0; q < count; ++q) {
        const TraceOpeningV& o = ops[q];
        uint64_t leaf = idx[q];
        uint64_t index = brev64(leaf, log_m);
        require(o.index == index, Z0P_ERR_FRI_VERIFICATION_FAILED);
        require(m % 8 == 0, Z0P_ERR_FRI_VERIFICATION_FAILED);
        uint64_t leaves = m / 8;
        uint64_t depth = ilog2(leaves);
        uint64_t drop = std::min<uint64_t>(p.cap_height, depth);
        require(column_cap.size() == (uint64_t(1) << drop) && o.path.size() == depth - drop &&
                    o.shift_path.size() == depth - drop,
                Z0P_ERR_BAD_MERKLE_PATH);
        uint64_t ro = (index & 7) * 8;
        require(load_le32(o.leaf + ro) == o.a && load_le32(o.leaf + ro + 4) == o.b, Z0P_ERR_BAD_MERKLE_PATH);
        uint64_t e = (index / 8) >> (depth - drop);
        require(e < column_cap.size(), Z0P_ERR_BAD_MERKLE_PATH);
        if (defer) {
            require(capped64_shape_ok(index / 8, depth, o.path), Z0P_ERR_BAD_MERKLE_PATH);
            defer_path(*defer, o.leaf, o.path, column_cap[e]);
        } else {
            require(verify_capped64(column_cap[e], o.leaf, index / 8, depth, o.path), Z0P_ERR_BAD_MERKLE_PATH);
        }
        uint64_t shifted = (index + offset) % m;
        uint64_t so = (shifted & 7) * 8;
        require(load_le32(o.shift_leaf + so) == o.an && load_le32(o.shift_leaf + so + 4) == o.bn,
                Z0P_ERR_BAD_MERKLE_PATH);
        uint64_t se = (shifted / 8) >> (depth - drop);
        require(se < column_cap.size(), Z0P_ERR_BAD_MERKLE_PATH);
        if (defer) {
            require(capped64_shape_ok(shifted / 8, depth, o.shift_path), Z0P_ERR_BAD_MERKLE_PATH);
            defer_path(*defer, o.shift_leaf, o.shift_path, column_cap[se]);
        } else {
            require(verify_capped64(column_cap[se], o.shift_leaf, shifted / 8, depth, o.shift_path),
                    Z0P_ERR_BAD_MERKLE_PATH);
        }
        uint32_t x = mul(GENERATOR, pow_c(omega_m, index));
        E4 expected = composition_at(o.a, o.b, o.an, o.bn, x, w_last, a0, b0, b_last, n, al);
        require(!pf.queries[q].empty(), Z0P_ERR_FRI_VERIFICATION_FAILED);
        require(e4_eq(pf.queries[q][0].values[leaf % 4], expected), Z0P_ERR_FRI_VERIFICATION_FAILED);
    }
}

}  // namespace

void air_verify(const z0p_air_params& p, Transcript& tr, const uint8_t* bytes, size_t len) {
    air_verify_impl(p, tr, bytes, len, nullptr);
}

void air_verify_deferred(const z0p_air_params& p, Transcript& tr, const uint8_t* bytes, size_t len,
                         PathChecks& defer) {
    air_verify_impl(p, tr, bytes, len, &defer);
}

void range_verify(const z0p_range_params& p, Transcript& tr, const uint8_t* bytes, size_t len) {
    // Proof::from_bytes (range.rs:280-306)
    struct RowOpening {
        uint64_t index = 0;
        uint32_t v = 0;
        std::vector<uint32_t> bits;
        Path path;
    };
    Reader r(bytes, len);
    r.version();
    std::vector<Digest> column_cap = read_cap(r);
    CommitmentV cm = read_commitment_body(r);
    FriProofV pf = read_fri_proof_body(r);
    std::vector<RowOpening> ops;
    uint32_t no = r.count(MAX_QUERIES);
    for (uint32_t i = 0; i < no; ++i) {
        RowOpening o;
        o.index = r.u64();
        o.v = r.field();
        uint32_t nb = r.count(MAX_ELEMS);
        for (uint32_t k = 0; k < nb; ++k) o.bits.push_back(r.field());
        o.path = read_path(r);
        ops.push_back(std::move(o));
    }
    r.end();

    // range::verify (range.rs:521-721)
    validate_range(p);
    const uint64_t n = p.n_rows;
    const uint64_t m = cm.lde_len;
    const uint64_t expected_m = n * p.blowup;  // validated: <= 2^27
    require(m >= expected_m && m % expected_m == 0, Z0P_ERR_FRI_VERIFICATION_FAILED);
    if (m == expected_m) {
        require(cm.degree_bound == n - 1, Z0P_ERR_FRI_VERIFICATION_FAILED);
    } else {
        // masked columns: at most the largest honest masked degree
        const uint64_t padded = m / p.blowup;
        require(2 * padded >= n + 2 && cm.degree_bound <= 2 * padded - (n + 2), Z0P_ERR_FRI_VERIFICATION_FAILED);
    }
    require(cm.poly_len == m, Z0P_ERR_FRI_VERIFICATION_FAILED);
    const uint32_t value = p.value % P;
    uint8_t pb[56];
    store_le64(pb, p.n_rows);
    store_le64(pb + 8, p.n_bits);
    store_le64(pb + 16, p.blowup);
    store_le64(pb + 24, p.num_queries);
    store_le64(pb + 32, p.final_size);
    store_le32(pb + 40, p.grinding_bits);
    store_le64(pb + 44, p.cap_height);
    store_le32(pb + 52, value);
    tr.absorb("range/params", pb, 56);
    require(ops.size() == p.num_queries, Z0P_ERR_FRI_VERIFICATION_FAILED);
    {
        std::vector<uint8_t> cb;
        for (auto& d : column_cap) cb.insert(cb.end(), d.begin(), d.end());
        tr.absorb("range/column", cb);
    }
    std::vector<E4> al(p.n_bits + 2);
    for (auto& a : al) a = tr.challenge_ext("range/alpha");
    const z0p_fri_params fp = range_fri_params(p);
    std::vector<uint64_t> idx = fri_verify_impl(cm, pf, fp, tr);
    require(pow2(m), Z0P_ERR_FRI_VERIFICATION_FAILED);
    const uint64_t log_m = ilog2(m);
    const uint32_t omega_m = root_of_unity(uint32_t(log_m));
    const size_t count = std::min(std::min(pf.queries.size(), ops.size()), idx.size());
    for (size_t q = 0; q < count; ++q) {
        const RowOpening& o = ops[q];
        const uint64_t leaf = idx[q];
        const uint64_t index = brev64(leaf, log_m);
        require(o.index == index && o.bits.size() == p.n_bits, Z0P_ERR_FRI_VERIFICATION_FAILED);
        std::vector<uint8_t> row(4 * (p.n_bits + 1));
        store_le32(row.data(), o.v);
        for (size_t j = 0; j < o.bits.size(); ++j) store_le32(row.data() + 4 * (j + 1), o.bits[j]);
        const uint64_t depth = log_m;
        const uint64_t drop = std::min<uint64_t>(p.cap_height, depth);
        require(column_cap.size() == (uint64_t(1) << drop) && o.path.size() == depth - drop, Z0P_ERR_BAD_MERKLE_PATH);
        const uint64_t e = index >> (depth - drop);
        require(e < column_cap.size(), Z0P_ERR_BAD_MERKLE_PATH);
        require(verify_capped_row(column_cap[e], row, index, depth, o.path), Z0P_ERR_BAD_MERKLE_PATH);
        const uint32_t x = mul(GENERATOR, pow_c(omega_m, index));
        const E4 expected = range_composition_at(o.v, o.bits, x, value, n, al);
        require(!pf.queries[q].empty(), Z0P_ERR_FRI_VERIFICATION_FAILED);
        require(e4_eq(pf.queries[q][0].values[leaf % 4], expected), Z0P_ERR_FRI_VERIFICATION_FAILED);
    }
}

void table_verify(const z0p_table_params& p, const uint32_t* table_in, uint64_t table_len, Transcript& tr,
                  const uint8_t* bytes, size_t len) {
    // Proof::from_bytes (table.rs:265-288)
    struct ValueOpening {
        uint64_t index = 0;
        uint32_t v = 0;
        Path path;
    };
    Reader r(bytes, len);
    r.version();
    std::vector<Digest> column_cap = read_cap(r);
    CommitmentV cm = read_commitment_body(r);
    FriProofV pf = read_fri_proof_body(r);
    std::vector<ValueOpening> ops;
    uint32_t no = r.count(MAX_QUERIES);
    for (uint32_t i = 0; i < no; ++i) {
        ValueOpening o;
        o.index = r.u64();
        o.v = r.field();
        o.path = read_path(r);
        ops.push_back(std::move(o));
    }
    r.end();

    // table::verify (table.rs:430-538)
    validate_table(p);
    require(table_len >= 2 && table_in, Z0P_ERR_MALFORMED_INPUT);
    std::vector<uint32_t> table(table_len);
    for (uint64_t k = 0; k < table_len; ++k) table[k] = table_in[k] % P;
    const uint64_t n = p.n_rows;
    const uint64_t m = cm.lde_len;
    const uint64_t expected_m = n * p.blowup;  // validated: <= 2^27
    require(m >= expected_m && m % expected_m == 0, Z0P_ERR_FRI_VERIFICATION_FAILED);
    if (m == expected_m) {
        require(cm.degree_bound == membership_bound(table_len, n, 0), Z0P_ERR_FRI_VERIFICATION_FAILED);
    } else {
        // masked column: p = m / blowup coefficients bound the honest degree
        const uint64_t padded = m / p.blowup;
        const unsigned __int128 prod = (unsigned __int128)table_len * (padded == 0 ? 0 : padded - 1);
        require(prod <= UINT64_MAX && prod >= n && cm.degree_bound <= uint64_t(prod) - n,
                Z0P_ERR_FRI_VERIFICATION_FAILED);
    }
    require(cm.poly_len == m, Z0P_ERR_FRI_VERIFICATION_FAILED);
    const uint32_t value = p.value % P;
    std::vector<uint8_t> pb(56 + 4 * table_len);
    store_le64(pb.data(), p.n_rows);
    store_le64(pb.data() + 8, p.blowup);
    store_le64(pb.data() + 16, p.num_queries);
    store_le64(pb.data() + 24, p.final_size);
    store_le32(pb.data() + 32, p.grinding_bits);
    store_le64(pb.data() + 36, p.cap_height);
    store_le32(pb.data() + 44, value);
    store_le64(pb.data() + 48, table_len);
    for (uint64_t k = 0; k < table_len; ++k) store_le32(pb.data() + 56 + 4 * k, table[k]);
    tr.absorb("table/params", pb);
    require(ops.size() == p.num_queries, Z0P_ERR_FRI_VERIFICATION_FAILED);
    {
        std::vector<uint8_t> cb;
        for (auto& d : column_cap) cb.insert(cb.end(), d.begin(), d.end());
        tr.absorb("table/column", cb);
    }
    const E4 a_member = tr.challenge_ext("table/alpha");
    const E4 a_bound = tr.challenge_ext("table/alpha");
    const z0p_fri_params fp = table_fri_params(p);
    std::vector<uint64_t> idx = fri_verify_impl(cm, pf, fp, tr);
    require(pow2(m), Z0P_ERR_FRI_VERIFICATION_FAILED);
    const uint64_t log_m = ilog2(m);
    const uint32_t omega_m = root_of_unity(uint32_t(log_m));
    const size_t count = std::min(std::min(pf.queries.size(), ops.size()), idx.size());
    for (size_t q = 0; q < count; ++q) {
        const ValueOpening& o = ops[q];
        const uint64_t leaf = idx[q];
        const uint64_t index = brev64(leaf, log_m);
        require(o.index == index, Z0P_ERR_FRI_VERIFICATION_FAILED);
        const uint64_t depth = log_m;
        const uint64_t drop = std::min<uint64_t>(p.cap_height, depth);
        require(column_cap.size() == (uint64_t(1) << drop) && o.path.size() == depth - drop, Z0P_ERR_BAD_MERKLE_PATH);
        const uint64_t e = index >> (depth - drop);
        require(e < column_cap.size(), Z0P_ERR_BAD_MERKLE_PATH);
        std::vector<uint8_t> row(4);
        store_le32(row.data(), o.v);
        require(verify_capped_row(column_cap[e], row, index, depth, o.path), Z0P_ERR_BAD_MERKLE_PATH);
        // table.rs:646-663
        const uint32_t x = mul(GENERATOR, pow_c(omega_m, index));
        uint32_t prod = 1;
        for (uint32_t t : table) prod = mul(prod, sub(o.v, t));
        const uint32_t s_m = mul(prod, inv_c(sub(pow_c(x, n), 1)));
        const uint32_t s_b = mul(sub(o.v, value), inv_c(sub(x, 1)));
        const E4 expected = e4_add(e4_mul_base(a_member, s_m), e4_mul_base(a_bound, s_b));
        require(!pf.queries[q].empty(), Z0P_ERR_FRI_VERIFICATION_FAILED);
        require(e4_eq(pf.queries[q][0].values[leaf % 4], expected), Z0P_ERR_FRI_VERIFICATION_FAILED);
    }
}

void fold_verify(const z0p_fold_params& p, Transcript& tr, const uint8_t* bytes, size_t len) {
    // IvcProof::from_bytes (fold.rs:312-359)
    struct StepOpeningV {
        uint64_t index = 0;
        uint32_t cols[4] = {0, 0, 0, 0};
        Path path, shift_path;
    };
    struct StepV {
        uint32_t claims[4];  // in_a, in_b, out_a, out_b
        std::vector<Digest> cap;
        std::vector<StepOpeningV> openings;
    };
    constexpr uint32_t MAX_STEPS = 1u << 20;  // codec.rs:34
    // Vectors grow one decoded item at a time: memory stays proportional to
    // the bytes actually read, never to an untrusted count (codec.rs:10-13).
    Reader r(bytes, len);
    r.version();
    std::vector<StepV> steps;
    const uint32_t nsteps = r.count(MAX_STEPS);
    for (uint32_t k = 0; k < nsteps; ++k) {
        StepV st;
        for (auto& c : st.claims) c = r.field();
        st.cap = read_cap(r);
        const uint32_t nops = r.count(MAX_QUERIES);
        for (uint32_t i = 0; i < nops; ++i) {
            StepOpeningV o;
            o.index = r.u64();
            for (auto& v : o.cols) v = r.field();
            o.path = read_path(r);
            o.shift_path = read_path(r);
            st.openings.push_back(std::move(o));
        }
        steps.push_back(std::move(st));
    }
    CommitmentV cm = read_commitment_body(r);
    FriProofV pf = read_fri_proof_body(r);
    r.end();

    // fold::verify (fold.rs:572-667)
    validate_fold(p);
    require(!steps.empty(), Z0P_ERR_EMPTY_INPUT);
    const uint64_t s = p.seg_steps, m = s * p.blowup;
    require(cm.lde_len == m && cm.poly_len == m && cm.degree_bound == s - 1, Z0P_ERR_FRI_VERIFICATION_FAILED);
    const uint64_t offset = m / s;
    uint8_t pb[56];
    store_le64(pb, p.seg_steps);
    store_le64(pb + 8, p.blowup);
    store_le64(pb + 16, p.num_queries);
    store_le64(pb + 24, p.final_size);
    store_le32(pb + 32, p.grinding_bits);
    store_le64(pb + 36, p.cap_height);
    const uint32_t in_a = p.in_a % P, in_b = p.in_b % P, out_b = p.out_b % P;
    store_le32(pb + 44, in_a);
    store_le32(pb + 48, in_b);
    store_le32(pb + 52, out_b);
    tr.absorb("fold/params", pb, 56);
    // check_chain (fold.rs:710-727)
    require(steps[0].claims[0] == in_a && steps[0].claims[1] == in_b, Z0P_ERR_FRI_VERIFICATION_FAILED);
    for (size_t k = 0; k + 1 < steps.size(); ++k)
        require(steps[k].claims[2] == steps[k + 1].claims[0] && steps[k].claims[3] == steps[k + 1].claims[1],
                Z0P_ERR_FRI_VERIFICATION_FAILED);
    require(steps.back().claims[3] == out_b, Z0P_ERR_FRI_VERIFICATION_FAILED);
    std::vector<std::array<E4, 6>> alphas(steps.size());
    for (size_t k = 0; k < steps.size(); ++k) {
        uint8_t cb[16];
        for (int i = 0; i < 4; ++i) store_le32(cb + 4 * i, steps[k].claims[i]);
        tr.absorb("fold/step", cb, 16);
        std::vector<uint8_t> capb;
        for (auto& d : steps[k].cap) capb.insert(capb.end(), d.begin(), d.end());
        tr.absorb("fold/trace", capb);
        for (auto& a : alphas[k]) a = tr.challenge_ext("fold/alpha");
        require(steps[k].openings.size() == p.num_queries, Z0P_ERR_FRI_VERIFICATION_FAILED);
    }
    std::vector<E4> gammas(steps.size());
    for (auto& g : gammas) g = tr.challenge_ext("fold/gamma");
    const z0p_fri_params fp = fold_fri_params(p);
    std::vector<uint64_t> idx = fri_verify_impl(cm, pf, fp, tr);
    const uint64_t log_m = ilog2(m);
    const uint32_t omega_m = root_of_unity(uint32_t(log_m));
    const uint32_t w_last = pow_c(root_of_unity(ilog2(s)), s - 1);
    const uint64_t depth = log_m, drop = std::min<uint64_t>(p.cap_height, depth);
    const size_t count = std::min(pf.queries.size(), idx.size());
    for (size_t q = 0; q < count; ++q) {
        const uint64_t leaf = idx[q];
        const uint64_t index = brev64(leaf, log_m);
        const uint32_t x = mul(GENERATOR, pow_c(omega_m, index));
        E4 combined{{0, 0, 0, 0}};
        for (size_t k = 0; k < steps.size(); ++k) {
            const StepV& st = steps[k];
            const StepOpeningV& o = st.openings[q];
            require(o.index == index, Z0P_ERR_FRI_VERIFICATION_FAILED);
            // check_step_openings (fold.rs:669-708)
            require(st.cap.size() == (uint64_t(1) << drop) && o.path.size() == depth - drop &&
                        o.shift_path.size() == depth - drop,
                    Z0P_ERR_BAD_MERKLE_PATH);
            std::vector<uint8_t> row(8);
            store_le32(row.data(), o.cols[0]);
            store_le32(row.data() + 4, o.cols[1]);
            const uint64_t e = index >> (depth - drop);
            require(e < st.cap.size() && verify_capped_row(st.cap[e], row, index, depth, o.path),
                    Z0P_ERR_BAD_MERKLE_PATH);
            const uint64_t shifted = (index + offset) % m;
            store_le32(row.data(), o.cols[2]);
            store_le32(row.data() + 4, o.cols[3]);
            const uint64_t se = shifted >> (depth - drop);
            require(se < st.cap.size() && verify_capped_row(st.cap[se], row, shifted, depth, o.shift_path),
                    Z0P_ERR_BAD_MERKLE_PATH);
            // composition_at (fold.rs:856-879)
            const uint32_t a = o.cols[0], b = o.cols[1], an = o.cols[2], bn = o.cols[3];
            const uint32_t zi = inv_c(mul(sub(pow_c(x, s), 1), inv_c(sub(x, w_last))));
            const uint32_t i1 = inv_c(sub(x, 1)), iw = inv_c(sub(x, w_last));
            const uint32_t terms[6] = {mul(sub(an, b), zi),
                                       mul(sub(bn, add(a, b)), zi),
                                       mul(sub(a, st.claims[0]), i1),
                                       mul(sub(b, st.claims[1]), i1),
                                       mul(sub(a, st.claims[2]), iw),
                                       mul(sub(b, st.claims[3]), iw)};
            E4 c{{0, 0, 0, 0}};
            for (int t = 0; t < 6; ++t) c = e4_add(c, e4_mul_base(alphas[k][t], terms[t]));
            combined = e4_add(combined, e4_mul(gammas[k], c));
        }
        require(!pf.queries[q].empty(), Z0P_ERR_FRI_VERIFICATION_FAILED);
        require(e4_eq(pf.queries[q][0].values[leaf % 4], combined), Z0P_ERR_FRI_VERIFICATION_FAILED);
    }
}

// Generic byte-leaf Merkle trees (merkle.rs:16-30, 165-235).
void merkle_hash_leaf(const uint8_t* data, size_t len, uint8_t out[32]) {
    Blake3 h;
    const uint8_t prefix = 0x00;
    h.update(&prefix, 1);
    h.update(data, len);
    h.digest(out);
}

void merkle_hash_node(const uint8_t left[32], const uint8_t right[32], uint8_t out[32]) {
    Digest l, r;
    std::memcpy(l.data(), left, 32);
    std::memcpy(r.data(), right, 32);
    const Digest d = hash_node(l, r);
    std::memcpy(out, d.data(), 32);
}

// `path` holds `steps` entries of 33 bytes, sibling digest then the
// direction byte (1: the running node is the left child), as on the wire.
bool merkle_verify_digest(const uint8_t root[32], const uint8_t leaf_digest[32], uint64_t index, const uint8_t* path,
                          size_t steps) {
    if (steps >= 64 || (index >> steps) != 0) return false;
    Digest cur;
    std::memcpy(cur.data(), leaf_digest, 32);
    for (size_t l = 0; l < steps; ++l) {
        const uint8_t* st = path + 33 * l;
        if (st[32] > 1) return false;
        const bool is_left = st[32] == 1;
        if (is_left != (((index >> l) & 1) == 0)) return false;
        Digest sib;
        std::memcpy(sib.data(), st, 32);
        cur = is_left ? hash_node(cur, sib) : hash_node(sib, cur);
    }
    return std::memcmp(cur.data(), root, 32) == 0;
}

}  // namespace z0p
