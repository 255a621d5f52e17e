// ORACLE — test infrastructure only. This C entry-point layer exists so the
// Python tests and the benchmark's cpu_baseline leg can drive the CPU
// restatement through ctypes. Nothing in the product links it.
//
// Return codes follow src/error.rs declaration order (0 = Ok).
#include <cstdlib>
#include <cstring>
#include <string>

#include "air.hpp"
#include "range.hpp"
#include "table.hpp"
#include "fold.hpp"

using namespace oracle;

extern "C" {

typedef struct {
    uint64_t n_steps, blowup, num_queries, final_size;
    uint32_t grinding_bits, reserved0;
    uint64_t cap_height;
    uint32_t a0, b0, b_last, reserved1;
} oracle_air_params;

typedef struct {
    uint64_t blowup, num_queries, final_size;
    uint32_t grinding_bits, reserved0;
    uint64_t cap_height;
} oracle_fri_params;

typedef struct {
    uint64_t n_rows, blowup, num_queries, final_size;
    uint32_t grinding_bits, reserved0;
    uint64_t cap_height;
    uint32_t value, reserved1;
} oracle_table_params;

typedef struct {
    uint64_t seg_steps, blowup, num_queries, final_size;
    uint32_t grinding_bits, reserved0;
    uint64_t cap_height;
    uint32_t in_a, in_b, out_b, reserved1;
} oracle_fold_params;

typedef struct {
    uint64_t n_rows, n_bits, blowup, num_queries, final_size;
    uint32_t grinding_bits, reserved0;
    uint64_t cap_height;
    uint32_t value, reserved1;
} oracle_range_params;

}  // extern "C"

namespace {

air::Params to_air(const oracle_air_params* p) {
    air::Params q;
    q.n_steps = p->n_steps;
    q.blowup = p->blowup;
    q.num_queries = p->num_queries;
    q.final_size = p->final_size;
    q.grinding_bits = p->grinding_bits;
    q.cap_height = p->cap_height;
    q.a0 = Field::make(p->a0);
    q.b0 = Field::make(p->b0);
    q.b_last = Field::make(p->b_last);
    return q;
}

void from_air(const air::Params& q, oracle_air_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->n_steps = q.n_steps;
    p->blowup = q.blowup;
    p->num_queries = q.num_queries;
    p->final_size = q.final_size;
    p->grinding_bits = q.grinding_bits;
    p->cap_height = q.cap_height;
    p->a0 = q.a0.v;
    p->b0 = q.b0.v;
    p->b_last = q.b_last.v;
}

range::Params to_range(const oracle_range_params* p) {
    range::Params q;
    q.n_rows = p->n_rows;
    q.n_bits = p->n_bits;
    q.blowup = p->blowup;
    q.num_queries = p->num_queries;
    q.final_size = p->final_size;
    q.grinding_bits = p->grinding_bits;
    q.cap_height = p->cap_height;
    q.value = Field::make(p->value);
    return q;
}

void from_range(const range::Params& q, oracle_range_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->n_rows = q.n_rows;
    p->n_bits = q.n_bits;
    p->blowup = q.blowup;
    p->num_queries = q.num_queries;
    p->final_size = q.final_size;
    p->grinding_bits = q.grinding_bits;
    p->cap_height = q.cap_height;
    p->value = q.value.v;
}

table::Params to_table(const oracle_table_params* p) {
    table::Params q;
    q.n_rows = p->n_rows;
    q.blowup = p->blowup;
    q.num_queries = p->num_queries;
    q.final_size = p->final_size;
    q.grinding_bits = p->grinding_bits;
    q.cap_height = p->cap_height;
    q.value = Field::make(p->value);
    return q;
}

void from_table(const table::Params& q, oracle_table_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->n_rows = q.n_rows;
    p->blowup = q.blowup;
    p->num_queries = q.num_queries;
    p->final_size = q.final_size;
    p->grinding_bits = q.grinding_bits;
    p->cap_height = q.cap_height;
    p->value = q.value.v;
}

fold::Params to_fold(const oracle_fold_params* p) {
    fold::Params q;
    q.seg_steps = p->seg_steps;
    q.blowup = p->blowup;
    q.num_queries = p->num_queries;
    q.final_size = p->final_size;
    q.grinding_bits = p->grinding_bits;
    q.cap_height = p->cap_height;
    q.in_a = Field::make(p->in_a);
    q.in_b = Field::make(p->in_b);
    q.out_b = Field::make(p->out_b);
    return q;
}

void from_fold(const fold::Params& q, oracle_fold_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->seg_steps = q.seg_steps;
    p->blowup = q.blowup;
    p->num_queries = q.num_queries;
    p->final_size = q.final_size;
    p->grinding_bits = q.grinding_bits;
    p->cap_height = q.cap_height;
    p->in_a = q.in_a.v;
    p->in_b = q.in_b.v;
    p->out_b = q.out_b.v;
}

fri::Params to_fri(const oracle_fri_params* p) {
    fri::Params q;
    q.blowup = p->blowup;
    q.num_queries = p->num_queries;
    q.final_size = p->final_size;
    q.grinding_bits = p->grinding_bits;
    q.cap_height = p->cap_height;
    return q;
}

uint8_t* dup_bytes(const std::vector<uint8_t>& v, size_t* len) {
    uint8_t* out = static_cast<uint8_t*>(std::malloc(v.size() ? v.size() : 1));
    if (!out) return nullptr;
    std::memcpy(out, v.data(), v.size());
    *len = v.size();
    return out;
}

std::vector<Field> to_fields(const uint32_t* v, size_t n) {
    std::vector<Field> out(n);
    for (size_t i = 0; i < n; ++i) out[i] = Field::make(v[i]);
    return out;
}

}  // namespace

#define GUARD(...)                          \
    try {                                   \
        __VA_ARGS__;                        \
        return 0;                           \
    } catch (const Z0pError& e) {           \
        return e.code;                      \
    } catch (...) {                         \
        return 100;                         \
    }

extern "C" {

void oracle_set_threads(size_t n) { set_max_workers(n); }
size_t oracle_get_threads() { return max_workers(); }
void oracle_free(void* p) { std::free(p); }

void oracle_blake3(const uint8_t* data, size_t n, uint8_t* out) {
    Digest d = blake3_256(data, n);
    std::memcpy(out, d.data(), 32);
}
void oracle_blake3_small(const uint8_t* data, size_t n, uint8_t* out) {
    Digest d = blake3_256_small(data, n);
    std::memcpy(out, d.data(), 32);
}
void oracle_blake3_leaf64(const uint8_t* block, uint8_t* out) {
    Digest d = leaf64_digest(block);
    std::memcpy(out, d.data(), 32);
}
void oracle_blake3_node(const uint8_t* l, const uint8_t* r, uint8_t* out) {
    Digest a, b;
    std::memcpy(a.data(), l, 32);
    std::memcpy(b.data(), r, 32);
    Digest d = node_digest(a, b);
    std::memcpy(out, d.data(), 32);
}
// Streaming hasher fed with arbitrary splits (hash.rs Hasher::update).
void oracle_blake3_split(const uint8_t* data, size_t n, const size_t* splits, size_t nsplits, uint8_t* out) {
    Hasher h;
    size_t off = 0;
    for (size_t i = 0; i < nsplits && off < n; ++i) {
        size_t take = std::min(splits[i], n - off);
        h.update(data + off, take);
        off += take;
    }
    h.update(data + off, n - off);
    Digest d = h.finalize_consume();
    std::memcpy(out, d.data(), 32);
}

uint32_t oracle_field_mul(uint32_t a, uint32_t b) { return (Field::raw(a) * Field::raw(b)).v; }
uint32_t oracle_field_inv(uint32_t a) { return Field::raw(a).inv().v; }
uint32_t oracle_field_pow(uint32_t a, uint64_t e) { return Field::raw(a).pow(e).v; }
uint32_t oracle_field_halve(uint32_t a) { return Field::raw(a).halve().v; }
int oracle_two_adic_generator(uint32_t bits, uint32_t* out) { GUARD(*out = two_adic_generator(bits).v) }

void oracle_ext4_mul(const uint32_t* a, const uint32_t* b, uint32_t* out) {
    Ext4 x = Ext4::make(Field::raw(a[0]), Field::raw(a[1]), Field::raw(a[2]), Field::raw(a[3]));
    Ext4 y = Ext4::make(Field::raw(b[0]), Field::raw(b[1]), Field::raw(b[2]), Field::raw(b[3]));
    Ext4 z = x * y;
    for (int i = 0; i < 4; ++i) out[i] = z.c[i].v;
}
void oracle_ext4_inv(const uint32_t* a, uint32_t* out) {
    Ext4 x = Ext4::make(Field::raw(a[0]), Field::raw(a[1]), Field::raw(a[2]), Field::raw(a[3]));
    Ext4 z = x.inv();
    for (int i = 0; i < 4; ++i) out[i] = z.c[i].v;
}
// Fermat inverse x^(p^4 - 2), the closed-form check of ext4.rs:339-346.
void oracle_ext4_pow(const uint32_t* a, uint64_t e_hi, uint64_t e_lo, uint32_t* out) {
    Ext4 x = Ext4::make(Field::raw(a[0]), Field::raw(a[1]), Field::raw(a[2]), Field::raw(a[3]));
    unsigned __int128 e = ((unsigned __int128)e_hi << 64) | e_lo;
    Ext4 z = x.pow(e);
    for (int i = 0; i < 4; ++i) out[i] = z.c[i].v;
}

int oracle_ntt_forward(uint32_t* values, size_t n) {
    GUARD({
        std::vector<Field> v = to_fields(values, n);
        forward(v);
        for (size_t i = 0; i < n; ++i) values[i] = v[i].v;
    })
}
int oracle_ntt_inverse(uint32_t* values, size_t n) {
    GUARD({
        std::vector<Field> v = to_fields(values, n);
        inverse(v);
        for (size_t i = 0; i < n; ++i) values[i] = v[i].v;
    })
}
int oracle_ntt_evaluate(const uint32_t* coeffs, size_t n, uint32_t point, uint32_t* out) {
    GUARD(*out = evaluate(to_fields(coeffs, n), Field::make(point)).v)
}
int oracle_evaluate_on_coset(const uint32_t* coeffs, size_t n, size_t blowup, uint32_t* out) {
    GUARD({
        std::vector<Field> v = evaluate_on_coset(to_fields(coeffs, n), blowup, fri::lde_shift());
        for (size_t i = 0; i < v.size(); ++i) out[i] = v[i].v;
    })
}

// All tree layers leaf to root, concatenated (2*leaves - 1 digests for a
// power-of-two leaf count).
int oracle_merkle_flat64(const uint8_t* flat, size_t len, uint8_t* out_layers) {
    GUARD({
        MerkleTree t = MerkleTree::from_flat64(flat, len);
        size_t off = 0;
        for (auto& layer : t.layers)
            for (auto& d : layer) {
                std::memcpy(out_layers + off, d.data(), 32);
                off += 32;
            }
    })
}

void oracle_fold_values(const uint32_t* quad, uint32_t inv_point, const uint32_t* ch, uint32_t unit, uint32_t* out) {
    Ext4 q[4];
    for (int k = 0; k < 4; ++k)
        q[k] = Ext4::make(Field::raw(quad[4 * k]), Field::raw(quad[4 * k + 1]), Field::raw(quad[4 * k + 2]),
                          Field::raw(quad[4 * k + 3]));
    std::array<Ext4, 3> c;
    for (int k = 0; k < 3; ++k)
        c[k] = Ext4::make(Field::raw(ch[4 * k]), Field::raw(ch[4 * k + 1]), Field::raw(ch[4 * k + 2]),
                          Field::raw(ch[4 * k + 3]));
    Ext4 r = fri::fold_values(q, Field::raw(inv_point), c, Field::raw(unit));
    for (int i = 0; i < 4; ++i) out[i] = r.c[i].v;
}

int oracle_point_at(size_t lde_len, size_t depth, size_t j, uint32_t* out) { GUARD(*out = fri::point_at(lde_len, depth, j).v) }

int oracle_linked_lde(const uint32_t* lde_a, size_t m, size_t n, uint32_t actual_b_last, uint32_t a0, uint32_t* out) {
    GUARD({
        std::vector<Field> r = air::linked_lde(to_fields(lde_a, m), n, Field::make(actual_b_last), Field::make(a0));
        for (size_t i = 0; i < m; ++i) out[i] = r[i].v;
    })
}

// Batched composition over the natural-order LDE (air.rs:847-929); out is m
// Ext4 values as 4 u32 each. scalar=1 selects composition_at per point.
int oracle_compose(const uint32_t* lde_a, const uint32_t* lde_b, size_t m, size_t n, uint32_t a0, uint32_t b0,
                   uint32_t b_last, const uint32_t* alphas, int scalar, uint32_t* out) {
    GUARD({
        std::vector<Field> cols[2] = {to_fields(lde_a, m), to_fields(lde_b, m)};
        std::array<Ext4, 5> al;
        for (int k = 0; k < 5; ++k)
            al[k] = Ext4::make(Field::raw(alphas[4 * k]), Field::raw(alphas[4 * k + 1]), Field::raw(alphas[4 * k + 2]),
                               Field::raw(alphas[4 * k + 3]));
        std::vector<Field> pts = air::lde_points(m);
        Field w_last = primitive_root(n).pow(n - 1);
        std::vector<Ext4> r;
        if (scalar) {
            r.resize(m);
            size_t off = m / n;
            for (size_t i = 0; i < m; ++i) {
                size_t nx = (i + off) % m;
                Field c[2] = {cols[0][i], cols[1][i]};
                Field nn[2] = {cols[0][nx], cols[1][nx]};
                r[i] = air::composition_at(c, nn, pts[i], w_last, Field::make(a0), Field::make(b0), Field::make(b_last),
                                           n, al);
            }
        } else {
            r = air::compose(cols, pts, w_last, Field::make(a0), Field::make(b0), Field::make(b_last), n, m / n, al);
        }
        for (size_t i = 0; i < m; ++i)
            for (int k = 0; k < 4; ++k) out[4 * i + k] = r[i].c[k].v;
    })
}

void* oracle_transcript_new(const uint8_t* label, size_t n) { return new Transcript(label, n); }
void* oracle_transcript_clone(void* t) { return new Transcript(*static_cast<Transcript*>(t)); }
void oracle_transcript_free(void* t) { delete static_cast<Transcript*>(t); }
void oracle_transcript_absorb(void* t, const char* label, const uint8_t* data, size_t n) {
    static_cast<Transcript*>(t)->absorb(label, data, n);
}
void oracle_transcript_challenge_bytes(void* t, const char* label, uint8_t* out) {
    Digest d = static_cast<Transcript*>(t)->squeeze_raw(label);
    std::memcpy(out, d.data(), 32);
}
void oracle_transcript_challenge_ext(void* t, const char* label, uint32_t* out) {
    Ext4 e = static_cast<Transcript*>(t)->challenge_ext(label);
    for (int i = 0; i < 4; ++i) out[i] = e.c[i].v;
}
int oracle_transcript_challenge_index(void* t, const char* label, size_t domain, size_t* out) {
    auto r = static_cast<Transcript*>(t)->challenge_index(label, domain);
    if (!r) return ERR_MALFORMED_INPUT;
    *out = *r;
    return 0;
}
uint64_t oracle_transcript_grind(void* t, const char* label, uint32_t bits) {
    return static_cast<Transcript*>(t)->grind(label, bits);
}
int oracle_transcript_verify_grind(void* t, const char* label, uint64_t nonce, uint32_t bits) {
    return static_cast<Transcript*>(t)->verify_grind(label, nonce, bits) ? 1 : 0;
}
// Digest of the absorbed history (the proof-of-work head prefix).
void oracle_transcript_history_digest(void* t, uint8_t* out) {
    Digest d = static_cast<Transcript*>(t)->history.finalize();
    std::memcpy(out, d.data(), 32);
}
uint64_t oracle_grind_head(const uint8_t* head, size_t n, uint32_t bits) {
    std::vector<uint8_t> h(head, head + n);
    size_t workers = worker_count(size_t(1) << bits, 4096);
    return workers > 1 ? grind_parallel(h, bits, workers) : grind_sequential(h, bits);
}

int oracle_air_params_for_security(uint64_t n, uint32_t a0, uint32_t b0, uint32_t b_last, uint32_t bits,
                                   oracle_air_params* out) {
    GUARD(from_air(air::Params::for_security(n, Field::make(a0), Field::make(b0), Field::make(b_last), bits), out))
}
int oracle_air_params_new(uint64_t n, uint64_t blowup, uint64_t nq, uint64_t fs, uint32_t a0, uint32_t b0,
                          uint32_t b_last, oracle_air_params* out) {
    GUARD(from_air(air::Params::make(n, blowup, nq, fs, Field::make(a0), Field::make(b0), Field::make(b_last)), out))
}
int oracle_fri_params_for_security(uint64_t poly_len, uint32_t bits, oracle_fri_params* out) {
    GUARD({
        fri::Params p = fri::Params::for_security(poly_len, bits);
        std::memset(out, 0, sizeof(*out));
        out->blowup = p.blowup;
        out->num_queries = p.num_queries;
        out->final_size = p.final_size;
        out->grinding_bits = p.grinding_bits;
        out->cap_height = p.cap_height;
    })
}

uint32_t oracle_fib_b_last(uint32_t a0, uint32_t b0, uint64_t n) {
    auto cols = air::build_trace(Field::make(a0), Field::make(b0), n);
    return cols.second.empty() ? 0 : cols.second.back().v;
}
int oracle_build_trace(uint32_t a0, uint32_t b0, uint64_t n, uint32_t* col_a, uint32_t* col_b) {
    auto cols = air::build_trace(Field::make(a0), Field::make(b0), n);
    for (size_t i = 0; i < n; ++i) {
        col_a[i] = cols.first[i].v;
        col_b[i] = cols.second[i].v;
    }
    return 0;
}

int oracle_air_prove(const oracle_air_params* p, void* t, uint8_t** out, size_t* len) {
    GUARD({
        air::Proof pr = air::prove(to_air(p), *static_cast<Transcript*>(t));
        *out = dup_bytes(pr.to_bytes(), len);
    })
}
int oracle_air_prove_with_trace(const oracle_air_params* p, const uint32_t* col_a, const uint32_t* col_b, void* t,
                                uint8_t** out, size_t* len) {
    GUARD({
        air::Params ap = to_air(p);
        air::Proof pr = air::prove_with_trace(ap, to_fields(col_a, ap.n_steps), to_fields(col_b, ap.n_steps),
                                              *static_cast<Transcript*>(t));
        *out = dup_bytes(pr.to_bytes(), len);
    })
}
int oracle_air_prove_zk(const oracle_air_params* p, size_t n_rand, void* rng, void* t, uint8_t** out, size_t* len) {
    GUARD({
        air::Proof pr = air::prove_zk(to_air(p), n_rand, *static_cast<ChaCha20*>(rng), *static_cast<Transcript*>(t));
        *out = dup_bytes(pr.to_bytes(), len);
    })
}

// rng.rs — stream handle for the tests.
void* oracle_chacha20_new(const uint8_t* key, const uint8_t* nonce) { return new ChaCha20(key, nonce); }
void oracle_chacha20_free(void* r) { delete static_cast<ChaCha20*>(r); }
void oracle_chacha20_fill(void* r, uint8_t* out, size_t n) { static_cast<ChaCha20*>(r)->fill_bytes(out, n); }
uint32_t oracle_chacha20_next_u32(void* r) { return static_cast<ChaCha20*>(r)->next_u32(); }
uint64_t oracle_chacha20_next_u64(void* r) { return static_cast<ChaCha20*>(r)->next_u64(); }
uint32_t oracle_chacha20_next_field(void* r) { return static_cast<ChaCha20*>(r)->next_field().v; }
void oracle_chacha20_block(const uint8_t* key, uint32_t counter, const uint8_t* nonce, uint8_t* out) {
    chacha_block(key, counter, nonce, out);
}
void oracle_chacha20_quarter(uint32_t* state16, int a, int b, int c, int d) { chacha_quarter(state16, a, b, c, d); }
// blind.rs:38-79; out holds the padded length (returned in *out_len).
int oracle_mask_coeffs(const uint32_t* coeffs, size_t len, size_t n, size_t n_rand, void* rng, uint32_t* out,
                       size_t cap, size_t* out_len) {
    GUARD({
        std::vector<Field> m = mask_coeffs(to_fields(coeffs, len), n, n_rand, *static_cast<ChaCha20*>(rng));
        *out_len = m.size();
        if (m.size() > cap) throw Z0pError(ERR_INDEX_OUT_OF_BOUNDS);
        for (size_t i = 0; i < m.size(); ++i) out[i] = m[i].v;
    })
}
int oracle_air_verify(const oracle_air_params* p, void* t, const uint8_t* proof, size_t len) {
    GUARD({
        air::Proof pr = air::Proof::from_bytes(proof, len);
        air::verify(to_air(p), pr, *static_cast<Transcript*>(t));
    })
}
namespace {
std::vector<Ext4> to_ext(const uint32_t* v, size_t n) {
    std::vector<Ext4> out(n);
    for (size_t i = 0; i < n; ++i)
        out[i] = Ext4::make(Field::make(v[4 * i]), Field::make(v[4 * i + 1]), Field::make(v[4 * i + 2]),
                            Field::make(v[4 * i + 3]));
    return out;
}
// fri::open's indices (fri.rs:645-666) as u64 when idx is not null.
void put_idx(const std::vector<size_t>& q, uint64_t* idx) {
    if (idx)
        for (size_t i = 0; i < q.size(); ++i) idx[i] = q[i];
}
int fri_finish(fri::ProverState st, Transcript& tr, uint8_t** cm, size_t* cm_len, uint8_t** pf, size_t* pf_len,
               uint64_t* idx) {
    fri::Commitment c = fri::public_commitment(st);
    auto opened = fri::open(st, tr);
    *cm = dup_bytes(c.to_bytes(), cm_len);
    *pf = dup_bytes(opened.first.to_bytes(), pf_len);
    put_idx(opened.second, idx);
    return 0;
}
// kind -1: commit (base coefficients), 0: commit_ext (Ext4 coefficients),
// 1: commit_evals (base LDE values), 2: commit_ext_evals (Ext4 LDE values);
// Ext4 values are 4 words each.
fri::ProverState commit_kind(int kind, const uint32_t* vals, size_t n, size_t degree_bound, const oracle_fri_params* p,
                             Transcript& tr) {
    if (kind < 0) return fri::commit(to_fields(vals, n), to_fri(p), tr);
    if (kind == 0) return fri::commit_ext(to_ext(vals, n), to_fri(p), tr);
    if (kind == 1) return fri::commit_evals(to_fields(vals, n), degree_bound, to_fri(p), tr);
    return fri::commit_ext_evals_owned(to_ext(vals, n), degree_bound, to_fri(p), tr);
}
}  // namespace

int oracle_fri_commit_open(const uint32_t* coeffs, size_t n, const oracle_fri_params* p, void* t, uint8_t** cm,
                           size_t* cm_len, uint8_t** pf, size_t* pf_len, uint64_t* idx) {
    GUARD({
        Transcript& tr = *static_cast<Transcript*>(t);
        fri_finish(fri::commit(to_fields(coeffs, n), to_fri(p), tr), tr, cm, cm_len, pf, pf_len, idx);
    })
}

int oracle_fri_commit_open_kind(int kind, const uint32_t* vals, size_t n, size_t degree_bound,
                                const oracle_fri_params* p, void* t, uint8_t** cm, size_t* cm_len, uint8_t** pf,
                                size_t* pf_len, uint64_t* idx) {
    GUARD({
        Transcript& tr = *static_cast<Transcript*>(t);
        fri_finish(commit_kind(kind, vals, n, degree_bound, p, tr), tr, cm, cm_len, pf, pf_len, idx);
    })
}

// fri::commit* alone (fri.rs:408-549): an owned ProverState for a later
// oracle_fri_state_open; kinds as commit_kind.
int oracle_fri_commit_state(int kind, const uint32_t* vals, size_t n, size_t degree_bound, const oracle_fri_params* p,
                            void* t, void** out) {
    GUARD({
        *out = new fri::ProverState(commit_kind(kind, vals, n, degree_bound, p, *static_cast<Transcript*>(t)));
    })
}
int oracle_fri_state_commitment(void* st, uint8_t** cm, size_t* cm_len) {
    GUARD({ *cm = dup_bytes(fri::public_commitment(*static_cast<fri::ProverState*>(st)).to_bytes(), cm_len); })
}
int oracle_fri_state_open(void* st, void* t, uint8_t** pf, size_t* pf_len, uint64_t* idx) {
    GUARD({
        auto opened = fri::open(*static_cast<fri::ProverState*>(st), *static_cast<Transcript*>(t));
        *pf = dup_bytes(opened.first.to_bytes(), pf_len);
        put_idx(opened.second, idx);
    })
}
void oracle_fri_state_free(void* st) { delete static_cast<fri::ProverState*>(st); }
// ntt.rs:394-420, coordinates interleaved as 4 words per value.
int oracle_evaluate_on_coset_ext(const uint32_t* coeffs4, size_t n, size_t blowup, uint32_t* out4) {
    GUARD({
        std::vector<Ext4> v = evaluate_on_coset_ext(to_ext(coeffs4, n), blowup, fri::lde_shift());
        for (size_t i = 0; i < v.size(); ++i)
            for (int c = 0; c < 4; ++c) out4[4 * i + c] = v[i].c[c].v;
    })
}

int oracle_fri_verify(const uint8_t* cm, size_t cm_len, const uint8_t* pf, size_t pf_len, const oracle_fri_params* p,
                      void* t, uint64_t* idx) {
    GUARD({
        fri::Commitment c = fri::Commitment::from_bytes(cm, cm_len);
        fri::Proof pr = fri::Proof::from_bytes(pf, pf_len);
        put_idx(fri::verify(c, pr, to_fri(p), *static_cast<Transcript*>(t)), idx);
    })
}

// range.rs: parameters, the three provers and the verifier. Columns arrive
// as n_columns blocks of `len` words, V first.
int oracle_range_params_new(uint64_t n_rows, uint64_t n_bits, uint64_t blowup, uint64_t nq, uint64_t fs,
                            uint32_t value, oracle_range_params* out) {
    GUARD(from_range(range::Params::make(n_rows, n_bits, blowup, nq, fs, Field::make(value)), out))
}
int oracle_range_params_for_security(uint64_t n_rows, uint64_t n_bits, uint32_t value, uint32_t bits,
                                     oracle_range_params* out) {
    GUARD(from_range(range::Params::for_security(n_rows, n_bits, Field::make(value), bits), out))
}
int oracle_range_prove(const oracle_range_params* p, void* t, uint8_t** out, size_t* len) {
    GUARD({
        range::Proof pr = range::prove(to_range(p), *static_cast<Transcript*>(t));
        *out = dup_bytes(pr.to_bytes(), len);
    })
}
int oracle_range_prove_with_columns(const oracle_range_params* p, const uint32_t* cols, size_t ncols, size_t len,
                                    void* t, uint8_t** out, size_t* out_len) {
    GUARD({
        std::vector<Field> v = ncols ? to_fields(cols, len) : std::vector<Field>();
        std::vector<std::vector<Field>> bits;
        for (size_t c = 1; c < ncols; ++c) bits.push_back(to_fields(cols + c * len, len));
        range::Proof pr = range::prove_with_columns(to_range(p), v, bits, *static_cast<Transcript*>(t));
        *out = dup_bytes(pr.to_bytes(), out_len);
    })
}
// range.rs:436 directly: coefficient columns with a declared degree bound.
int oracle_range_prove_coeffs(const oracle_range_params* p, const uint32_t* cols, size_t ncols, size_t len,
                              size_t degree_bound, void* t, uint8_t** out, size_t* out_len) {
    GUARD({
        range::Params rp = to_range(p);
        rp.validate();
        std::vector<std::vector<Field>> c;
        for (size_t k = 0; k < ncols; ++k) c.push_back(to_fields(cols + k * len, len));
        range::Proof pr = range::prove_from_columns(rp, c, degree_bound, *static_cast<Transcript*>(t));
        *out = dup_bytes(pr.to_bytes(), out_len);
    })
}
int oracle_range_prove_zk(const oracle_range_params* p, size_t n_rand, void* rng, void* t, uint8_t** out,
                          size_t* len) {
    GUARD({
        range::Proof pr =
            range::prove_zk(to_range(p), n_rand, *static_cast<ChaCha20*>(rng), *static_cast<Transcript*>(t));
        *out = dup_bytes(pr.to_bytes(), len);
    })
}
int oracle_range_verify(const oracle_range_params* p, void* t, const uint8_t* proof, size_t len) {
    GUARD({
        range::Proof pr = range::Proof::from_bytes(proof, len);
        range::verify(to_range(p), pr, *static_cast<Transcript*>(t));
    })
}
// range.rs:752 at one point, for the composition pins.
int oracle_range_composition_at(uint32_t v, const uint32_t* bits, size_t n_bits, uint32_t point, uint32_t value,
                                uint64_t n, const uint32_t* alphas4, uint32_t* out4) {
    GUARD({
        std::vector<Ext4> al = to_ext(alphas4, n_bits + 2);
        std::vector<Ext4> ab(al.begin() + 1, al.begin() + 1 + n_bits);
        Ext4 r = range::composition_at(Field::make(v), to_fields(bits, n_bits), Field::make(point), Field::make(value),
                                       n, al[0], ab, al[n_bits + 1]);
        for (int c = 0; c < 4; ++c) out4[c] = r.c[c].v;
    })
}

// table.rs: parameters, the provers and the verifier; the table is `tlen`
// words (reduced mod p).
int oracle_table_params_new(uint64_t n_rows, uint64_t blowup, uint64_t nq, uint64_t fs, uint32_t value,
                            oracle_table_params* out) {
    GUARD(from_table(table::Params::make(n_rows, blowup, nq, fs, Field::make(value)), out))
}
int oracle_table_params_for_security(uint64_t n_rows, uint64_t table_len, uint32_t value, uint32_t bits,
                                     oracle_table_params* out) {
    GUARD(from_table(table::Params::for_security(n_rows, table_len, Field::make(value), bits), out))
}
uint64_t oracle_table_membership_bound(uint64_t table_len, uint64_t n, uint64_t n_rand) {
    return table::membership_bound(table_len, n, n_rand);
}
int oracle_table_prove(const oracle_table_params* p, const uint32_t* tab, size_t tlen, void* t, uint8_t** out,
                       size_t* len) {
    GUARD({
        table::Proof pr = table::prove(to_table(p), to_fields(tab, tlen), *static_cast<Transcript*>(t));
        *out = dup_bytes(pr.to_bytes(), len);
    })
}
int oracle_table_prove_coeffs(const oracle_table_params* p, const uint32_t* tab, size_t tlen, const uint32_t* coeffs,
                              size_t clen, size_t degree_bound, void* t, uint8_t** out, size_t* len) {
    GUARD({
        table::Params tp = to_table(p);
        table::Proof pr = table::prove_from_coeffs(tp, to_fields(tab, tlen), to_fields(coeffs, clen), degree_bound,
                                                   *static_cast<Transcript*>(t));
        *out = dup_bytes(pr.to_bytes(), len);
    })
}
int oracle_table_prove_zk(const oracle_table_params* p, const uint32_t* tab, size_t tlen, size_t n_rand, void* rng,
                          void* t, uint8_t** out, size_t* len) {
    GUARD({
        table::Proof pr = table::prove_zk(to_table(p), to_fields(tab, tlen), n_rand, *static_cast<ChaCha20*>(rng),
                                          *static_cast<Transcript*>(t));
        *out = dup_bytes(pr.to_bytes(), len);
    })
}
int oracle_table_verify(const oracle_table_params* p, const uint32_t* tab, size_t tlen, void* t,
                        const uint8_t* proof, size_t len) {
    GUARD({
        table::Proof pr = table::Proof::from_bytes(proof, len);
        table::verify(to_table(p), to_fields(tab, tlen), pr, *static_cast<Transcript*>(t));
    })
}

// fold.rs: parameters, the accumulator and the verifier.
int oracle_fold_params_new(uint64_t seg, uint64_t blowup, uint64_t nq, uint64_t fs, uint32_t in_a, uint32_t in_b,
                           uint32_t out_b, oracle_fold_params* out) {
    GUARD(from_fold(fold::Params::make(seg, blowup, nq, fs, Field::make(in_a), Field::make(in_b), Field::make(out_b)),
                    out))
}
int oracle_fold_params_for_security(uint64_t seg, uint32_t in_a, uint32_t in_b, uint32_t out_b, uint32_t bits,
                                    oracle_fold_params* out) {
    GUARD(from_fold(fold::Params::for_security(seg, Field::make(in_a), Field::make(in_b), Field::make(out_b), bits),
                    out))
}
void* oracle_fold_acc_new(const oracle_fold_params* p) { return new fold::Accumulator(to_fold(p)); }
void oracle_fold_acc_free(void* a) { delete static_cast<fold::Accumulator*>(a); }
size_t oracle_fold_acc_len(void* a) { return static_cast<fold::Accumulator*>(a)->len(); }
// Accumulator::params is fixed at creation; tests change it through this
// call to drive the reference's error paths.
void oracle_fold_acc_set_params(void* a, const oracle_fold_params* p) {
    static_cast<fold::Accumulator*>(a)->params = to_fold(p);
}
int oracle_fold_acc_add_segment(void* a, const uint32_t* col_a, size_t len_a, const uint32_t* col_b, size_t len_b,
                                void* t) {
    GUARD(static_cast<fold::Accumulator*>(a)->add_segment(to_fields(col_a, len_a), to_fields(col_b, len_b),
                                                          *static_cast<Transcript*>(t)))
}
int oracle_fold_acc_finalize(void* a, void* t, uint8_t** out, size_t* len) {
    GUARD({
        fold::IvcProof pr = static_cast<fold::Accumulator*>(a)->finalize(*static_cast<Transcript*>(t));
        *out = dup_bytes(pr.to_bytes(), len);
    })
}
// fold.rs:927-992 (forged_segment_a_jump_rejected): step `k`'s out_a claim
// replaced, the transcript replayed from `label` and every composition
// recomputed, so the proof is internally consistent apart from the claim.
int oracle_fold_acc_finalize_forged_out_a(void* a, size_t k, uint32_t out_a, const uint8_t* label, size_t label_len,
                                          uint8_t** out, size_t* len) {
    GUARD({
        fold::Accumulator forged = *static_cast<fold::Accumulator*>(a);
        if (k >= forged.steps.size()) throw Z0pError(ERR_INDEX_OUT_OF_BOUNDS);
        forged.steps[k].claims.out_a = Field::make(out_a);
        Transcript tr(label, label_len);
        fold::absorb_params(forged.params, tr);
        size_t s = forged.params.seg_steps, m = forged.params.lde_len();
        Field w_last = primitive_root(s).pow(s - 1);
        size_t drop = capped_shape(m, forged.params.cap_height).first;
        for (auto& st : forged.steps) {
            fold::absorb_step(st.claims, tr);
            tr.absorb("fold/trace", cap_bytes(st.tree.cap(drop)));
            std::array<Ext4, fold::ALPHAS> al;
            for (auto& x : al) x = tr.challenge_ext("fold/alpha");
            st.composition = fold::compose_step(st.ldes, w_last, st.claims, s, m / s, al);
        }
        *out = dup_bytes(forged.finalize(tr).to_bytes(), len);
    })
}
int oracle_fold_verify(const oracle_fold_params* p, void* t, const uint8_t* proof, size_t len) {
    GUARD({
        fold::IvcProof pr = fold::IvcProof::from_bytes(proof, len);
        fold::verify(to_fold(p), pr, *static_cast<Transcript*>(t));
    })
}

void oracle_stage_timers_clear() { timers().clear(); }
// Writes "name=ms;name=ms;..." into buf; returns the needed length.
size_t oracle_stage_timers(char* buf, size_t cap) {
    std::string s;
    for (auto& kv : timers().ms) s += kv.first + "=" + std::to_string(kv.second) + ";";
    if (buf && cap) {
        size_t n = std::min(cap - 1, s.size());
        std::memcpy(buf, s.data(), n);
        buf[n] = 0;
    }
    return s.size() + 1;
}

}  // extern "C"
