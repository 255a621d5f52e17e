// Host orchestration of the GPU prover: Fibonacci AIR (air.rs:365-599) and
// the FRI commitment (fri.rs:408-666), driving the kernels on the context
// stream and the Fiat-Shamir transcript on the host.
//
// The transcript is the only serial dependency: each FRI layer's cap must be
// absorbed before its three fold challenges exist (fri.rs:594-612), the
// proof-of-work gates the query indices (fri.rs:645-656). Each such point is
// one small device-to-host copy plus a stream synchronisation; all bulk data
// stays resident in device memory.
#include "prover.hpp"

#include <algorithm>
#include <cstring>

#include "blake3.hpp"
#include "params.hpp"

namespace z0p {

uint32_t brev_host(uint64_t x, int bits) {
    uint64_t r = 0;
    for (int i = 0; i < bits; ++i) {
        r = (r << 1) | (x & 1);
        x >>= 1;
    }
    return uint32_t(r);
}

namespace {

// (F(k), F(k+1)) canonical by fast doubling.
void fib_pair(uint64_t k, uint32_t& f, uint32_t& g) {
    f = 0;
    g = 1;
    for (int b = 63; b >= 0; --b) {
        uint32_t f2 = mul(f, sub(dbl(g), f));
        uint32_t g2 = add(mul(f, f), mul(g, g));
        if ((k >> b) & 1) {
            f = g2;
            g = add(f2, g2);
        } else {
            f = f2;
            g = g2;
        }
    }
}

template <typename T>
T* upload(std::vector<void*>& owned, const std::vector<T>& host) {
    void* p = nullptr;
    hip_check(hipMalloc(&p, std::max<size_t>(host.size(), 1) * sizeof(T)));
    owned.push_back(p);
    if (!host.empty()) hip_check(hipMemcpy(p, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice));
    return static_cast<T*>(p);
}

// Montgomery-form geometric sequence base^0..base^(len-1).
std::vector<uint32_t> geometric_m(uint32_t base, size_t len) {
    std::vector<uint32_t> out(len);
    uint32_t bm = to_mont(base), acc = R1;
    for (size_t i = 0; i < len; ++i) {
        out[i] = acc;
        acc = mont_mul(acc, bm);
    }
    return out;
}

}  // namespace

uint32_t fib_b_last(uint32_t a0, uint32_t b0, uint64_t n) {
    // B[n-1] = F(n-1) a0 + F(n) b0
    uint32_t f, g;
    fib_pair(n - 1, f, g);
    return add(mul(f, a0), mul(g, b0));
}

const GeometryTables& Ctx::geometry(int logn, int logb) {
    auto key = std::make_pair(logn, logb);
    auto it = tables.find(key);
    if (it != tables.end()) return *it->second;
    auto t = std::make_unique<GeometryTables>();
    t->logn = logn;
    t->logb = logb;
    const int logm = logn + logb;
    const uint64_t n = uint64_t(1) << logn, B = uint64_t(1) << logb, m = n << logb;
    const uint32_t wm = root_of_unity(uint32_t(logm));
    const uint32_t wm_inv = inv_c(wm);
    const size_t hi_len = std::max<uint64_t>(1, m >> 12);
    t->w.lo = upload(t->owned, geometric_m(wm, 4096));
    t->w.hi = upload(t->owned, geometric_m(pow_c(wm, 4096), hi_len));
    t->winv.lo = upload(t->owned, geometric_m(wm_inv, 4096));
    t->winv.hi = upload(t->owned, geometric_m(pow_c(wm_inv, 4096), hi_len));
    std::vector<uint32_t> shift(B), zinv(B);
    for (uint64_t j = 0; j < B; ++j) {
        uint32_t s = mul(GENERATOR, pow_c(wm, j));
        shift[j] = to_mont(s);
        zinv[j] = to_mont(inv_c(sub(pow_c(s, n), 1)));
    }
    t->coset.shift = upload(t->owned, shift);
    t->zinv = upload(t->owned, zinv);
    NttShape sh = ntt_shape(logn);
    if (sh.log1 == 0) {
        std::vector<uint32_t> small;
        small.reserve(B * n);
        for (uint64_t j = 0; j < B; ++j) {
            auto g = geometric_m(from_mont(shift[j]), n);
            small.insert(small.end(), g.begin(), g.end());
        }
        t->coset.small = upload(t->owned, small);
    } else {
        const uint64_t n1 = uint64_t(1) << sh.log1, n2 = uint64_t(1) << sh.log2;
        std::vector<uint32_t> row(B * n2), col;
        col.reserve(B * n1);
        for (uint64_t j = 0; j < B; ++j) {
            uint32_t s = from_mont(shift[j]);
            auto g = geometric_m(pow_c(s, n1), n2);
            for (uint64_t c = 0; c < n2; ++c) row[j * n2 + c] = g[brev_host(c, sh.log2)];
            auto gc = geometric_m(s, n1);
            col.insert(col.end(), gc.begin(), gc.end());
        }
        t->coset.row = upload(t->owned, row);
        t->coset.col = upload(t->owned, col);
    }
    auto* raw = t.get();
    tables[key] = std::move(t);
    return *raw;
}

Ctx::~Ctx() {
    tables.clear();
    if (arena) (void)hipFree(arena);
    if (pinned) (void)hipHostFree(pinned);
    if (tw4096) (void)hipFree(tw4096);
    if (tw4096_inv) (void)hipFree(tw4096_inv);
    if (stream) (void)hipStreamDestroy(stream);
}

Ctx* ctx_create(int device) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0 || device < 0 || device >= count)
        throw Status(Z0P_ERR_NO_DEVICE);
    hip_check(hipSetDevice(device));
    auto c = std::make_unique<Ctx>();
    c->device = device;
    hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    const uint32_t w = root_of_unity(12), wi = inv_c(w);
    auto fwd = geometric_m(w, 2048), inv = geometric_m(wi, 2048);
    hip_check(hipMalloc(&c->tw4096, 2048 * 4));
    hip_check(hipMalloc(&c->tw4096_inv, 2048 * 4));
    hip_check(hipMemcpy(c->tw4096, fwd.data(), 2048 * 4, hipMemcpyHostToDevice));
    hip_check(hipMemcpy(c->tw4096_inv, inv.data(), 2048 * 4, hipMemcpyHostToDevice));
    c->reserve_pinned(size_t(4) << 20);
    return c.release();
}

// ---------------------------------------------------------------------------
// FRI over a device-resident bit-reversed layer 0.

TreeGeom tree_geom(uint64_t leaves, uint64_t cap_height) {
    TreeGeom g;
    g.leaves = leaves;
    g.depth = int(ilog2(leaves));
    g.drop = int(std::min<uint64_t>(cap_height, uint64_t(g.depth)));
    g.target = g.depth - g.drop;
    return g;
}

// Build the parent levels of a tree whose levels [0, have] exist.
void build_levels(Ctx& ctx, uint32_t* tree, const TreeGeom& g, int have) {
    Launch L = ctx.launch();
    int cur = have;
    while (cur < g.target) {
        int k = std::min(9, g.target - cur);
        launch_merkle_levels(L, tree + g.level_offset_words(cur), g.leaves >> cur, k);
        cur += k;
    }
}

namespace {

using Digest = std::array<uint8_t, 32>;

void digest_words_to_bytes(const uint32_t* w, uint8_t* out) {
    for (int k = 0; k < 8; ++k) store_le32(out + 4 * k, w[k]);
}

// Host Merkle cap for the (small) final layer: commit_layer (fri.rs:1401-1428).
std::vector<Digest> host_layer_cap(const std::vector<E4>& vals, uint64_t cap_height) {
    std::vector<Digest> level;
    if (vals.size() % 4 == 0) {
        for (size_t i = 0; i < vals.size(); i += 4) {
            uint32_t m[16];
            for (int k = 0; k < 4; ++k)
                for (int c = 0; c < 4; ++c) m[4 * k + c] = vals[i + k].v[c];
            uint32_t d[8];
            leaf64_words(m, d);
            Digest dg;
            digest_words_to_bytes(d, dg.data());
            level.push_back(dg);
        }
    } else {
        // one 0x00-prefixed 16-byte leaf per value (merkle.rs:62-79, hash.rs:352-382)
        for (const E4& v : vals) {
            uint8_t buf[17];
            buf[0] = 0;
            for (int c = 0; c < 4; ++c) store_le32(buf + 1 + 4 * c, v.v[c]);
            Digest dg;
            blake3_hash(buf, 17, dg.data());
            level.push_back(dg);
        }
    }
    size_t target = 1;
    while (target < level.size()) target <<= 1;
    while (level.size() < target) level.push_back(level.back());
    int depth = int(ilog2(level.size()));
    int drop = int(std::min<uint64_t>(cap_height, uint64_t(depth)));
    for (int l = 0; l < depth - drop; ++l) {
        std::vector<Digest> next(level.size() / 2);
        for (size_t i = 0; i < next.size(); ++i) {
            uint32_t a[8], b[8], o[8];
            for (int k = 0; k < 8; ++k) {
                a[k] = load_le32(level[2 * i].data() + 4 * k);
                b[k] = load_le32(level[2 * i + 1].data() + 4 * k);
            }
            node_words(a, b, o);
            digest_words_to_bytes(o, next[i].data());
        }
        level.swap(next);
    }
    return level;
}

struct FriLayerDev {
    uint32_t* values = nullptr;  // Ext4 (4 words) or base (1 word) per value
    bool base = false;
    uint64_t len = 0;
    uint32_t* tree = nullptr;
    TreeGeom geom;
};

struct FriResult {
    std::vector<FriLayerDev> layers;  // opened layers (all but the final)
    std::vector<std::vector<Digest>> caps;
    std::vector<E4> final_values;
    uint64_t nonce = 0;
    std::vector<uint64_t> queries;
};

uint64_t fri_arena_words(uint64_t m, bool base0, uint64_t final_size, uint64_t cap_height) {
    uint64_t w = 0;
    for (uint64_t len = m; len > final_size; len /= 4) {
        bool base = base0 && len == m;
        if (len != m) w += len * 4;
        (void)base;
        w += tree_geom(len / 4, cap_height).words();
        w += 64;  // alignment slack
    }
    w += final_size * 4 + 64;
    return w;
}

// fri.rs:551-633 from layer 0 already on the device, then fri.rs:645-656
// (grind and query sampling). Leaves tree data resident for openings.
FriResult fri_commit_and_sample(Ctx& ctx, uint32_t* layer0, bool base0, uint64_t m, uint64_t poly_len,
                                uint64_t degree_bound, const z0p_fri_params& fp, PowTable winv, Transcript& tr) {
    FriResult R;
    Launch L = ctx.launch();
    uint8_t dims[24];
    store_le64(dims, m);
    store_le64(dims + 8, poly_len);
    store_le64(dims + 16, degree_bound);
    tr.absorb("fri-dims", dims, 24);
    const uint32_t unit_m = to_mont(root_of_unity(2));
    const uint32_t ginv = inv_c(GENERATOR);

    uint32_t* cur = layer0;
    bool cur_base = base0;
    uint64_t len = m;
    int depth = 0;
    while (true) {
        if (len == fp.final_size) {
            // Final layer: values to the host, cap computed there.
            size_t words = cur_base ? len : len * 4;
            hip_check(hipMemcpyAsync(ctx.pinned, cur, words * 4, hipMemcpyDeviceToHost, ctx.stream));
            ctx.sync();
            const uint32_t* w = reinterpret_cast<const uint32_t*>(ctx.pinned);
            R.final_values.resize(len);
            for (uint64_t i = 0; i < len; ++i)
                R.final_values[i] = cur_base ? e4_base(w[i]) : E4{{w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]}};
            auto cap = host_layer_cap(R.final_values, fp.cap_height);
            for (auto& d : cap) tr.absorb("fri-layer-root", d.data(), 32);
            R.caps.push_back(cap);
            break;
        }
        FriLayerDev lay;
        lay.values = cur;
        lay.base = cur_base;
        lay.len = len;
        lay.geom = tree_geom(len / 4, fp.cap_height);
        lay.tree = ctx.alloc<uint32_t>(lay.geom.words());
        int first_levels = std::min(9, lay.geom.target);
        launch_fri_leaves(L, cur, cur_base, lay.tree, lay.geom.leaves, first_levels);
        build_levels(ctx, lay.tree, lay.geom, first_levels);
        uint64_t cap_n = uint64_t(1) << lay.geom.drop;
        hip_check(hipMemcpyAsync(ctx.pinned, lay.tree + lay.geom.level_offset_words(lay.geom.target), cap_n * 32,
                                 hipMemcpyDeviceToHost, ctx.stream));
        ctx.sync();
        std::vector<Digest> cap(cap_n);
        for (uint64_t i = 0; i < cap_n; ++i) {
            std::memcpy(cap[i].data(), ctx.pinned + 32 * i, 32);
            tr.absorb("fri-layer-root", cap[i].data(), 32);
        }
        R.caps.push_back(cap);
        R.layers.push_back(lay);
        FoldConsts fc{};
        for (int k = 0; k < 3; ++k) {
            E4 ch = tr.challenge_ext("fri-fold");
            for (int c = 0; c < 4; ++c) fc.ch_m[k][c] = to_mont(ch.v[c]);
        }
        fc.unit_m = unit_m;
        fc.ginv_m = to_mont(pow_c(ginv, uint64_t(1) << (2 * depth)));
        fc.log_len = int(ilog2(len));
        fc.depth = depth;
        uint32_t* next = ctx.alloc<uint32_t>((len / 4) * 4);
        launch_fold(L, cur, cur_base, next, winv, fc);
        cur = next;
        cur_base = false;
        len /= 4;
        ++depth;
    }
    std::vector<uint8_t> fin(R.final_values.size() * 16);
    for (size_t i = 0; i < R.final_values.size(); ++i)
        for (int c = 0; c < 4; ++c) store_le32(fin.data() + 16 * i + 4 * c, R.final_values[i].v[c]);
    tr.absorb("fri-final", fin);

    // Proof-of-work (transcript.rs:206-228) on the device.
    if (fp.grinding_bits > 0) {
        std::vector<uint8_t> head = tr.pow_head("fri-pow");
        require(head.size() == 47, Z0P_ERR_DEVICE);
        uint32_t hw[16] = {0};
        for (size_t i = 0; i < head.size(); ++i) hw[i >> 2] |= uint32_t(head[i]) << (8 * (i & 3));
        unsigned long long* res = ctx.alloc<unsigned long long>(1);
        uint64_t found = ~uint64_t(0);
        for (uint64_t window = 0; found == ~uint64_t(0); ++window) {
            // Each launch searches [0, limit); later windows only widen the
            // bound, and the kernel returns the smallest passing nonce below it.
            uint64_t limit = uint64_t(1) << std::min<uint64_t>(63, 24 + 4 * window);
            uint64_t init = ~uint64_t(0);
            std::memcpy(ctx.pinned, &init, 8);
            hip_check(hipMemcpyAsync(res, ctx.pinned, 8, hipMemcpyHostToDevice, ctx.stream));
            launch_grind(L, hw, 47, fp.grinding_bits, limit, res);
            hip_check(hipMemcpyAsync(ctx.pinned + 8, res, 8, hipMemcpyDeviceToHost, ctx.stream));
            ctx.sync();
            std::memcpy(&found, ctx.pinned + 8, 8);
        }
        R.nonce = found;
        uint8_t nb[8];
        store_le64(nb, found);
        tr.absorb("fri-pow", nb, 8);
    }
    for (uint64_t q = 0; q < fp.num_queries; ++q) R.queries.push_back(tr.challenge_index("fri-query", m));
    return R;
}

// Opening gather plan: (src word offset from arena base, words, dst word).
struct Gather {
    std::vector<uint64_t> src;
    std::vector<uint32_t> dst, len;
    uint32_t total = 0;
    uint32_t add(const Ctx& ctx, const uint32_t* p, uint32_t words) {
        src.push_back(uint64_t(reinterpret_cast<const uint8_t*>(p) - ctx.arena) / 4);
        dst.push_back(total);
        len.push_back(words);
        uint32_t at = total;
        total += words;
        return at;
    }
};

struct FriOpenPlan {
    // per query, per opened layer: offset of the 4 values and of each path step
    std::vector<std::vector<uint32_t>> values_at;
    std::vector<std::vector<std::vector<uint32_t>>> path_at;
};

FriOpenPlan plan_fri_openings(Ctx& ctx, const FriResult& R, Gather& g) {
    FriOpenPlan P;
    for (uint64_t leaf : R.queries) {
        std::vector<uint32_t> vals;
        std::vector<std::vector<uint32_t>> paths;
        uint64_t index = leaf;
        for (const FriLayerDev& lay : R.layers) {
            uint64_t base = index / 4 * 4;
            vals.push_back(lay.base ? g.add(ctx, lay.values + base, 4) : g.add(ctx, lay.values + base * 4, 16));
            std::vector<uint32_t> steps;
            uint64_t qi = base / 4;
            for (int l = 0; l < lay.geom.target; ++l) {
                uint64_t sib = (qi >> l) ^ 1;
                steps.push_back(g.add(ctx, lay.tree + lay.geom.level_offset_words(l) + sib * 8, 8));
            }
            paths.push_back(steps);
            index /= 4;
        }
        P.values_at.push_back(vals);
        P.path_at.push_back(paths);
    }
    return P;
}

void run_gather(Ctx& ctx, const Gather& g, std::vector<uint32_t>& out) {
    const uint32_t items = uint32_t(g.src.size());
    size_t desc_bytes = size_t(items) * 16;
    size_t out_bytes = size_t(g.total) * 4;
    ctx.reserve_pinned(desc_bytes + out_bytes + 4096);
    uint64_t* d_src = ctx.alloc<uint64_t>(items);
    uint32_t* d_dst = ctx.alloc<uint32_t>(items);
    uint32_t* d_len = ctx.alloc<uint32_t>(items);
    uint32_t* d_out = ctx.alloc<uint32_t>(g.total);
    uint8_t* hp = ctx.pinned;
    std::memcpy(hp, g.src.data(), size_t(items) * 8);
    std::memcpy(hp + size_t(items) * 8, g.dst.data(), size_t(items) * 4);
    std::memcpy(hp + size_t(items) * 12, g.len.data(), size_t(items) * 4);
    hip_check(hipMemcpyAsync(d_src, hp, size_t(items) * 8, hipMemcpyHostToDevice, ctx.stream));
    hip_check(hipMemcpyAsync(d_dst, hp + size_t(items) * 8, size_t(items) * 4, hipMemcpyHostToDevice, ctx.stream));
    hip_check(hipMemcpyAsync(d_len, hp + size_t(items) * 12, size_t(items) * 4, hipMemcpyHostToDevice, ctx.stream));
    launch_gather(ctx.launch(), reinterpret_cast<const uint32_t*>(ctx.arena), d_src, d_dst, d_len, items, d_out);
    hip_check(hipMemcpyAsync(hp + desc_bytes, d_out, out_bytes, hipMemcpyDeviceToHost, ctx.stream));
    ctx.sync();
    out.resize(g.total);
    std::memcpy(out.data(), hp + desc_bytes, out_bytes);
}

void write_cap(Writer& w, const std::vector<Digest>& cap) {
    w.u32(uint32_t(cap.size()));
    for (auto& d : cap) w.raw(d.data(), 32);
}

void write_path(Writer& w, const std::vector<uint32_t>& at, const std::vector<uint32_t>& words, uint64_t index) {
    w.u32(uint32_t(at.size()));
    for (size_t l = 0; l < at.size(); ++l) {
        uint8_t d[32];
        digest_words_to_bytes(&words[at[l]], d);
        w.raw(d, 32);
        w.byte(((index >> l) & 1) == 0 ? 1 : 0);
    }
}

// fri::Commitment::write_body (fri.rs:896-904).
void write_fri_commitment_body(Writer& w, const FriResult& R, uint64_t m, uint64_t poly_len, uint64_t degree_bound) {
    w.u64(m);
    w.u64(poly_len);
    w.u64(degree_bound);
    w.u32(uint32_t(R.caps.size()));
    for (auto& c : R.caps) write_cap(w, c);
}

// fri::Proof::write_body (fri.rs:352-368).
void write_fri_proof_body(Writer& w, const FriResult& R, const FriOpenPlan& P, const std::vector<uint32_t>& words) {
    w.u32(uint32_t(R.final_values.size()));
    for (auto& v : R.final_values) w.e4(v);
    w.u64(R.nonce);
    w.u32(uint32_t(R.queries.size()));
    for (size_t q = 0; q < R.queries.size(); ++q) {
        w.u32(uint32_t(R.layers.size()));
        uint64_t index = R.queries[q];
        for (size_t d = 0; d < R.layers.size(); ++d) {
            const uint32_t* v = &words[P.values_at[q][d]];
            if (R.layers[d].base) {
                for (int k = 0; k < 4; ++k) w.e4(e4_base(v[k]));
            } else {
                for (int k = 0; k < 16; ++k) w.u32(v[k]);
            }
            write_path(w, P.path_at[q][d], words, index / 4);
            index /= 4;
        }
    }
}

}  // namespace


// ---------------------------------------------------------------------------
// air::prove (air.rs:365-372 -> 408-422 -> 485-599), linked path.
std::vector<uint8_t> air_prove(Ctx& ctx, const z0p_air_params& p, Transcript& tr) {
    validate_air(p);
    const uint64_t n = p.n_steps, B = p.blowup, m = n * B;
    const int logn = int(ilog2(n)), logb = int(ilog2(B)), logm = logn + logb;
    const uint32_t a0 = p.a0 % P, b0 = p.b0 % P, b_last = p.b_last % P;
    const uint32_t actual_b_last = fib_b_last(a0, b0, n);

    // absorb_params (air.rs:754-766): 56 bytes, grinding_bits as u32.
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

    const z0p_fri_params fp = air_fri_params(p);
    validate_fri(fp);
    fri_num_folds(m, fp.final_size, n - 1);

    const GeometryTables& gt = ctx.geometry(logn, logb);
    const TreeGeom tg = tree_geom(m / 8, p.cap_height);
    const uint64_t words = 2 * n + std::max(n, m) + 2 * m + tg.words() + 4 * m +
                           fri_arena_words(m, false, fp.final_size, fp.cap_height) + (1u << 16);
    ctx.reserve_arena(words * 4 + (size_t(p.num_queries) * 4096 * 16) + (16u << 20));
    ctx.reset_arena();
    Launch L = ctx.launch();

    uint32_t* col = ctx.alloc<uint32_t>(n);
    uint32_t* coeff = ctx.alloc<uint32_t>(n);
    uint32_t* scratch = ctx.alloc<uint32_t>(std::max(n, m));
    uint32_t* A = ctx.alloc<uint32_t>(m);
    uint32_t* Bc = ctx.alloc<uint32_t>(m);
    uint32_t* trace_tree = ctx.alloc<uint32_t>(tg.words());
    uint32_t* consts = ctx.alloc<uint32_t>(B);

    // Linked LDE constant (air.rs:815-820): coeff = (b_last - a0) w_last / n,
    // times (x^n - 1) = (s_j^n - 1) per coset.
    const uint32_t w_n = root_of_unity(uint32_t(logn));
    const uint32_t w_last = pow_c(w_n, n - 1);
    const uint32_t coeff_c = mul(mul(sub(actual_b_last, a0), w_last), inv_c(uint32_t(n % P)));
    {
        const uint32_t wm = root_of_unity(uint32_t(logm));
        uint32_t* hc = reinterpret_cast<uint32_t*>(ctx.pinned);
        for (uint64_t j = 0; j < B; ++j) {
            uint32_t s = mul(GENERATOR, pow_c(wm, j));
            hc[j] = mul(coeff_c, sub(pow_c(s, n), 1));
        }
        hip_check(hipMemcpyAsync(consts, hc, B * 4, hipMemcpyHostToDevice, ctx.stream));
    }

    launch_fib_trace(L, col, n, a0, b0);
    launch_intt_to_br(L, col, coeff, scratch, logn, ctx.tw4096_inv, gt.winv, logb, to_mont(inv_c(uint32_t(n % P))));
    launch_lde_br(L, coeff, A, scratch, logn, logb, ctx.tw4096, gt.w, gt.coset);
    launch_linked_b(L, A, Bc, logn, logb, gt.w, gt.coset.shift, consts, to_mont(w_last));
    int have = launch_trace_leaves(L, A, Bc, trace_tree, logn, logb, tg.target);
    build_levels(ctx, trace_tree, tg, have);

    const uint64_t cap_n = uint64_t(1) << tg.drop;
    hip_check(hipMemcpyAsync(ctx.pinned + 4096, trace_tree + tg.level_offset_words(tg.target), cap_n * 32,
                             hipMemcpyDeviceToHost, ctx.stream));
    ctx.sync();
    std::vector<Digest> column_cap(cap_n);
    for (uint64_t i = 0; i < cap_n; ++i) std::memcpy(column_cap[i].data(), ctx.pinned + 4096 + 32 * i, 32);
    {
        std::vector<uint8_t> cb(cap_n * 32);
        for (uint64_t i = 0; i < cap_n; ++i) std::memcpy(cb.data() + 32 * i, column_cap[i].data(), 32);
        tr.absorb("fib/trace", cb);
    }
    ComposeConsts cc{};
    for (int k = 0; k < 5; ++k) {
        E4 al = tr.challenge_ext("fib/alpha");
        for (int c = 0; c < 4; ++c) cc.alpha_m[k][c] = to_mont(al.v[c]);
    }
    cc.w_last_m = to_mont(w_last);
    cc.a0 = a0;
    cc.b0 = b0;
    cc.b_last = b_last;
    uint32_t* layer0 = ctx.alloc<uint32_t>(4 * m);
    launch_compose(L, A, Bc, layer0, logn, logb, gt.w, gt.coset.shift, gt.zinv, cc);

    FriResult R = fri_commit_and_sample(ctx, layer0, false, m, m, n - 1, fp, gt.winv, tr);

    // Openings: FRI quads/paths plus the trace rows (air.rs:567-591).
    Gather g;
    FriOpenPlan P = plan_fri_openings(ctx, R, g);
    struct TraceOpen {
        uint64_t index, shifted;
        uint32_t leaf_at, shift_leaf_at;
        std::vector<uint32_t> path, shift_path;
    };
    std::vector<TraceOpen> tops;
    auto row_pos = [&](uint64_t i) -> uint64_t {
        uint64_t j = i & (B - 1), t = i >> logb;
        return (uint64_t(brev_host(j, logb)) << logn) | brev_host(t, logn);
    };
    auto add_leaf = [&](uint64_t base) {
        uint32_t first = 0;
        for (int r = 0; r < 8; ++r) {
            uint64_t pos = row_pos(base + r);
            uint32_t at = g.add(ctx, A + pos, 1);
            g.add(ctx, Bc + pos, 1);
            if (r == 0) first = at;
        }
        return first;
    };
    auto add_path = [&](uint64_t leaf) {
        std::vector<uint32_t> at;
        for (int l = 0; l < tg.target; ++l)
            at.push_back(g.add(ctx, trace_tree + tg.level_offset_words(l) + ((leaf >> l) ^ 1) * 8, 8));
        return at;
    };
    for (uint64_t leaf : R.queries) {
        TraceOpen o;
        o.index = brev_host(leaf, logm);
        o.shifted = (o.index + B) % m;
        uint64_t base = o.index & ~uint64_t(7), sbase = o.shifted & ~uint64_t(7);
        o.leaf_at = add_leaf(base);
        o.shift_leaf_at = add_leaf(sbase);
        o.path = add_path(base / 8);
        o.shift_path = add_path(sbase / 8);
        tops.push_back(std::move(o));
    }
    std::vector<uint32_t> words_out;
    run_gather(ctx, g, words_out);

    // Proof::to_bytes (air.rs:262-281).
    Writer w;
    w.byte(WIRE_VERSION);
    write_cap(w, column_cap);
    write_fri_commitment_body(w, R, m, m, n - 1);
    write_fri_proof_body(w, R, P, words_out);
    w.u32(uint32_t(tops.size()));
    for (const TraceOpen& o : tops) {
        const uint32_t* lw = &words_out[o.leaf_at];
        const uint32_t* sw = &words_out[o.shift_leaf_at];
        w.u64(o.index);
        w.u32(lw[2 * (o.index & 7)]);
        w.u32(lw[2 * (o.index & 7) + 1]);
        w.u32(sw[2 * (o.shifted & 7)]);
        w.u32(sw[2 * (o.shifted & 7) + 1]);
        write_path(w, o.path, words_out, o.index / 8);
        write_path(w, o.shift_path, words_out, o.shifted / 8);
        for (int k = 0; k < 16; ++k) w.u32(lw[k]);
        for (int k = 0; k < 16; ++k) w.u32(sw[k]);
    }
    return std::move(w.out);
}

// fri::commit + public_commitment + open (fri.rs:408-433, 863-870, 645-666).
void fri_commit_open(Ctx& ctx, const uint32_t* coeffs, uint64_t n, const z0p_fri_params& fp, Transcript& tr,
                     std::vector<uint8_t>& commitment, std::vector<uint8_t>& proof) {
    validate_fri(fp);
    require(n != 0, Z0P_ERR_EMPTY_INPUT);
    require(pow2(n) && n <= MAX_DOMAIN, Z0P_ERR_INVALID_DOMAIN_SIZE);
    require(n <= MAX_DOMAIN / fp.blowup, Z0P_ERR_INVALID_DOMAIN_SIZE);
    const uint64_t B = fp.blowup, m = n * B;
    const int logn = int(ilog2(n)), logb = int(ilog2(B));
    fri_num_folds(m, fp.final_size, n - 1);
    const GeometryTables& gt = ctx.geometry(logn, logb);
    const uint64_t words = 2 * n + m + m + fri_arena_words(m, true, fp.final_size, fp.cap_height) + (1u << 16);
    ctx.reserve_arena(words * 4 + (size_t(fp.num_queries) * 4096 * 16) + (16u << 20));
    ctx.reset_arena();
    Launch L = ctx.launch();
    uint32_t* raw = ctx.alloc<uint32_t>(n);
    uint32_t* coeff_br = ctx.alloc<uint32_t>(n);
    uint32_t* scratch = ctx.alloc<uint32_t>(m);
    uint32_t* layer0 = ctx.alloc<uint32_t>(m);
    // Coefficients are reduced mod p on entry (Field::new semantics).
    ctx.reserve_pinned(size_t(n) * 4 + 4096);
    uint32_t* hc = reinterpret_cast<uint32_t*>(ctx.pinned);
    for (uint64_t i = 0; i < n; ++i) hc[i] = coeffs[i] % P;
    hip_check(hipMemcpyAsync(raw, hc, n * 4, hipMemcpyHostToDevice, ctx.stream));
    launch_bitrev(L, raw, coeff_br, logn);
    launch_lde_br(L, coeff_br, layer0, scratch, logn, logb, ctx.tw4096, gt.w, gt.coset);
    ctx.sync();  // pinned staging is reused below

    FriResult R = fri_commit_and_sample(ctx, layer0, true, m, n, n - 1, fp, gt.winv, tr);
    Gather g;
    FriOpenPlan Pl = plan_fri_openings(ctx, R, g);
    std::vector<uint32_t> words_out;
    run_gather(ctx, g, words_out);

    Writer wc;
    wc.byte(WIRE_VERSION);
    write_fri_commitment_body(wc, R, m, n, n - 1);
    commitment = std::move(wc.out);
    Writer wp;
    wp.byte(WIRE_VERSION);
    write_fri_proof_body(wp, R, Pl, words_out);
    proof = std::move(wp.out);
}

}  // namespace z0p
