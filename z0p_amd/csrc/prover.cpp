// Host orchestration of the GPU prover: Fibonacci AIR (air.rs:365-599), the
// range-check AIR (range.rs:334-508) and the FRI commitment (fri.rs:408-666).
//
// Every proof is a small state machine whose phases are cut by the
// Fiat-Shamir transcript: the trace cap gates the composition challenges
// (air.rs:525-540), each FRI layer cap gates that layer's fold challenges
// (fri.rs:594-612), and the proof-of-work gates the query indices
// (fri.rs:645-656). One scheduler loop advances every admitted proof by one
// phase per iteration. It launches each kernel type once for all proofs
// (descriptor tables), gathers the few words the transcript needs (caps,
// final layers, nonces, openings) with one collect kernel and one
// device-to-host copy, synchronises once, then runs the host transcript
// steps. Bulk data never leaves device memory. Proofs are admitted as pool
// memory frees up, so proofs of different sizes and phases share every
// launch; a single proof is a batch of one.
#include "prover.hpp"

#include <algorithm>
#include <array>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <variant>

#include "blake3.hpp"
#include "params.hpp"

namespace z0p {

// The field modulus under a name that the proof variables (P) do not shadow.
constexpr uint32_t P_MOD = P;

uint32_t brev_host(uint64_t x, int bits) {
    return bits == 0 ? 0u : uint32_t(__builtin_bitreverse64(x) >> (64 - bits));
}

namespace {

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

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace

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
    std::vector<uint32_t> shift(B), zinv(B), zvan(B);
    for (uint64_t j = 0; j < B; ++j) {
        uint32_t s = mul(GENERATOR, pow_c(wm, j));
        shift[j] = to_mont(s);
        uint32_t v = sub(pow_c(s, n), 1);
        zvan[j] = to_mont(v);
        zinv[j] = to_mont(inv_c(v));
    }
    t->coset.shift = upload(t->owned, shift);
    t->zinv = upload(t->owned, zinv);
    t->zvan = upload(t->owned, zvan);
    {
        // Point tables in bit-reversed order, so consecutive positions load
        // consecutive words; generated on the device from the power tables.
        void* xn = nullptr;
        void* finv = nullptr;
        const uint64_t quads = std::max<uint64_t>(m / 4, 1);
        hip_check(hipMalloc(&xn, n * 4));
        t->owned.push_back(xn);
        hip_check(hipMalloc(&finv, quads * 4));
        t->owned.push_back(finv);
        Launch L = launch();
        launch_pow_brev(L, t->w, static_cast<uint32_t*>(xn), logn, logb);
        launch_pow_brev(L, t->winv, static_cast<uint32_t*>(finv), int(ilog2(quads)), 0);
        hip_check(hipStreamSynchronize(stream));
        t->xn = static_cast<uint32_t*>(xn);
        t->finv = static_cast<uint32_t*>(finv);
    }
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

const PointInverses& Ctx::point_inverses(int logn, int logb) {
    const GeometryTables& g = geometry(logn, logb);
    GeometryTables& t = *tables.at(std::make_pair(logn, logb));
    if (t.pinv.i1) return t.pinv;
    const uint64_t m = uint64_t(1) << (logn + logb);
    void *i1 = nullptr, *iw = nullptr, *iz = nullptr;
    hip_check(hipMalloc(&i1, m * 4));
    t.owned.push_back(i1);
    hip_check(hipMalloc(&iw, m * 4));
    t.owned.push_back(iw);
    hip_check(hipMalloc(&iz, m * 4));
    t.owned.push_back(iz);
    PointInvJob j{};
    j.i1 = static_cast<uint32_t*>(i1);
    j.iw = static_cast<uint32_t*>(iw);
    j.iz = static_cast<uint32_t*>(iz);
    j.xn = g.xn;
    j.shift_m = g.coset.shift;
    j.zinv_m = g.zinv;
    j.w_last_m = to_mont(pow_c(root_of_unity(uint32_t(logn)), (uint64_t(1) << logn) - 1));
    j.logn = logn;
    j.logb = logb;
    launch_point_inv(launch(), j);
    hip_check(hipStreamSynchronize(stream));
    t.pinv = PointInverses{j.i1, j.iw, j.iz};
    return t.pinv;
}

Ctx::~Ctx() {
    tables.clear();
    if (alt_stream) (void)hipStreamSynchronize(alt_stream);
    if (stream) (void)hipStreamSynchronize(stream);
    release_graphs();
    for (hipEvent_t& e : verify_done)
        if (e) (void)hipEventDestroy(e);
    if (stage_host) (void)hipHostFree(stage_host);
    if (alt_stage_host) (void)hipHostFree(alt_stage_host);
    if (flag_host) (void)hipHostFree(flag_host);
    if (alt_flag_host) (void)hipHostFree(alt_flag_host);
    if (setup_done) (void)hipEventDestroy(setup_done);
    if (alt_setup_done) (void)hipEventDestroy(alt_setup_done);
    if (mailbox_host) (void)hipHostFree(mailbox_host);
    if (scratch) (void)hipFree(scratch);
    if (grind_buf) (void)hipFree(grind_buf);
    for (auto& kv : spare) (void)hipFree(kv.second);
    if (tw4096) (void)hipFree(tw4096);
    if (tw4096_inv) (void)hipFree(tw4096_inv);
    if (alt_stream) (void)hipStreamDestroy(alt_stream);
    if (stream) (void)hipStreamDestroy(stream);
}

Ctx* ctx_create(int device) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0 || device < 0 || device >= count)
        throw Status(Z0P_ERR_NO_DEVICE);
    hip_check(hipSetDevice(device));
    noalloc_scratch_init(device);
    auto c = std::make_unique<Ctx>();
    c->device = device;
    hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    hip_check(hipStreamCreateWithFlags(&c->alt_stream, hipStreamNonBlocking));
    hip_check(hipEventCreateWithFlags(&c->setup_done, hipEventDisableTiming));
    hip_check(hipEventCreateWithFlags(&c->alt_setup_done, hipEventDisableTiming));
    hip_check(hipEventRecord(c->setup_done, c->stream));
    hip_check(hipEventRecord(c->alt_setup_done, c->alt_stream));
    const uint32_t w = root_of_unity(12), wi = inv_c(w);
    auto fwd = geometric_m(w, 2048), inv = geometric_m(wi, 2048);
    hip_check(hipMalloc(&c->tw4096, 2048 * 4));
    hip_check(hipMalloc(&c->tw4096_inv, 2048 * 4));
    hip_check(hipMemcpy(c->tw4096, fwd.data(), 2048 * 4, hipMemcpyHostToDevice));
    hip_check(hipMemcpy(c->tw4096_inv, inv.data(), 2048 * 4, hipMemcpyHostToDevice));
    c->ring.init(size_t(32) << 20);
    if (const char* e = std::getenv("Z0P_PIPELINE")) c->pipeline_single = std::strcmp(e, "0") != 0;
    if (const char* e = std::getenv("Z0P_WAVES")) c->waves_min = size_t(std::strtoull(e, nullptr, 10));
    c->reserve_stage(size_t(1) << 16);
    hip_check(hipHostMalloc(reinterpret_cast<void**>(&c->flag_host), 256, hipHostMallocMapped | hipHostMallocCoherent));
    hip_check(hipHostGetDevicePointer(reinterpret_cast<void**>(&c->flag_dev), c->flag_host, 0));
    *c->flag_host = 0;
    hip_check(hipHostMalloc(reinterpret_cast<void**>(&c->alt_flag_host), 256,
                            hipHostMallocMapped | hipHostMallocCoherent));
    hip_check(hipHostGetDevicePointer(reinterpret_cast<void**>(&c->alt_flag_dev), c->alt_flag_host, 0));
    *c->alt_flag_host = 0;
    const size_t mb = Ctx::MAILBOX_HEAD + Ctx::MAILBOX_SLOTS * Ctx::MAILBOX_SLOT;
    hip_check(hipHostMalloc(reinterpret_cast<void**>(&c->mailbox_host), mb, hipHostMallocMapped | hipHostMallocCoherent));
    hip_check(hipHostGetDevicePointer(reinterpret_cast<void**>(&c->mailbox_dev), c->mailbox_host, 0));
    std::memset(c->mailbox_host, 0, mb);
    return c.release();
}

// Binet form of the Fibonacci AIR columns. p = 1 mod 5, so 5 is a square and
// x^2 = x + 1 has the roots phi, psi = (1 +- sqrt 5)/2 in the field; with
// F(k) = (phi^k - psi^k)/sqrt 5 (F(-1) = 1, as phi psi = -1),
// A[i] = F(i-1) a0 + F(i) b0 = c1 phi^i + c2 psi^i for
// c1 = (a0/phi + b0)/sqrt 5 and c2 = -(a0/psi + b0)/sqrt 5.
struct FibRoots {
    uint32_t phi = 0, psi = 0, inv_s5 = 0;
    FibRoots() {
        // Tonelli-Shanks square root of 5: p - 1 = 2^27 * 15.
        const uint32_t q = (P - 1) >> TWO_ADICITY;
        uint32_t z = 2;
        while (pow_c(z, (P - 1) / 2) != P - 1) ++z;
        uint32_t m = TWO_ADICITY, c = pow_c(z, q), t = pow_c(5, q), r = pow_c(5, (q + 1) / 2);
        while (t != 1) {
            uint32_t i = 0, t2 = t;
            while (t2 != 1) t2 = mul(t2, t2), ++i;
            const uint32_t b = pow_c(c, uint64_t(1) << (m - i - 1));
            m = i, c = mul(b, b), t = mul(t, c), r = mul(r, b);
        }
        const uint32_t half = inv_c(2);
        phi = mul(add(1, r), half);
        psi = mul(sub(1, r), half);
        inv_s5 = inv_c(r);
    }
};

// The closed-form coefficient job for the column c1 phi^i + c2 psi^i, i < n.
FibCoeffJob fib_coeff_job(uint32_t* out, const GeometryTables& gt, int logn, int logb, uint32_t c1, uint32_t c2) {
    static const FibRoots F;
    const uint64_t n = uint64_t(1) << logn;
    const uint32_t phin = pow_c(F.phi, n), psin = pow_c(F.psi, n);
    // phi and psi are no 2^k-th roots of unity (phi^(2^27) != 1), so no
    // denominator 1 - phi omega^-k vanishes for n <= 2^27.
    require(phin != 1 && psin != 1, Z0P_ERR_DEVICE);
    const uint32_t ninv = inv_c(uint32_t(n % P));
    FibCoeffJob j{};
    j.out = out;
    j.winv = gt.winv;
    j.phi_m = to_mont(F.phi);
    j.psi_m = to_mont(F.psi);
    j.k1_m = to_mont(mul(mul(c1, sub(1, phin)), ninv));
    j.k2_m = to_mont(mul(mul(c2, sub(1, psin)), ninv));
    j.logn = logn;
    j.logb = logb;
    return j;
}

// (c1, c2) of column A for claims (a0, b0); column B[i] = A[i+1] is
// (c1 phi, c2 psi).
std::pair<uint32_t, uint32_t> fib_binet(uint32_t a0, uint32_t b0) {
    static const FibRoots F;
    const uint32_t c1 = mul(add(mul(a0, inv_c(F.phi)), b0), F.inv_s5);
    const uint32_t c2 = sub(0, mul(add(mul(a0, inv_c(F.psi)), b0), F.inv_s5));
    return {c1, c2};
}
std::pair<uint32_t, uint32_t> fib_binet_next(std::pair<uint32_t, uint32_t> c) {
    static const FibRoots F;
    return {mul(c.first, F.phi), mul(c.second, F.psi)};
}

TreeGeom tree_geom(uint64_t leaves, uint64_t cap_height) {
    TreeGeom g;
    g.leaves = leaves;
    g.depth = int(ilog2(leaves));
    g.drop = int(std::min<uint64_t>(cap_height, uint64_t(g.depth)));
    g.target = g.depth - g.drop;
    return g;
}

// Levels a leaf (or node) workgroup reduces in LDS above the level holding
// its parents: its share of that level is min(256, parents) digests, and the
// reduction stops at the cap level (`remaining` levels left) and at the
// schedule's limit.
int subtree_levels_for(uint64_t parents, int remaining, const Schedule& s) {
    int k = 0;
    while (k < SUBTREE_MAX_LEVELS && (uint64_t(2) << k) <= std::min<uint64_t>(parents, 256)) ++k;
    return std::max(0, std::min({k, remaining, s.subtree_levels}));
}

void tree_levels(std::vector<Step>& chain, uint32_t* tree, const TreeGeom& g, int have, const Schedule& s,
                 uint32_t* counters) {
    while (have < g.target) {
        const uint64_t count = g.leaves >> have;
        const int remaining = g.target - have;
        uint32_t* lvl = tree + g.level_offset_words(have);
        if (count > (uint64_t(1) << s.tail_max_log) && s.subtree_levels > 0 && count >= 8) {
            // Wide levels in throughput mode: each workgroup turns 1024
            // digests into two levels of nodes, then reduces its 256 nodes of
            // the second level further in LDS.
            LeavesJob nj{};
            nj.kind = LEAF_NODES;
            nj.src = lvl;
            nj.levels = tree + g.level_offset_words(have + 1);
            nj.leaves = count / 2;
            nj.with_parent = remaining >= 2 ? 1 : 0;
            nj.sub_levels = nj.with_parent ? subtree_levels_for(count / 4, remaining - 2, s) : 0;
            chain.push_back(nj);
            have += 1 + nj.with_parent + nj.sub_levels;
        } else if (count > (uint64_t(1) << s.tail_max_log)) {
            const int fuse = std::min(s.levels_fuse, remaining);
            chain.push_back(LevelsJob{lvl, count, fuse, 0});
            have += fuse;
        } else if (counters && s.tree_top) {
            // Every narrow level in one staged launch.
            chain.push_back(TreeTopJob{tree, g.leaves, have, g.target, counters, 0});
            have = g.target;
        } else {
            const int gl = s.tail_group_log(count);
            const int k = std::min(gl, remaining);
            chain.push_back(TailJob{lvl, count, k, gl, 0});
            have += k;
        }
    }
}

void tree_steps(std::vector<Step>& chain, uint32_t* tree, const TreeGeom& g, LeavesJob leaves, const Schedule& s,
                uint32_t* counters) {
    leaves.levels = tree;
    leaves.leaves = g.leaves;
    leaves.with_parent = g.target >= 1 ? 1 : 0;
    leaves.sub_levels = leaves.with_parent && leaf_contiguous(leaves.kind)
                            ? subtree_levels_for(g.leaves / 2, g.target - 1, s)
                            : 0;
    chain.push_back(leaves);
    tree_levels(chain, tree, g, leaves.with_parent + leaves.sub_levels, s, counters);
}

namespace {

// Per-round buckets: step r of every chain, grouped by kernel type.
struct Buckets {
    std::vector<ZeroJob> zero;
    std::vector<FibCoeffJob> fib;
    std::vector<BitrevJob> bitrev;
    std::vector<InterleaveJob> interleave;
    std::vector<NttJob> ntt;
    std::vector<LinkedJob> linked;
    std::vector<ComposeJob> compose;
    std::vector<RangeComposeJob> rcompose;
    std::vector<TableComposeJob> tcompose;
    std::vector<FoldCombineJob> fcombine;
    std::vector<LeavesJob> leaves;
    std::vector<LevelsJob> levels;
    std::vector<TailJob> tail;
    std::vector<TreeTopJob> top;
    std::vector<FoldJob> fold;
    std::vector<SmallCommitJob> small;
    std::vector<GrindJob> grind;
    std::vector<OpenJob> open;

    std::vector<ZeroJob>& of(const ZeroJob&) { return zero; }
    std::vector<FibCoeffJob>& of(const FibCoeffJob&) { return fib; }
    std::vector<BitrevJob>& of(const BitrevJob&) { return bitrev; }
    std::vector<InterleaveJob>& of(const InterleaveJob&) { return interleave; }
    std::vector<NttJob>& of(const NttJob&) { return ntt; }
    std::vector<LinkedJob>& of(const LinkedJob&) { return linked; }
    std::vector<ComposeJob>& of(const ComposeJob&) { return compose; }
    std::vector<RangeComposeJob>& of(const RangeComposeJob&) { return rcompose; }
    std::vector<TableComposeJob>& of(const TableComposeJob&) { return tcompose; }
    std::vector<FoldCombineJob>& of(const FoldCombineJob&) { return fcombine; }
    std::vector<LeavesJob>& of(const LeavesJob&) { return leaves; }
    std::vector<LevelsJob>& of(const LevelsJob&) { return levels; }
    std::vector<TailJob>& of(const TailJob&) { return tail; }
    std::vector<TreeTopJob>& of(const TreeTopJob&) { return top; }
    std::vector<FoldJob>& of(const FoldJob&) { return fold; }
    std::vector<SmallCommitJob>& of(const SmallCommitJob&) { return small; }
    std::vector<GrindJob>& of(const GrindJob&) { return grind; }
    std::vector<OpenJob>& of(const OpenJob&) { return open; }

    void add(const Step& s) {
        std::visit([this](const auto& j) { of(j).push_back(j); }, s);
    }
    void launch(const Launch& L) const {
        if (!zero.empty()) launch_zero(L, zero);
        if (!fib.empty()) launch_fib_coeffs(L, fib);
        if (!bitrev.empty()) launch_bitrev(L, bitrev);
        if (!ntt.empty()) launch_ntt(L, ntt);
        if (!interleave.empty()) launch_interleave(L, interleave);
        if (!linked.empty()) launch_linked(L, linked);
        if (!compose.empty()) launch_compose(L, compose);
        if (!rcompose.empty()) launch_range_compose(L, rcompose);
        if (!tcompose.empty()) launch_table_compose(L, tcompose);
        if (!fcombine.empty()) launch_fold_combine(L, fcombine);
        if (!leaves.empty()) launch_leaves(L, leaves);
        if (!levels.empty()) launch_levels(L, levels);
        if (!tail.empty()) launch_tail(L, tail);
        if (!top.empty()) launch_tree_top(L, top);
        if (!fold.empty()) launch_fold(L, fold);
        if (!small.empty()) launch_small_commit(L, small);
        if (!grind.empty()) launch_grind(L, grind);
        if (!open.empty()) launch_open(L, open);
    }
};

}  // namespace

void launch_chains(const Launch& L, const std::vector<const std::vector<Step>*>& chains,
                   const std::vector<CollectItem>& collect, uint64_t* flag, uint64_t seq, const HostGate& gate) {
    size_t rounds = 0;
    for (auto* c : chains) rounds = std::max(rounds, c->size());
    for (size_t r = 0; r < rounds; ++r) {
        Buckets b;
        for (auto* c : chains)
            if (r < c->size()) b.add((*c)[r]);
        b.launch(L);
    }
    if (flag || !collect.empty() || gate.go) launch_collect(L, collect, flag, seq, gate);
}

// ---------------------------------------------------------------------------
// Proof state machine.

namespace {

using Digest = std::array<uint8_t, 32>;

void digest_words_to_bytes(const uint32_t* w, uint8_t* out) {
    for (int k = 0; k < 8; ++k) store_le32(out + 4 * k, w[k]);
}

std::vector<Digest> read_digests(const uint32_t* w, uint64_t count) {
    std::vector<Digest> out(count);
    for (uint64_t i = 0; i < count; ++i) digest_words_to_bytes(w + 8 * i, out[i].data());
    return out;
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

void write_cap(Writer& w, const std::vector<Digest>& cap) {
    w.u32(uint32_t(cap.size()));
    for (auto& d : cap) w.raw(d.data(), 32);
}

// Authentication path from `count` digests of 8 words; one direction byte
// per step, 1 when the opened node is the left child (merkle.rs path order).
const uint32_t* write_path(Writer& w, const uint32_t* p, int count, uint64_t index) {
    static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "digest words are stored little-endian");
    w.u32(uint32_t(count));
    uint8_t* o = w.room(33 * size_t(count));
    for (int l = 0; l < count; ++l, o += 33) {
        std::memcpy(o, p + 8 * l, 32);
        o[32] = ((index >> l) & 1) == 0 ? 1 : 0;
    }
    return p + 8 * count;
}

}  // namespace

enum Phase : int { PH_FRONT = 0, PH_COMPOSE, PH_FOLD, PH_GRIND, PH_OPEN, PH_DONE };
// The one-leaf-per-row AIRs: value (and bit) columns on the LDE, a joint row
// tree, an AIR-specific composition, then the shared FRI path.
enum RowAir : int { ROW_NONE = 0, ROW_RANGE, ROW_TABLE };

struct Proof {
    // request
    bool air = true;
    int row = ROW_NONE;  // a row AIR (air == false)
    z0p_air_params ap{};
    z0p_range_params rp{};
    z0p_table_params tp{};
    z0p_fri_params fp{};
    Transcript* tr = nullptr;
    const uint32_t* coeff_src = nullptr;  // FRI: host coefficients or LDE values
    int fri_in = FRI_COEFFS;
    // Layer 0 holds base-field values (4 bytes each): the FRI commitments
    // of base coefficients or base LDE values.
    bool base0() const { return !air && !row && (fri_in == FRI_COEFFS || fri_in == FRI_EVALS); }
    uint64_t src_words() const {
        if (air) return n;
        if (row) return host_a ? ncols * n : 0;
        if (fri_in == FRI_FOLD_COMBINE) return 0;
        return (fri_in == FRI_COEFFS_EXT || fri_in == FRI_EVALS_EXT ? 4 : 1) * n;
    }
    // AIR with both columns interpolated and extended (air.rs:385-406,
    // 439-464) instead of the linked second column.
    bool two_col = false;
    const uint32_t *host_a = nullptr, *host_b = nullptr;  // explicit trace columns
    std::vector<uint32_t> mask;                           // hiding: A's n_rand maskers, then B's
    uint64_t n_rand = 0;
    // Row AIRs: ncols columns, V first (range: n_bits + 1; table: 1); the
    // constant columns of the provers as their constant coefficients
    // (host_a == nullptr), or the caller's evaluation columns in host_a.
    uint64_t ncols = 0;
    std::vector<uint32_t> const_cols;
    uint32_t* rows = nullptr;  // ncols bit-reversed LDE columns, m words apart
    RangeConsts rc{};
    TableConsts tc{};
    std::vector<uint32_t> table;  // the public table, canonical
    uint32_t* table_dev = nullptr;  // the table in Montgomery form
    // FRI_FOLD_COMBINE: the committed IVC steps (device table)
    const FoldStepDesc* fold_steps = nullptr;
    uint32_t fold_nsteps = 0;
    uint64_t tag = 0;
    // geometry
    uint64_t n = 0, B = 0, m = 0, poly_len = 0, degree_bound = 0;
    int logn = 0, logb = 0;
    size_t folds = 0;
    const GeometryTables* gt = nullptr;
    const PointInverses* pinv = nullptr;  // AIR: the geometry's per-point inverses
    TreeGeom tg;                     // AIR trace tree
    std::vector<uint64_t> lens;      // FRI layer d length, d = 0..folds
    std::vector<TreeGeom> geoms;     // FRI layer d tree, d < folds
    // device memory: one pool block, or a block the proof owns (own_mem:
    // a FRI prover state that outlives its commit call)
    uint8_t* mem = nullptr;
    size_t mem_bytes = 0;
    bool own_mem = false;
    // fri::commit alone (fri.rs:408-549): stop once "fri-final" is absorbed,
    // keeping the layers and trees for a later fri::open
    bool split_commit = false;
    bool resume = false;  // fri::open of a committed state: admitted as it is
    size_t stage_need = 0;  // words of staging any one iteration may use
    uint32_t *col = nullptr, *coeff = nullptr, *scratch = nullptr, *A = nullptr, *Bc = nullptr, *trace_tree = nullptr;
    uint32_t *colB = nullptr, *coeffB = nullptr, *mask_dev = nullptr;
    uint32_t* planar = nullptr;  // FRI_COEFFS_EXT: the four coordinate LDEs
    std::vector<uint32_t*> vals, trees;
    unsigned long long* grind_res = nullptr;
    uint32_t* counters = nullptr;  // staged tree-top counters (cleared at the start of FRONT)
    std::vector<uint64_t> counter_off;  // per tree: offset of its counters
    uint32_t* tree_counters(size_t d) { return counters + counter_off[d]; }  // d = folds: the trace tree
    uint64_t ncounters = 0;
    uint8_t* consts = nullptr;  // per-phase transcript constants (pre-enqueued phases)
    uint32_t wpq = 0, fri_wpq = 0;  // opening words per query (all / FRI part)
    // AIR constants
    uint32_t a0 = 0, b0 = 0, b_last = 0, coeff_c = 0, w_last = 0;
    ComposeConsts cc{};
    // transcript-derived state
    std::vector<Digest> column_cap;
    std::vector<std::vector<Digest>> caps;
    std::vector<E4> final_values;
    FoldConsts fold{};
    GrindHead pow{};     // proof-of-work head (device layout)
    GrindJob pow_job{};  // its tail fields: fixed by the label length
    int window = 0;
    uint64_t nonce = 0;
    std::vector<uint64_t> queries;
    // progress
    Phase phase = PH_FRONT;
    size_t depth = 0;      // layer folded in the current PH_FOLD iteration
    size_t stage_off = 0;  // words, this iteration
    bool in_flight = false;  // launched; its host step has not run yet
    // result
    int status = Z0P_OK;
    Bytes proof, commitment;
    double t_submit = 0, t_admit = 0, t_done = 0;
};

namespace {

size_t layout(Proof& P, uint8_t* base) {
    size_t off = 0;
    auto take = [&](uint64_t words) -> uint32_t* {
        uint32_t* p = base ? reinterpret_cast<uint32_t*>(base + off) : nullptr;
        off += (size_t(words) * 4 + 255) & ~size_t(255);
        return p;
    };
    P.col = take(P.src_words());
    if (P.row) {
        P.coeff = take(P.ncols * P.n);
        if (P.row == ROW_TABLE) P.table_dev = take(P.table.size());
        if (P.n_rand) P.mask_dev = take(P.ncols * P.n_rand);
        P.rows = take(P.ncols * P.m);
        P.trace_tree = take(P.tg.words());
    } else {
        P.coeff = take(!P.air && P.fri_in == FRI_COEFFS_EXT ? 4 * P.n : P.n);
    }
    if (!P.air && !P.row && P.fri_in == FRI_COEFFS_EXT) P.planar = take(4 * P.m);
    if (P.two_col) {
        P.colB = take(P.n);
        P.coeffB = take(P.n);
        if (P.n_rand) P.mask_dev = take(2 * P.n_rand);
    }
    P.scratch = take(P.m);
    if (P.air) {
        P.A = take(P.m);
        P.Bc = take(P.m);
        P.trace_tree = take(P.tg.words());
    }
    P.vals.assign(P.folds + 1, nullptr);
    P.trees.assign(P.folds, nullptr);
    for (size_t d = 0; d <= P.folds; ++d) {
        P.vals[d] = take(d == 0 && P.base0() ? P.m : P.lens[d] * 4);
        if (d < P.folds) P.trees[d] = take(P.geoms[d].words());
    }
    P.grind_res = reinterpret_cast<unsigned long long*>(take(4));  // running minimum, pooled chunk counter
    // Staged-top counters: FRI layer d's tree at counter_off[d], the trace
    // (or row) tree at counter_off[folds].
    P.counter_off.assign(P.folds + 1, 0);
    P.ncounters = 0;
    for (size_t d = 0; d < P.folds; ++d) {
        P.counter_off[d] = P.ncounters;
        P.ncounters += tree_stage_counters(P.geoms[d].leaves, 0, P.geoms[d].target);
    }
    P.counter_off[P.folds] = P.ncounters;
    if (P.tg.leaves) P.ncounters += tree_stage_counters(P.tg.leaves, 0, P.tg.target);
    P.counters = take(std::max<uint64_t>(P.ncounters, 1));
    P.consts = reinterpret_cast<uint8_t*>(take(Ctx::MAILBOX_SLOTS * Ctx::MAILBOX_SLOT / 4));
    return off;
}

void init_fri_shape(Proof& P, const z0p_fri_params& fp) {
    P.fp = fp;
    P.folds = fri_num_folds(P.m, fp.final_size, P.degree_bound);
    require(P.folds <= size_t(MAX_OPEN_LAYERS), Z0P_ERR_INVALID_DOMAIN_SIZE);
    P.lens.clear();
    P.geoms.clear();
    for (uint64_t len = P.m, d = 0; d <= P.folds; ++d, len /= 4) {
        P.lens.push_back(len);
        if (d < P.folds) P.geoms.push_back(tree_geom(len / 4, fp.cap_height));
    }
    uint32_t w = 0;
    for (size_t d = 0; d < P.folds; ++d) w += (d == 0 && P.base0() ? 4 : 16) + 8 * uint32_t(P.geoms[d].target);
    P.fri_wpq = w;
    if (P.air) w += 32 + 16 * uint32_t(P.tg.target);
    if (P.row) w += uint32_t(P.ncols) + 8 * uint32_t(P.tg.target);
    P.wpq = w;
    P.stage_need = size_t(fp.num_queries) * P.wpq + 256 * 8 + P.lens[P.folds] * 4 + 256;
}

// AIR geometry over an m-point LDE: m / n cosets of n points (m = n * blowup,
// or larger for a masked trace), composition degree below degree_bound + 1.
void init_air_shape(Ctx& c, Proof& P, const z0p_air_params& p, Transcript* tr, uint64_t m, uint64_t degree_bound) {
    const z0p_fri_params fp = air_fri_params(p);
    validate_fri(fp);
    P.air = true;
    P.ap = p;
    P.tr = tr;
    P.n = p.n_steps;
    P.m = m;
    P.B = m / P.n;
    P.logn = int(ilog2(P.n));
    P.logb = int(ilog2(P.B));
    P.poly_len = P.m;
    P.degree_bound = degree_bound;
    P.tg = tree_geom(P.m / 8, p.cap_height);
    init_fri_shape(P, fp);
    P.a0 = p.a0 % P_MOD;
    P.b0 = p.b0 % P_MOD;
    P.b_last = p.b_last % P_MOD;
    // Linked LDE constant (air.rs:815-820): coeff = (b_last - a0) w_last / n.
    const uint32_t actual_b_last = fib_b_last(P.a0, P.b0, P.n);
    P.w_last = pow_c(root_of_unity(uint32_t(P.logn)), P.n - 1);
    P.coeff_c = mul(mul(sub(actual_b_last, P.a0), P.w_last), inv_c(uint32_t(P.n % P_MOD)));
    P.gt = &c.geometry(P.logn, P.logb);
    P.pinv = &c.point_inverses(P.logn, P.logb);
    P.mem_bytes = layout(P, nullptr);
}

// air::prove (air.rs:365-372): the linked second column.
void init_air(Ctx& c, Proof& P, const z0p_air_params& p, Transcript* tr) {
    validate_air(p);
    init_air_shape(c, P, p, tr, p.n_steps * p.blowup, p.n_steps - 1);
}

// air::prove_with_trace (air.rs:385-406): caller columns of n canonical values.
void init_air_trace(Ctx& c, Proof& P, const z0p_air_params& p, const uint32_t* col_a, uint64_t len_a,
                    const uint32_t* col_b, uint64_t len_b, Transcript* tr) {
    validate_air(p);
    require(len_a == p.n_steps && len_b == p.n_steps, Z0P_ERR_MALFORMED_INPUT);
    for (uint64_t i = 0; i < len_a; ++i) require(col_a[i] < P_MOD && col_b[i] < P_MOD, Z0P_ERR_MALFORMED_INPUT);
    P.two_col = true;
    P.host_a = col_a;
    P.host_b = col_b;
    init_air_shape(c, P, p, tr, p.n_steps * p.blowup, p.n_steps - 1);
}

// air::prove_zk (air.rs:439-464) with blind::mask_coeffs (blind.rs:38-79):
// each column becomes T + (x^n - 1) R over next_pow2(n + n_rand)
// coefficients. The checks and the rng draws come in the reference's order.
void init_air_zk(Ctx& c, Proof& P, const z0p_air_params& p, uint64_t n_rand, ChaCha20& rng, Transcript* tr) {
    validate_air(p);
    require(p.num_queries <= UINT64_MAX / 16, Z0P_ERR_MALFORMED_INPUT);
    require(n_rand >= 16 * p.num_queries, Z0P_ERR_MALFORMED_INPUT);
    const uint64_t n = p.n_steps;
    require(n_rand <= MAX_DOMAIN, Z0P_ERR_INVALID_DOMAIN_SIZE);
    uint64_t padded = 1;
    while (padded < n + n_rand) padded <<= 1;
    require(padded <= MAX_DOMAIN, Z0P_ERR_INVALID_DOMAIN_SIZE);
    P.mask.resize(2 * n_rand);
    for (auto& v : P.mask) v = rng.next_field();
    require(padded <= MAX_DOMAIN / p.blowup, Z0P_ERR_INVALID_DOMAIN_SIZE);  // evaluate_on_coset (ntt.rs:341-352)
    P.two_col = true;
    P.n_rand = n_rand;
    init_air_shape(c, P, p, tr, padded * p.blowup, n + n_rand - 1);
}

// Row AIR geometry over an m-point LDE (m / n cosets of n points) of ncols
// columns, one tree leaf per LDE row, composition degree below
// degree_bound + 1, committed by fri::commit_ext_evals (fri.rs:527-549).
void init_row_shape(Ctx& c, Proof& P, int kind, const z0p_fri_params& fp, uint64_t n, uint64_t ncols,
                    Transcript* tr, uint64_t m, uint64_t degree_bound) {
    validate_fri(fp);
    require(degree_bound < m, Z0P_ERR_MALFORMED_INPUT);  // rate < 1 (fri.rs:535-540)
    P.air = false;
    P.row = kind;
    P.tr = tr;
    P.n = n;
    P.ncols = ncols;
    P.m = m;
    P.B = m / P.n;
    P.logn = int(ilog2(P.n));
    P.logb = int(ilog2(P.B));
    P.poly_len = P.m;
    P.degree_bound = degree_bound;
    P.tg = tree_geom(P.m, fp.cap_height);
    init_fri_shape(P, fp);
    P.gt = &c.geometry(P.logn, P.logb);
    P.mem_bytes = layout(P, nullptr);
}

void init_range_shape(Ctx& c, Proof& P, const z0p_range_params& p, Transcript* tr, uint64_t m, uint64_t degree_bound) {
    P.rp = p;
    P.rp.value = p.value % P_MOD;
    P.rc.value = P.rp.value;
    P.rc.nbits = uint32_t(p.n_bits);
    init_row_shape(c, P, ROW_RANGE, range_fri_params(p), p.n_rows, p.n_bits + 1, tr, m, degree_bound);
}

// range::build_columns (range.rs:312-327) as coefficients: V = value and
// C_j = bit j of value interpolate to constants.
void range_const_cols(Proof& P, uint32_t value, uint64_t n_bits) {
    P.const_cols.assign(n_bits + 1, 0);
    P.const_cols[0] = value;
    for (uint64_t j = 0; j < n_bits; ++j) P.const_cols[j + 1] = (value >> j) & 1u;
}

// range::prove (range.rs:334-342).
void init_range(Ctx& c, Proof& P, const z0p_range_params& p, Transcript* tr) {
    validate_range(p);
    range_const_cols(P, p.value % P_MOD, p.n_bits);
    init_range_shape(c, P, p, tr, p.n_rows * p.blowup, p.n_rows - 1);
}

// range::prove_with_columns (range.rs:357-382): ncols columns of len
// canonical values, V first.
void init_range_cols(Ctx& c, Proof& P, const z0p_range_params& p, const uint32_t* cols, uint64_t ncols, uint64_t len,
                     Transcript* tr) {
    validate_range(p);
    require(ncols == p.n_bits + 1 && len == p.n_rows, Z0P_ERR_MALFORMED_INPUT);
    for (uint64_t i = 0; i < ncols * len; ++i) require(cols[i] < P_MOD, Z0P_ERR_MALFORMED_INPUT);
    P.host_a = cols;
    init_range_shape(c, P, p, tr, p.n_rows * p.blowup, p.n_rows - 1);
}

// range::prove_zk (range.rs:401-428): each column masked by
// blind::mask_coeffs (blind.rs:38-79), V's maskers first. The checks and the
// rng draws come in the reference's order.
void init_range_zk(Ctx& c, Proof& P, const z0p_range_params& p, uint64_t n_rand, ChaCha20& rng, Transcript* tr) {
    validate_range(p);
    require(n_rand >= p.num_queries, Z0P_ERR_MALFORMED_INPUT);
    const uint64_t n = p.n_rows;
    require(n_rand <= MAX_DOMAIN, Z0P_ERR_INVALID_DOMAIN_SIZE);
    uint64_t padded = 1;
    while (padded < n + n_rand) padded <<= 1;
    require(padded <= MAX_DOMAIN, Z0P_ERR_INVALID_DOMAIN_SIZE);
    P.mask.resize((p.n_bits + 1) * n_rand);
    for (auto& v : P.mask) v = rng.next_field();
    require(padded <= MAX_DOMAIN / p.blowup, Z0P_ERR_INVALID_DOMAIN_SIZE);  // evaluate_on_coset (ntt.rs:341-352)
    range_const_cols(P, p.value % P_MOD, p.n_bits);
    P.n_rand = n_rand;
    init_range_shape(c, P, p, tr, padded * p.blowup, n + 2 * n_rand - 2);
}

// table.rs:291-298: at least two entries.
void init_table_common(Proof& P, const z0p_table_params& p, const uint32_t* table, uint64_t table_len) {
    validate_table(p);
    require(table_len >= 2 && table, Z0P_ERR_MALFORMED_INPUT);
    P.tp = p;
    P.tp.value = p.value % P_MOD;
    P.tc.value = P.tp.value;
    P.table.resize(table_len);
    for (uint64_t k = 0; k < table_len; ++k) P.table[k] = table[k] % P_MOD;
    P.const_cols.assign(1, P.tp.value);
}

// table::prove (table.rs:305-318): the constant value column.
void init_table(Ctx& c, Proof& P, const z0p_table_params& p, const uint32_t* table, uint64_t table_len,
                Transcript* tr) {
    init_table_common(P, p, table, table_len);
    init_row_shape(c, P, ROW_TABLE, table_fri_params(p), p.n_rows, 1, tr, p.n_rows * p.blowup,
                   membership_bound(table_len, p.n_rows, 0));
}

// table::prove_zk (table.rs:329-355): the column masked by blind::mask_coeffs
// (blind.rs:38-79); checks and rng draws in the reference's order.
void init_table_zk(Ctx& c, Proof& P, const z0p_table_params& p, const uint32_t* table, uint64_t table_len,
                   uint64_t n_rand, ChaCha20& rng, Transcript* tr) {
    init_table_common(P, p, table, table_len);
    require(n_rand >= p.num_queries, Z0P_ERR_MALFORMED_INPUT);
    const uint64_t n = p.n_rows;
    require(n_rand <= MAX_DOMAIN, Z0P_ERR_INVALID_DOMAIN_SIZE);
    uint64_t padded = 1;
    while (padded < n + n_rand) padded <<= 1;
    require(padded <= MAX_DOMAIN, Z0P_ERR_INVALID_DOMAIN_SIZE);
    P.mask.resize(n_rand);
    for (auto& v : P.mask) v = rng.next_field();
    require(padded <= MAX_DOMAIN / p.blowup, Z0P_ERR_INVALID_DOMAIN_SIZE);  // evaluate_on_coset (ntt.rs:341-352)
    P.n_rand = n_rand;
    init_row_shape(c, P, ROW_TABLE, table_fri_params(p), n, 1, tr, padded * p.blowup,
                   membership_bound(table_len, n, n_rand));
}

// fri::commit / commit_ext (fri.rs:408-461): count coefficients, extended
// by the blowup. fri::commit_evals / commit_ext_evals (fri.rs:473-549): count
// LDE values with a declared degree bound. Checks in the reference's order.
void init_fri(Ctx& c, Proof& P, const uint32_t* src, uint64_t count, int kind, uint64_t degree_bound,
              const z0p_fri_params& fp, Transcript* tr) {
    validate_fri(fp);
    require(count != 0, Z0P_ERR_EMPTY_INPUT);
    require(pow2(count) && count <= MAX_DOMAIN, Z0P_ERR_INVALID_DOMAIN_SIZE);
    P.air = false;
    P.tr = tr;
    P.coeff_src = src;
    P.fri_in = kind;
    P.n = count;
    if (kind == FRI_COEFFS || kind == FRI_COEFFS_EXT) {
        require(count <= MAX_DOMAIN / fp.blowup, Z0P_ERR_INVALID_DOMAIN_SIZE);
        P.B = fp.blowup;
        P.m = count * fp.blowup;
        P.poly_len = count;
        P.degree_bound = count - 1;
    } else {
        require(degree_bound < count, Z0P_ERR_MALFORMED_INPUT);
        P.B = 1;  // the values are layer 0 itself; one "coset" of m points
        P.m = count;
        P.poly_len = count;
        P.degree_bound = degree_bound;
    }
    P.logn = int(ilog2(P.n));
    P.logb = int(ilog2(P.B));
    init_fri_shape(P, fp);
    P.gt = &c.geometry(P.logn, P.logb);
    P.mem_bytes = layout(P, nullptr);
}

// fold::Accumulator::finalize (fold.rs:510-524): fri::commit_ext_evals of the
// combined compositions over the steps' shared m-point domain (m / n cosets
// of n = seg_steps points), declared degree seg_steps - 1.
void init_fold_fri(Ctx& c, Proof& P, const FoldStepDesc* steps, uint32_t nsteps, uint64_t seg_steps, uint64_t blowup,
                   const z0p_fri_params& fp, Transcript* tr) {
    validate_fri(fp);
    require(nsteps > 0 && steps, Z0P_ERR_EMPTY_INPUT);
    require(pow2(seg_steps) && seg_steps >= 2 && pow2(blowup) && seg_steps * blowup <= MAX_DOMAIN,
            Z0P_ERR_INVALID_DOMAIN_SIZE);
    P.air = false;
    P.tr = tr;
    P.fri_in = FRI_FOLD_COMBINE;
    P.fold_steps = steps;
    P.fold_nsteps = nsteps;
    P.n = seg_steps;
    P.B = blowup;
    P.m = seg_steps * blowup;
    P.poly_len = P.m;
    P.degree_bound = seg_steps - 1;
    require(P.degree_bound < P.m, Z0P_ERR_MALFORMED_INPUT);
    P.logn = int(ilog2(P.n));
    P.logb = int(ilog2(P.B));
    init_fri_shape(P, fp);
    P.gt = &c.geometry(P.logn, P.logb);
    P.mem_bytes = layout(P, nullptr);
}

// Host work when a proof is admitted (its memory assigned).
void start(Ctx& c, Proof& P) {
    layout(P, P.mem);
    Transcript& tr = *P.tr;
    if (P.air) {
        // absorb_params (air.rs:754-766): 56 bytes, grinding_bits as u32.
        const z0p_air_params& p = P.ap;
        uint8_t pb[56];
        store_le64(pb, p.n_steps);
        store_le64(pb + 8, p.blowup);
        store_le64(pb + 16, p.num_queries);
        store_le64(pb + 24, p.final_size);
        store_le32(pb + 32, p.grinding_bits);
        store_le64(pb + 36, p.cap_height);
        store_le32(pb + 44, P.a0);
        store_le32(pb + 48, P.b0);
        store_le32(pb + 52, P.b_last);
        tr.absorb("fib/params", pb, 56);
        if (P.host_a) {
            hip_check(hipMemcpyAsync(P.col, P.host_a, P.n * 4, hipMemcpyHostToDevice, c.stream));
            hip_check(hipMemcpyAsync(P.colB, P.host_b, P.n * 4, hipMemcpyHostToDevice, c.stream));
        }
        if (P.n_rand)
            hip_check(hipMemcpyAsync(P.mask_dev, P.mask.data(), P.mask.size() * 4, hipMemcpyHostToDevice, c.stream));
    } else if (P.row) {
        if (P.row == ROW_RANGE) {
            // absorb_params (range.rs:723-735): 56 bytes, grinding_bits as u32.
            const z0p_range_params& p = P.rp;
            uint8_t pb[56];
            store_le64(pb, p.n_rows);
            store_le64(pb + 8, p.n_bits);
            store_le64(pb + 16, p.blowup);
            store_le64(pb + 24, p.num_queries);
            store_le64(pb + 32, p.final_size);
            store_le32(pb + 40, p.grinding_bits);
            store_le64(pb + 44, p.cap_height);
            store_le32(pb + 52, p.value);
            tr.absorb("range/params", pb, 56);
        } else {
            // absorb_params (table.rs:540-555): 48 bytes, |T| as u64, the entries.
            const z0p_table_params& p = P.tp;
            std::vector<uint8_t> pb(56 + 4 * P.table.size());
            store_le64(pb.data(), p.n_rows);
            store_le64(pb.data() + 8, p.blowup);
            store_le64(pb.data() + 16, p.num_queries);
            store_le64(pb.data() + 24, p.final_size);
            store_le32(pb.data() + 32, p.grinding_bits);
            store_le64(pb.data() + 36, p.cap_height);
            store_le32(pb.data() + 44, p.value);
            store_le64(pb.data() + 48, P.table.size());
            std::vector<uint32_t> tm(P.table.size());
            for (size_t k = 0; k < P.table.size(); ++k) {
                store_le32(pb.data() + 56 + 4 * k, P.table[k]);
                tm[k] = to_mont(P.table[k]);
            }
            tr.absorb("table/params", pb);
            hip_check(hipMemcpyAsync(P.table_dev, tm.data(), tm.size() * 4, hipMemcpyHostToDevice, c.stream));
            hip_check(hipStreamSynchronize(c.stream));  // tm is a local
        }
        if (P.host_a) {
            hip_check(hipMemcpyAsync(P.col, P.host_a, P.ncols * P.n * 4, hipMemcpyHostToDevice, c.stream));
        } else {
            // Constant columns: coefficient 0 of each column, zeros above.
            hip_check(hipMemsetAsync(P.coeff, 0, P.ncols * P.n * 4, c.stream));
            hip_check(hipMemcpy2DAsync(P.coeff, P.n * 4, P.const_cols.data(), 4, 4, P.ncols, hipMemcpyHostToDevice,
                                       c.stream));
        }
        if (P.n_rand)
            hip_check(hipMemcpyAsync(P.mask_dev, P.mask.data(), P.mask.size() * 4, hipMemcpyHostToDevice, c.stream));
    } else {
        uint8_t dims[24];
        store_le64(dims, P.m);
        store_le64(dims + 8, P.poly_len);
        store_le64(dims + 16, P.degree_bound);
        tr.absorb("fri-dims", dims, 24);
        if (P.src_words())
            hip_check(hipMemcpyAsync(P.col, P.coeff_src, P.src_words() * 4, hipMemcpyHostToDevice, c.stream));
    }
    P.phase = PH_FRONT;
}

uint32_t* stage_take(Ctx& c, size_t& used, size_t words, size_t& off) {
    off = used;
    used += (words + 63) & ~size_t(63);
    require(used <= c.stage_words, Z0P_ERR_DEVICE);
    return c.stage_dev + off;
}

void collect_cap(Ctx& c, Proof& P, uint32_t* tree, const TreeGeom& g, std::vector<CollectItem>& items,
                 size_t& used) {
    const uint32_t words = uint32_t(8u << g.drop);
    uint32_t* dst = stage_take(c, used, words, P.stage_off);
    items.push_back(CollectItem{tree + g.level_offset_words(g.target), dst, words, 0});
}

// Layer k (> 0, with a tree) committed in one launch with its fold
// (k_fold_commit): 2^8..2^fold_commit_max_log values (whole workgroups of
// 256 outputs); it takes precedence over the one-workgroup small commit.
bool fold_commit_layer(const Proof& P, size_t k, const Schedule& s) {
    return k < P.folds && s.fold_commit_max_log > 0 && P.lens[k] >= (uint64_t(1) << 8) &&
           P.lens[k] <= (uint64_t(1) << s.fold_commit_max_log) && P.geoms[k].target >= 1;
}

LeavesJob leaves_of(int kind, const uint32_t* src) {
    LeavesJob j{};
    j.kind = kind;
    j.src = src;
    return j;
}

// The device work of this proof's current phase. `cst` is the device
// address of this phase's transcript constants when they are supplied
// later (pre-enqueued phases); otherwise they are uploaded now.
void build(Ctx& c, Proof& P, std::vector<Step>& ch, std::vector<CollectItem>& items, size_t& used,
           const void* cst = nullptr) {
    const GeometryTables& gt = *P.gt;
    // The staged tree tops find their counters zero: FRONT clears them first
    // (its own trees use them), and every use leaves them zero.
    if (P.phase == PH_FRONT && P.counters) ch.push_back(ZeroJob{P.counters, std::max<uint64_t>(P.ncounters, 1), 0});
    switch (P.phase) {
        case PH_FRONT: {
            if (P.row) {
                const uint32_t ninv = to_mont(inv_c(uint32_t(P.n % P_MOD)));
                for (uint64_t k = 0; k < P.ncols; ++k) {
                    uint32_t* coeff = P.coeff + k * P.n;
                    if (P.host_a) {
                        NttPlan ip = plan_intt(P.col + k * P.n, coeff, P.scratch, P.logn, c.tw4096_inv, gt.winv,
                                               P.logb, ninv);
                        for (auto& j : ip.pass1) ch.push_back(j);
                        for (auto& j : ip.pass2) ch.push_back(j);
                    }
                    NttPlan lp = plan_lde(coeff, P.rows + k * P.m, P.scratch, P.logn, P.logb, c.tw4096, gt.w, gt.coset);
                    for (auto& j : lp.pass1) {
                        if (P.n_rand) {
                            j.mask = P.mask_dev + k * P.n_rand;
                            j.zvan = gt.zvan;
                            j.mask_len = uint32_t(P.n_rand);
                        }
                        ch.push_back(j);
                    }
                    for (auto& j : lp.pass2) ch.push_back(j);
                }
                LeavesJob t = leaves_of(LEAF_ROWS, P.rows);
                t.ncols = int(P.ncols);
                t.col_stride = P.m;
                t.logn = P.logn;
                t.logb = P.logb;
                tree_steps(ch, P.trace_tree, P.tg, t, c.sched, P.tree_counters(P.folds));
                collect_cap(c, P, P.trace_tree, P.tg, items, used);
            } else if (P.air && P.two_col) {
                if (!P.host_a) {
                    // Both Fibonacci columns' coefficients in closed form.
                    const auto ca = fib_binet(P.a0, P.b0);
                    const auto cb = fib_binet_next(ca);
                    ch.push_back(fib_coeff_job(P.coeff, gt, P.logn, P.logb, ca.first, ca.second));
                    ch.push_back(fib_coeff_job(P.coeffB, gt, P.logn, P.logb, cb.first, cb.second));
                } else {
                    const uint32_t ninv = to_mont(inv_c(uint32_t(P.n % P_MOD)));
                    for (int k = 0; k < 2; ++k) {
                        NttPlan ip = plan_intt(k ? P.colB : P.col, k ? P.coeffB : P.coeff, P.scratch, P.logn,
                                               c.tw4096_inv, gt.winv, P.logb, ninv);
                        for (auto& j : ip.pass1) ch.push_back(j);
                        for (auto& j : ip.pass2) ch.push_back(j);
                    }
                }
                for (int k = 0; k < 2; ++k) {
                    NttPlan lp = plan_lde(k ? P.coeffB : P.coeff, k ? P.Bc : P.A, P.scratch, P.logn, P.logb,
                                          c.tw4096, gt.w, gt.coset);
                    for (auto& j : lp.pass1) {
                        if (P.n_rand) {
                            j.mask = P.mask_dev + k * P.n_rand;
                            j.zvan = gt.zvan;
                            j.mask_len = uint32_t(P.n_rand);
                        }
                        ch.push_back(j);
                    }
                    for (auto& j : lp.pass2) ch.push_back(j);
                }
                LeavesJob t = leaves_of(LEAF_TRACE, P.A);
                t.src_b = P.Bc;
                t.logn = P.logn;
                t.logb = P.logb;
                tree_steps(ch, P.trace_tree, P.tg, t, c.sched, P.tree_counters(P.folds));
                collect_cap(c, P, P.trace_tree, P.tg, items, used);
            } else if (P.air) {
                // interpolate(build_trace(a0, b0, n).0) (air.rs:368-369) in closed form.
                const auto ca = fib_binet(P.a0, P.b0);
                ch.push_back(fib_coeff_job(P.coeff, gt, P.logn, P.logb, ca.first, ca.second));
                NttPlan lp = plan_lde(P.coeff, P.A, P.scratch, P.logn, P.logb, c.tw4096, gt.w, gt.coset);
                for (auto& j : lp.pass1) ch.push_back(j);
                for (auto& j : lp.pass2) ch.push_back(j);
                LeavesJob t = leaves_of(LEAF_TRACE, P.A);
                t.src_b = P.Bc;
                t.logn = P.logn;
                t.logb = P.logb;
                if (c.sched.trace_fuse && P.tg.leaves >= 2) {
                    // The linked column computed by the trace-leaves kernel.
                    t.kind = LEAF_TRACE_LINKED;
                    t.b_out = P.Bc;
                    t.iw = P.pinv->iw;
                    t.zvan_m = gt.zvan;
                    t.coeff = P.coeff_c;
                } else {
                    LinkedJob lj{};
                    lj.a = P.A;
                    lj.b = P.Bc;
                    lj.iw = P.pinv->iw;
                    lj.zvan_m = gt.zvan;
                    lj.coeff = P.coeff_c;
                    lj.logn = P.logn;
                    lj.logb = P.logb;
                    ch.push_back(lj);
                }
                tree_steps(ch, P.trace_tree, P.tg, t, c.sched, P.tree_counters(P.folds));
                collect_cap(c, P, P.trace_tree, P.tg, items, used);
            } else {
                switch (P.fri_in) {
                    case FRI_COEFFS: {
                        ch.push_back(BitrevJob{P.col, P.coeff, P.logn, 0});
                        NttPlan lp = plan_lde(P.coeff, P.vals[0], P.scratch, P.logn, P.logb, c.tw4096, gt.w, gt.coset);
                        for (auto& j : lp.pass1) ch.push_back(j);
                        for (auto& j : lp.pass2) ch.push_back(j);
                        break;
                    }
                    case FRI_COEFFS_EXT: {
                        // Each coordinate extended over the base field (ntt.rs:394-420),
                        // then interleaved into Ext4 layer 0.
                        for (int k = 0; k < 4; ++k) {
                            BitrevJob b{P.col + k, P.coeff + k * P.n, P.logn, 0};
                            b.in_stride = 4;
                            ch.push_back(b);
                        }
                        for (int k = 0; k < 4; ++k) {
                            NttPlan lp = plan_lde(P.coeff + k * P.n, P.planar + k * P.m, P.scratch, P.logn, P.logb,
                                                  c.tw4096, gt.w, gt.coset);
                            for (auto& j : lp.pass1) ch.push_back(j);
                            for (auto& j : lp.pass2) ch.push_back(j);
                        }
                        ch.push_back(InterleaveJob{P.planar, P.vals[0], P.m, 0});
                        break;
                    }
                    case FRI_EVALS:
                        ch.push_back(BitrevJob{P.col, P.vals[0], P.logn, 0});
                        break;
                    case FRI_FOLD_COMBINE: {
                        FoldCombineJob fj{};
                        fj.steps = P.fold_steps;
                        fj.nsteps = P.fold_nsteps;
                        fj.layer0 = P.vals[0];
                        fj.xn = gt.xn;
                        fj.shift_m = gt.coset.shift;
                        fj.zinv_m = gt.zinv;
                        fj.w_last_m = to_mont(pow_c(root_of_unity(uint32_t(P.logn)), P.n - 1));
                        fj.logn = P.logn;
                        fj.logb = P.logb;
                        ch.push_back(fj);
                        break;
                    }
                    default: {  // FRI_EVALS_EXT
                        BitrevJob b{P.col, P.vals[0], P.logn, 0};
                        b.in_stride = 4;
                        b.width = 4;
                        ch.push_back(b);
                        break;
                    }
                }
                tree_steps(ch, P.trees[0], P.geoms[0], leaves_of(P.base0() ? LEAF_BASE : LEAF_EXT, P.vals[0]),
                           c.sched, P.tree_counters(0));
                collect_cap(c, P, P.trees[0], P.geoms[0], items, used);
            }
            break;
        }
        case PH_COMPOSE: {
            if (P.row == ROW_TABLE) {
                TableComposeJob tj{};
                tj.v = P.rows;
                tj.layer0 = P.vals[0];
                tj.xn = gt.xn;
                tj.shift_m = gt.coset.shift;
                tj.zinv_m = gt.zinv;
                tj.c = static_cast<const TableConsts*>(cst ? cst : c.ring.push(&P.tc, sizeof(P.tc)));
                tj.table_m = P.table_dev;
                tj.table_len = uint32_t(P.table.size());
                tj.logn = P.logn;
                tj.logb = P.logb;
                ch.push_back(tj);
                tree_steps(ch, P.trees[0], P.geoms[0], leaves_of(LEAF_EXT, P.vals[0]), c.sched, P.tree_counters(0));
                collect_cap(c, P, P.trees[0], P.geoms[0], items, used);
                break;
            }
            if (P.row == ROW_RANGE) {
                RangeComposeJob rj{};
                rj.cols = P.rows;
                rj.col_stride = P.m;
                rj.layer0 = P.vals[0];
                rj.xn = gt.xn;
                rj.shift_m = gt.coset.shift;
                rj.zinv_m = gt.zinv;
                rj.c = static_cast<const RangeConsts*>(cst ? cst : c.ring.push(&P.rc, sizeof(P.rc)));
                rj.logn = P.logn;
                rj.logb = P.logb;
                ch.push_back(rj);
                tree_steps(ch, P.trees[0], P.geoms[0], leaves_of(LEAF_EXT, P.vals[0]), c.sched, P.tree_counters(0));
                collect_cap(c, P, P.trees[0], P.geoms[0], items, used);
                break;
            }
            ComposeJob cj{};
            cj.a = P.A;
            cj.b = P.Bc;
            cj.layer0 = P.vals[0];
            cj.pinv = *P.pinv;
            cj.c = static_cast<const ComposeConsts*>(cst ? cst : c.ring.push(&P.cc, sizeof(P.cc)));
            cj.logn = P.logn;
            cj.logb = P.logb;
            if (c.sched.compose_fuse) {
                // Layer 0's leaves (and their parents) hashed by the composition kernel.
                cj.tree = P.trees[0];
                cj.with_parent = P.geoms[0].target >= 1 ? 1 : 0;
                Schedule cs = c.sched;
                cs.subtree_levels = c.sched.compose_subtree;
                cj.sub_levels = cj.with_parent ? subtree_levels_for(P.geoms[0].leaves / 2, P.geoms[0].target - 1, cs)
                                               : 0;
                ch.push_back(cj);
                tree_levels(ch, P.trees[0], P.geoms[0], cj.with_parent + cj.sub_levels, c.sched, P.tree_counters(0));
            } else {
                ch.push_back(cj);
                tree_steps(ch, P.trees[0], P.geoms[0], leaves_of(LEAF_EXT, P.vals[0]), c.sched, P.tree_counters(0));
            }
            collect_cap(c, P, P.trees[0], P.geoms[0], items, used);
            break;
        }
        case PH_FOLD: {
            const size_t d = P.depth;
            const FoldConsts* fc = static_cast<const FoldConsts*>(cst ? cst : c.ring.push(&P.fold, sizeof(P.fold)));
            FoldJob fj{};
            fj.layer = P.vals[d];
            fj.next = P.vals[d + 1];
            fj.finv = gt.finv;
            fj.c = fc;
            fj.log_len = int(ilog2(P.lens[d]));
            fj.base_field = d == 0 && P.base0() ? 1 : 0;
            if (fold_commit_layer(P, d + 1, c.sched)) {
                // Layer d+1's whole commitment in this launch (k_fold_commit).
                fj.tree = P.trees[d + 1];
                fj.counters = P.tree_counters(d + 1);
                fj.target = P.geoms[d + 1].target;
                ch.push_back(fj);
                collect_cap(c, P, P.trees[d + 1], P.geoms[d + 1], items, used);
                break;
            }
            if (d + 1 < P.folds && c.sched.small_max_log > 0 && P.lens[d + 1] <= (uint64_t(1) << c.sched.small_max_log)) {
                // Small next layer: fold, leaves and tree in one workgroup.
                SmallCommitJob sj{};
                sj.layer = P.vals[d];
                sj.next = P.vals[d + 1];
                sj.finv = gt.finv;
                sj.c = fc;
                sj.tree = P.trees[d + 1];
                sj.leaves = uint32_t(P.geoms[d + 1].leaves);
                sj.target = P.geoms[d + 1].target;
                sj.base_field = d == 0 && P.base0() ? 1 : 0;
                ch.push_back(sj);
                collect_cap(c, P, P.trees[d + 1], P.geoms[d + 1], items, used);
                break;
            }
            if (d + 1 == P.folds) {
                ch.push_back(fj);
                const uint32_t words = uint32_t(P.lens[d + 1] * 4);
                uint32_t* dst = stage_take(c, used, words, P.stage_off);
                items.push_back(CollectItem{P.vals[d + 1], dst, words, 0});
            } else if (c.sched.fold_fuse_max_log > 0 && P.lens[d + 1] <= (uint64_t(1) << c.sched.fold_fuse_max_log)) {
                // Layer d+1's leaves (and parents) hashed by the fold kernel.
                fj.tree = P.trees[d + 1];
                fj.with_parent = P.geoms[d + 1].target >= 1 ? 1 : 0;
                fj.sub_levels = fj.with_parent ? subtree_levels_for(P.geoms[d + 1].leaves / 2,
                                                                    P.geoms[d + 1].target - 1, c.sched)
                                               : 0;
                ch.push_back(fj);
                tree_levels(ch, P.trees[d + 1], P.geoms[d + 1], fj.with_parent + fj.sub_levels, c.sched,
                            P.tree_counters(d + 1));
                collect_cap(c, P, P.trees[d + 1], P.geoms[d + 1], items, used);
            } else {
                ch.push_back(fj);
                tree_steps(ch, P.trees[d + 1], P.geoms[d + 1], leaves_of(LEAF_EXT, P.vals[d + 1]), c.sched,
                           P.tree_counters(d + 1));
                collect_cap(c, P, P.trees[d + 1], P.geoms[d + 1], items, used);
            }
            break;
        }
        case PH_GRIND: {
            GrindJob g{};
            if (P.pow_job.len0 == 0) {
                // The tail layout depends only on the head's length (47 bytes
                // for "fri-pow"), so a phase enqueued ahead knows it already.
                const uint8_t zeros[47] = {0};
                GrindHead unused;
                require(grind_prepare(zeros, sizeof(zeros), unused, P.pow_job), Z0P_ERR_DEVICE);
            }
            g = P.pow_job;
            g.head = static_cast<const GrindHead*>(cst ? cst : c.ring.push(&P.pow, sizeof(P.pow)));
            g.bits = P.fp.grinding_bits;
            g.nblocks = uint32_t(std::min(2048, std::max(16, c.sched.grind_blocks)));  // one proof; batches re-spread
            // Window w scans [2^(20+4w), 2^(24+4w)) after the windows before it.
            g.start = P.window == 0 ? 0 : uint64_t(1) << std::min<uint64_t>(63, 20 + 4 * uint64_t(P.window));
            g.limit = uint64_t(1) << std::min<uint64_t>(63, 24 + 4 * uint64_t(P.window));
            g.result = P.grind_res;
            ch.push_back(g);
            uint32_t* dst = stage_take(c, used, 2, P.stage_off);
            items.push_back(CollectItem{reinterpret_cast<const uint32_t*>(P.grind_res), dst, 2, 0});
            break;
        }
        case PH_OPEN: {
            OpenJob o{};
            for (size_t d = 0; d < P.folds; ++d) {
                o.layer[d] = P.vals[d];
                o.tree[d] = P.trees[d];
                o.tree_leaves[d] = P.geoms[d].leaves;
                o.target[d] = P.geoms[d].target;
                o.base[d] = d == 0 && P.base0() ? 1 : 0;
            }
            o.nlayers = int(P.folds);
            if (P.air) {
                o.a = P.A;
                o.b = P.Bc;
                o.trace_tree = P.trace_tree;
                o.trace_leaves = P.tg.leaves;
                o.trace_target = P.tg.target;
            } else if (P.row) {
                o.cols = P.rows;
                o.ncols = int(P.ncols);
                o.col_stride = P.m;
                o.trace_tree = P.trace_tree;
                o.trace_leaves = P.tg.leaves;
                o.trace_target = P.tg.target;
            }
            o.logn = P.logn;
            o.logb = P.logb;
            if (cst) {
                // Enqueued ahead: [u64 live][u64 queries...] arrive in the phase's constants.
                o.live = static_cast<const uint64_t*>(cst);
                o.queries = o.live + 1;
            } else {
                o.queries = static_cast<const uint64_t*>(c.ring.push(P.queries.data(), P.queries.size() * 8));
            }
            o.nq = uint32_t(P.fp.num_queries);
            o.words_per_query = P.wpq;
            o.out = stage_take(c, used, size_t(P.wpq) * P.fp.num_queries, P.stage_off);
            ch.push_back(o);
            break;
        }
        case PH_DONE:
            break;
    }
}

// The layer just committed is layer d: absorb its cap, draw the three fold
// challenges of fold d (fri.rs:594-612).
void commit_layer(Proof& P, const uint32_t* w, size_t d) {
    const TreeGeom& g = P.geoms[d];
    auto cap = read_digests(w, uint64_t(1) << g.drop);
    for (auto& dg : cap) P.tr->absorb("fri-layer-root", dg.data(), 32);
    P.caps.push_back(std::move(cap));
    E4 r[3];
    for (int k = 0; k < 3; ++k) r[k] = P.tr->challenge_ext("fri-fold");
    const FoldConsts fc = make_fold_consts(r, int(ilog2(P.lens[d])), int(d));
    P.fold = fc;
    P.depth = d;
    P.phase = PH_FOLD;
}

void begin_open(Proof& P);
void write_fri_commitment_body(Writer& w, const Proof& P);

void sample_queries(Proof& P) {
    P.queries.clear();
    for (uint64_t q = 0; q < P.fp.num_queries; ++q) P.queries.push_back(P.tr->challenge_index("fri-query", P.m));
    P.phase = PH_OPEN;
}

void finish_final(Proof& P, const uint32_t* w) {
    const uint64_t len = P.lens[P.folds];
    P.final_values.resize(len);
    for (uint64_t i = 0; i < len; ++i) P.final_values[i] = E4{{w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]}};
    auto cap = host_layer_cap(P.final_values, P.fp.cap_height);
    for (auto& d : cap) P.tr->absorb("fri-layer-root", d.data(), 32);
    P.caps.push_back(std::move(cap));
    std::vector<uint8_t> fin(len * 16);
    for (uint64_t i = 0; i < len; ++i)
        for (int c = 0; c < 4; ++c) store_le32(fin.data() + 16 * i + 4 * c, P.final_values[i].v[c]);
    P.tr->absorb("fri-final", fin);
    if (P.split_commit) {
        // fri::commit ends here (fri.rs:629-632); fri::open resumes later.
        Writer wc;
        wc.byte(WIRE_VERSION);
        write_fri_commitment_body(wc, P);
        P.commitment = wc.finish();
        P.phase = PH_DONE;
        return;
    }
    begin_open(P);
}

// fri::open (fri.rs:645-666): the proof-of-work on the device when
// grinding, then the query indices.
void begin_open(Proof& P) {
    if (P.fp.grinding_bits > 0) {
        // Proof-of-work (transcript.rs:206-228) on the device.
        std::vector<uint8_t> head = P.tr->pow_head("fri-pow");
        require(head.size() == 47 && grind_prepare(head.data(), head.size(), P.pow, P.pow_job), Z0P_ERR_DEVICE);
        P.window = 0;
        P.phase = PH_GRIND;
    } else {
        sample_queries(P);
    }
}

// fri::Commitment::write_body (fri.rs:896-904).
void write_fri_commitment_body(Writer& w, const Proof& P) {
    w.u64(P.m);
    w.u64(P.poly_len);
    w.u64(P.degree_bound);
    w.u32(uint32_t(P.caps.size()));
    for (auto& c : P.caps) write_cap(w, c);
}

// fri::Proof::write_body (fri.rs:352-368) from the opened words.
void write_fri_proof_body(Writer& w, const Proof& P, const uint32_t* words) {
    w.u32(uint32_t(P.final_values.size()));
    for (auto& v : P.final_values) w.e4(v);
    w.u64(P.nonce);
    w.u32(uint32_t(P.queries.size()));
    for (size_t q = 0; q < P.queries.size(); ++q) {
        const uint32_t* p = words + q * P.wpq;
        w.u32(uint32_t(P.folds));
        uint64_t index = P.queries[q];
        for (size_t d = 0; d < P.folds; ++d) {
            if (d == 0 && P.base0()) {
                for (int k = 0; k < 4; ++k) w.e4(e4_base(p[k]));
                p += 4;
            } else {
                w.words(p, 16);
                p += 16;
            }
            p = write_path(w, p, P.geoms[d].target, index / 4);
            index /= 4;
        }
    }
}

// Proof::to_bytes (air.rs:262-281) / the FRI commitment and proof.
void serialize(Proof& P, const uint32_t* words) {
    // Opened words dominate the size: 4 bytes each plus a direction byte
    // per path digest (one per 8 words); caps and headers add a few kilobytes.
    const size_t estimate = size_t(P.wpq) * P.queries.size() * 33 / 8 + 32768;
    if (P.air) {
        Writer w;
        w.reserve(estimate);
        w.byte(WIRE_VERSION);
        write_cap(w, P.column_cap);
        write_fri_commitment_body(w, P);
        write_fri_proof_body(w, P, words);
        w.u32(uint32_t(P.queries.size()));
        const uint64_t B = P.B, m = P.m;
        const int logm = P.logn + P.logb;
        for (size_t q = 0; q < P.queries.size(); ++q) {
            const uint32_t* lw = words + q * P.wpq + P.fri_wpq;
            const uint32_t* sw = lw + 16;
            const uint32_t* paths = lw + 32;
            const uint64_t index = brev_host(P.queries[q], logm);
            const uint64_t shifted = (index + B) % m;
            w.u64(index);
            w.u32(lw[2 * (index & 7)]);
            w.u32(lw[2 * (index & 7) + 1]);
            w.u32(sw[2 * (shifted & 7)]);
            w.u32(sw[2 * (shifted & 7) + 1]);
            paths = write_path(w, paths, P.tg.target, index / 8);
            write_path(w, paths, P.tg.target, shifted / 8);
            w.words(lw, 16);
            w.words(sw, 16);
        }
        P.proof = w.finish();
    } else if (P.row) {
        // range::Proof::to_bytes (range.rs:240-260), table::Proof::to_bytes
        // (table.rs:244-262: no bit columns).
        Writer w;
        w.reserve(estimate);
        w.byte(WIRE_VERSION);
        write_cap(w, P.column_cap);
        write_fri_commitment_body(w, P);
        write_fri_proof_body(w, P, words);
        w.u32(uint32_t(P.queries.size()));
        const int logm = P.logn + P.logb;
        for (size_t q = 0; q < P.queries.size(); ++q) {
            const uint32_t* rw = words + q * P.wpq + P.fri_wpq;
            const uint64_t index = brev_host(P.queries[q], logm);
            w.u64(index);
            w.u32(rw[0]);
            if (P.row == ROW_RANGE) {
                w.u32(uint32_t(P.ncols - 1));
                w.words(rw + 1, P.ncols - 1);
            }
            write_path(w, rw + P.ncols, P.tg.target, index);
        }
        P.proof = w.finish();
    } else {
        Writer wc;
        wc.byte(WIRE_VERSION);
        write_fri_commitment_body(wc, P);
        P.commitment = wc.finish();
        Writer wp;
        wp.reserve(estimate);
        wp.byte(WIRE_VERSION);
        write_fri_proof_body(wp, P, words);
        P.proof = wp.finish();
    }
}

// Host transcript step after the iteration's device work completed.
void host_step(Ctx& c, Proof& P) {
    const uint32_t* w = c.stage_host + P.stage_off;
    switch (P.phase) {
        case PH_FRONT:
            if (P.air) {
                P.column_cap = read_digests(w, uint64_t(1) << P.tg.drop);
                std::vector<uint8_t> cb(P.column_cap.size() * 32);
                for (size_t i = 0; i < P.column_cap.size(); ++i) std::memcpy(cb.data() + 32 * i, P.column_cap[i].data(), 32);
                P.tr->absorb("fib/trace", cb);
                for (int k = 0; k < 5; ++k) {
                    E4 al = P.tr->challenge_ext("fib/alpha");
                    for (int q = 0; q < 4; ++q) P.cc.alpha_m[k][q] = to_mont(al.v[q]);
                }
                P.cc.w_last_m = to_mont(P.w_last);
                P.cc.a0 = P.a0;
                P.cc.b0 = P.b0;
                P.cc.b_last = P.b_last;
                uint8_t dims[24];
                store_le64(dims, P.m);
                store_le64(dims + 8, P.poly_len);
                store_le64(dims + 16, P.degree_bound);
                P.tr->absorb("fri-dims", dims, 24);
                P.phase = PH_COMPOSE;
            } else if (P.row) {
                // range.rs:464-475: the row cap, then 1 + n_bits + 1 alphas;
                // table.rs:391-397: the column cap, then 2 alphas.
                P.column_cap = read_digests(w, uint64_t(1) << P.tg.drop);
                std::vector<uint8_t> cb(P.column_cap.size() * 32);
                for (size_t i = 0; i < P.column_cap.size(); ++i) std::memcpy(cb.data() + 32 * i, P.column_cap[i].data(), 32);
                const bool rng_air = P.row == ROW_RANGE;
                P.tr->absorb(rng_air ? "range/column" : "table/column", cb);
                const uint64_t nalpha = rng_air ? P.ncols + 1 : 2;
                for (uint64_t k = 0; k < nalpha; ++k) {
                    E4 al = P.tr->challenge_ext(rng_air ? "range/alpha" : "table/alpha");
                    for (int q = 0; q < 4; ++q) {
                        if (rng_air) P.rc.alpha_m[k][q] = to_mont(al.v[q]);
                        else P.tc.alpha_m[k][q] = to_mont(al.v[q]);
                    }
                }
                uint8_t dims[24];
                store_le64(dims, P.m);
                store_le64(dims + 8, P.poly_len);
                store_le64(dims + 16, P.degree_bound);
                P.tr->absorb("fri-dims", dims, 24);
                P.phase = PH_COMPOSE;
            } else {
                commit_layer(P, w, 0);
            }
            break;
        case PH_COMPOSE:
            commit_layer(P, w, 0);
            break;
        case PH_FOLD:
            if (P.depth + 1 == P.folds) finish_final(P, w);
            else commit_layer(P, w, P.depth + 1);
            break;
        case PH_GRIND: {
            const uint64_t found = uint64_t(w[0]) | (uint64_t(w[1]) << 32);
            if (found != ~uint64_t(0)) {
                P.nonce = found;
                uint8_t nb[8];
                store_le64(nb, found);
                P.tr->absorb("fri-pow", nb, 8);
                sample_queries(P);
            } else {
                // Each window searches [0, limit); later windows only widen
                // the bound, and the kernel returns the smallest passing nonce.
                ++P.window;
            }
            break;
        }
        case PH_OPEN: {
            const double t0 = now_ms();
            serialize(P, w);
            c.prof.host("~host_serialize", now_ms() - t0);
            P.phase = PH_DONE;
            break;
        }
        case PH_DONE:
            break;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Scheduler.

Engine::Engine(Ctx& c, size_t budget, size_t max_active) : c_(c), budget_(budget), max_active_(max_active) {
    if (max_active_ == 0) max_active_ = 1;
    if (budget_ && c_.pool.in_use() == 0 && c_.pool.capacity() < budget_) {
        c_.sync();
        c_.pool.reset_capacity(budget_);
    }
}

size_t auto_budget(Ctx& c, size_t want) {
    size_t free_b = 0, total_b = 0;
    hip_check(hipMemGetInfo(&free_b, &total_b));
    if (c.pool.in_use() == 0) free_b += c.pool.capacity();
    return std::min(want, free_b / 5 * 4);
}

Engine::~Engine() {
    // Proof memory goes back to the pool once no launch can still use it.
    if (!flight_.empty()) {
        (void)hipStreamSynchronize(c_.alt_stream);
        (void)hipStreamSynchronize(c_.stream);
    }
    for (Proof* p : active_)
        if (p->mem && !p->own_mem) c_.pool.release(p->mem, p->mem_bytes);
}

void Engine::submit(Proof* p) {
    p->t_submit = now_ms();
    queue_.push_back(p);
}

void Engine::finish(Proof* p, int status) {
    if (p->mem && !p->own_mem) {
        c_.pool.release(p->mem, p->mem_bytes);
        p->mem = nullptr;
    }
    p->status = status;
    p->phase = PH_DONE;
    p->t_done = now_ms();
    if (on_done_) on_done_(p);
}

void Engine::admit() {
    while (!queue_.empty() && active_.size() < max_active_) {
        Proof* p = queue_.front();
        if (p->own_mem) {
            // Memory of its own: no pool block and no budget share. A
            // resumed state keeps its device data and transcript position.
            queue_.pop_front();
            p->t_admit = now_ms();
            if (!p->resume) start(c_, *p);
            active_.push_back(p);
            continue;
        }
        uint8_t* mem = nullptr;
        if (budget_ && c_.pool.in_use() != 0 && c_.pool.in_use() + p->mem_bytes > budget_) break;
        if (!c_.pool.alloc(p->mem_bytes, mem)) {
            if (c_.pool.in_use() != 0) break;  // wait for memory to free up
            // Idle pool: grow it to fit (bounded by the budget, if any).
            size_t want = budget_ ? std::max(budget_, p->mem_bytes) : p->mem_bytes;
            if (budget_ && p->mem_bytes > budget_) {
                queue_.pop_front();
                finish(p, Z0P_ERR_OUT_OF_MEMORY);
                continue;
            }
            c_.sync();
            c_.pool.reset_capacity(want);
            if (!c_.pool.alloc(p->mem_bytes, mem)) {
                queue_.pop_front();
                finish(p, Z0P_ERR_OUT_OF_MEMORY);
                continue;
            }
        }
        queue_.pop_front();
        p->mem = mem;
        p->t_admit = now_ms();
        start(c_, *p);
        active_.push_back(p);
    }
}

// Builds and enqueues the current phase of every proof in `ws` as one launch
// sequence; returns the sequence number its completion flag will carry.
uint64_t Engine::launch(const std::vector<Proof*>& ws) {
    size_t need = 0;
    for (Proof* p : ws) need += p->stage_need;
    c_.reserve_stage(need);
    std::vector<std::vector<Step>> chains(ws.size());
    std::vector<CollectItem> items;
    size_t used = 0;
    for (size_t i = 0; i < ws.size(); ++i) build(c_, *ws[i], chains[i], items, used);
    // Grind sweeps: a job's workgroups scan its nonces in order, so a
    // smaller sweep per job wastes fewer hashes above the smallest passing
    // nonce; spread the schedule's workgroups over the jobs grinding now.
    size_t grinding = 0;
    for (auto& ch : chains)
        for (auto& st : ch) grinding += std::holds_alternative<GrindJob>(st) ? 1 : 0;
    if (grinding) {
        const uint32_t per = uint32_t(std::min<size_t>(2048, std::max<size_t>(16, size_t(c_.sched.grind_blocks) / grinding)));
        for (auto& ch : chains)
            for (auto& st : ch)
                if (auto* g = std::get_if<GrindJob>(&st)) g->nblocks = per;
    }
    std::vector<const std::vector<Step>*> ptrs;
    for (auto& ch : chains) ptrs.push_back(&ch);
    const uint64_t seq = ++c_.seq;
    c_.enqueue([&](const Launch& L) { launch_chains(L, ptrs, items, c_.flag_dev, seq); });
    for (Proof* p : ws) p->in_flight = true;
    return seq;
}

// Waits for launch `seq` of the proofs in `ws`, runs their host transcript
// steps and ends those that are done (or failed).
void Engine::complete(const std::vector<Proof*>& ws, uint64_t seq, bool reset_ring) {
    const double t_wait = now_ms();
    c_.wait_flag(seq, true, reset_ring);
    const double t_host = now_ms();
    ++iterations_;
    // Host transcript steps: independent per proof, so a batch spreads them
    // over the host pool.
    std::vector<int> err(ws.size(), Z0P_OK);
    auto step = [&](size_t i) {
        try {
            host_step(c_, *ws[i]);
        } catch (const Status& s) {
            err[i] = s.code;
        } catch (...) {
            err[i] = Z0P_ERR_DEVICE;
        }
    };
    if (ws.size() >= 4) {
        c_.host_workers().run(ws.size(), step);
    } else {
        for (size_t i = 0; i < ws.size(); ++i) step(i);
    }
    c_.prof.host("~host_wait_device", t_host - t_wait);
    c_.prof.host("~host_transcript", now_ms() - t_host);
    bool ended = false;
    for (size_t i = 0; i < ws.size(); ++i) {
        Proof* p = ws[i];
        p->in_flight = false;
        if (err[i] != Z0P_OK || p->phase == PH_DONE) {
            finish(p, err[i]);
            ended = true;
        }
    }
    if (ended)
        active_.erase(std::remove_if(active_.begin(), active_.end(), [](Proof* p) { return p->phase == PH_DONE; }),
                      active_.end());
}

bool Engine::iterate() {
    // Two waves on two streams: the stream of the wave launched before the
    // one in flight is idle (that wave has completed), so proofs are
    // admitted there, ahead of the next wave on that stream, and never
    // behind the wave in flight.
    const bool streams = c_.wave_streams();
    const bool on_idle = streams && !flight_.empty();
    if (on_idle) c_.swap_wave(true);
    admit();
    if (streams) hip_check(hipEventRecord(c_.setup_done, c_.stream));
    if (flight_.empty()) {
        if (active_.empty()) return !queue_.empty();
        c_.ring.reset();  // every launch has completed
        c_.release_graphs();
        if (pipelined_ && c_.pipeline_single && active_.size() == 1 && queue_.empty() &&
            active_[0]->phase == PH_FRONT) {
            Proof* p = active_[0];
            try {
                pipeline(*p);
                if (p->phase == PH_DONE) {
                    finish(p, Z0P_OK);
                    active_.clear();
                }
            } catch (const Status& s) {
                finish(p, s.code);
                active_.clear();
            }
            return !active_.empty() || !queue_.empty();
        }
        if (c_.waves_min == 0 || active_.size() < c_.waves_min) {
            // One wave: every active proof's phase in one launch.
            const double t_build = now_ms();
            const std::vector<Proof*> ws = active_;
            const uint64_t seq = launch(ws);
            c_.prof.host("~host_build_launch", now_ms() - t_build);
            complete(ws, seq, true);
            return !active_.empty() || !queue_.empty();
        }
    }
    // Two waves: the host runs one wave's transcript steps while the device
    // runs the other wave's phase. Launch the proofs that are not in flight
    // (oldest first), stopping near half the active weight so the waves stay
    // even, then complete the wave launched before. Each wave has its own
    // staging buffer and half of the descriptor ring.
    const double t_build = now_ms();
    size_t total = 0;
    for (Proof* p : active_) total += p->mem_bytes;
    std::vector<Proof*> ws;
    size_t took = 0;
    for (Proof* p : active_) {
        if (p->in_flight) continue;
        if (!ws.empty() && 2 * took >= total) break;
        ws.push_back(p);
        took += p->mem_bytes;
    }
    uint64_t seq = 0;
    if (!ws.empty()) {
        if (!on_idle) c_.swap_wave(streams);
        const size_t half = c_.ring.capacity() / 2;
        ring_half_ ^= 1;
        c_.ring.reset_to(ring_half_ * half, (ring_half_ + 1) * half);
        // Proofs admitted on the other stream (and not launched there) have
        // their uploads ordered before this wave.
        if (streams) hip_check(hipStreamWaitEvent(c_.stream, c_.alt_setup_done, 0));
        seq = launch(ws);
    }
    c_.prof.host("~host_build_launch", now_ms() - t_build);
    if (!flight_.empty()) {
        // The earlier wave staged its words in the other buffer (and, with
        // two streams, ran on the other stream).
        if (!ws.empty() || on_idle) c_.swap_wave(streams);
        complete(flight_, flight_seq_, false);
        if (!ws.empty()) c_.swap_wave(streams);
    }
    flight_ = std::move(ws);
    flight_seq_ = seq;
    return !active_.empty() || !queue_.empty();
}

// All phases from FRONT through the first proof-of-work window enqueued at
// once. Phase k > 0 starts behind a device gate that opens when the host
// has run phase k-1's transcript step and written phase k's constants
// (composition alphas, fold challenges, PoW head) to the mailbox.
void Engine::pipeline(Proof& P) {
    struct Ph {
        Phase ph;
        size_t depth;
        size_t stage_off;
    };
    std::vector<Ph> phs{{PH_FRONT, 0, 0}};
    if (P.air || P.row) phs.push_back({PH_COMPOSE, 0, 0});
    for (size_t d = 0; d < P.folds; ++d) phs.push_back({PH_FOLD, d, 0});
    if (P.fp.grinding_bits > 0 && !P.split_commit) phs.push_back({PH_GRIND, 0, 0});
    constexpr size_t SLOT = Ctx::MAILBOX_SLOT;
    // The openings too, when their query list fits the remaining slots; they
    // run only if the first proof-of-work window finds the nonce.
    const size_t open_slots = (8 + 8 * size_t(P.fp.num_queries) + SLOT - 1) / SLOT;
    const bool with_open = !P.split_commit && phs.size() + open_slots <= Ctx::MAILBOX_SLOTS;
    if (with_open) phs.push_back({PH_OPEN, 0, 0});
    require(phs.size() <= Ctx::MAILBOX_SLOTS, Z0P_ERR_DEVICE);
    static_assert(sizeof(ComposeConsts) <= SLOT && sizeof(RangeConsts) <= SLOT && sizeof(TableConsts) <= SLOT &&
                      sizeof(FoldConsts) <= SLOT &&
                      sizeof(GrindHead) <= SLOT,
                  "mailbox slot");
    c_.reserve_stage(P.stage_need * phs.size());
    std::vector<std::vector<Step>> chains(phs.size());
    std::vector<std::vector<CollectItem>> items(phs.size());
    size_t used = 0;
    const double t_build = now_ms();
    auto build_phases = [&](size_t k0, size_t k1) {
        for (size_t k = k0; k < k1; ++k) {
            P.phase = phs[k].ph;
            P.depth = phs[k].depth;
            P.window = 0;
            build(c_, P, chains[k], items[k], used, k == 0 ? nullptr : P.consts + k * SLOT);
            phs[k].stage_off = P.stage_off;
        }
    };
    uint64_t* gate = reinterpret_cast<uint64_t*>(c_.mailbox_host);
    uint32_t* timeout = reinterpret_cast<uint32_t*>(c_.mailbox_host + 8);
    __atomic_store_n(timeout, 0u, __ATOMIC_RELEASE);
    // The openings' slot starts cancelled (live = 0): if a failure opens
    // every gate to drain the enqueued work, k_open then skips instead of
    // reading an earlier proof's queries.
    if (with_open) std::memset(c_.mailbox_host + Ctx::MAILBOX_HEAD + (phs.size() - 1) * SLOT, 0, 8);
    const uint64_t g0 = c_.gate, g_end = g0 + phs.size(), seq0 = c_.seq;
    auto launch_phases = [&](const Launch& L, size_t k0, size_t k1) {
        for (size_t k = k0; k < k1; ++k) {
            // Phase k's collect also gates phase k + 1 (its constants).
            HostGate gate{};
            if (k + 1 < phs.size()) {
                gate.go = reinterpret_cast<const uint64_t*>(c_.mailbox_dev);
                gate.want = g0 + k + 1;
                gate.src = reinterpret_cast<const uint32_t*>(c_.mailbox_dev + Ctx::MAILBOX_HEAD + (k + 1) * SLOT);
                gate.dst = reinterpret_cast<uint32_t*>(P.consts + (k + 1) * SLOT);
                // only the next phase's constants cross PCIe
                size_t bytes = sizeof(GrindHead);  // the proof-of-work head
                switch (phs[k + 1].ph) {
                    case PH_COMPOSE:
                        bytes = P.row == ROW_RANGE ? sizeof(RangeConsts)
                                : P.row == ROW_TABLE ? sizeof(TableConsts)
                                                     : sizeof(ComposeConsts);
                        break;
                    case PH_FOLD: bytes = sizeof(FoldConsts); break;
                    case PH_OPEN: bytes = 8 + 8 * size_t(P.fp.num_queries); break;
                    default: break;
                }
                gate.words = uint32_t((bytes + 3) / 4);
                gate.spin_limit = c_.gate_spin;
                gate.timeout = reinterpret_cast<uint32_t*>(c_.mailbox_dev + 8);
            }
            const std::vector<const std::vector<Step>*> one{&chains[k]};
            launch_chains(L, one, items[k], c_.flag_dev, seq0 + 1 + k, gate);
        }
    };
    // The first phase goes to the device before the later ones are built:
    // the device works on it while the host builds and launches the rest.
    auto drain = [&] {
        // Open every remaining gate so the enqueued work drains.
        __atomic_store_n(gate, g_end, __ATOMIC_RELEASE);
        (void)hipStreamSynchronize(c_.stream);
        c_.gate = g_end;
        c_.ring.reset();
    };
    build_phases(0, 1);
    c_.enqueue([&](const Launch& L) { launch_phases(L, 0, 1); });
    try {
        build_phases(1, phs.size());
        P.phase = PH_FRONT;
        P.depth = 0;
        if (phs.size() > 1) c_.enqueue([&](const Launch& L) { launch_phases(L, 1, phs.size()); });
    } catch (...) {
        drain();
        throw;
    }
    c_.seq = seq0 + phs.size();
    c_.prof.host("~host_build_launch", now_ms() - t_build);
    try {
        bool cancelled = false;
        for (size_t k = 0; k < phs.size(); ++k) {
            const double t_wait = now_ms();
            c_.wait_flag(seq0 + 1 + k, false);
            if (__atomic_load_n(timeout, __ATOMIC_ACQUIRE)) throw Status(Z0P_ERR_DEVICE);
            const double t_host = now_ms();
            if (cancelled) break;  // the openings drained without running
            const Phase resume = P.phase;
            P.phase = phs[k].ph;
            P.depth = phs[k].depth;
            P.stage_off = phs[k].stage_off;
            if (phs[k].ph == PH_OPEN && resume != PH_OPEN) P.phase = resume;
            else host_step(c_, P);
            if (k + 1 < phs.size()) {
                const Ph& nx = phs[k + 1];
                uint8_t* slot = c_.mailbox_host + Ctx::MAILBOX_HEAD + (k + 1) * SLOT;
                if (nx.ph == PH_OPEN) {
                    // Nonce found: hand over the queries; otherwise cancel the
                    // openings and continue the search stepwise.
                    // The query list lands before `live` is published, so a
                    // slot that reads live = 1 always holds this proof's list.
                    const uint64_t live = P.phase == PH_OPEN ? 1 : 0;
                    if (live) {
                        require(P.queries.size() == P.fp.num_queries, Z0P_ERR_DEVICE);
                        std::memcpy(slot + 8, P.queries.data(), 8 * P.queries.size());
                    }
                    __atomic_store_n(reinterpret_cast<uint64_t*>(slot), live, __ATOMIC_RELEASE);
                    cancelled = !live;
                } else {
                    require(P.phase == nx.ph && (nx.ph != PH_FOLD || P.depth == nx.depth), Z0P_ERR_DEVICE);
                    if (nx.ph == PH_COMPOSE && P.row == ROW_RANGE) std::memcpy(slot, &P.rc, sizeof(P.rc));
                    else if (nx.ph == PH_COMPOSE && P.row == ROW_TABLE) std::memcpy(slot, &P.tc, sizeof(P.tc));
                    else if (nx.ph == PH_COMPOSE) std::memcpy(slot, &P.cc, sizeof(P.cc));
                    else if (nx.ph == PH_FOLD) std::memcpy(slot, &P.fold, sizeof(P.fold));
                    else std::memcpy(slot, &P.pow, sizeof(P.pow));
                }
                __atomic_store_n(gate, g0 + k + 1, __ATOMIC_RELEASE);
            }
            c_.prof.host("~host_wait_device", t_host - t_wait);
            c_.prof.host("~host_transcript", now_ms() - t_host);
        }
        // The per-launch events that have completed; the rest resolve when
        // the statistics are read (no wait on the proof's critical path).
        if (c_.prof.enabled) c_.prof.resolve_ready();
    } catch (...) {
        drain();
        throw;
    }
    c_.gate = g_end;
}

void Engine::fail_all(int status) {
    // Nothing enqueued (a wave in flight, a launch cut short) may still use
    // the memory the proofs hand back.
    (void)hipStreamSynchronize(c_.alt_stream);
    (void)hipStreamSynchronize(c_.stream);
    flight_.clear();
    for (Proof* p : active_) {
        p->in_flight = false;
        finish(p, status);
    }
    active_.clear();
    while (!queue_.empty()) {
        finish(queue_.front(), status);
        queue_.pop_front();
    }
}

void Engine::run() {
    try {
        while (iterate()) {
        }
    } catch (const Status& s) {
        // Device failure: every unfinished proof fails with it.
        fail_all(s.code);
        throw;
    }
}

// ---------------------------------------------------------------------------
// Entry points.

Proof* proof_new_air(Ctx& c, const z0p_air_params& p, Transcript* tr) {
    auto P = std::make_unique<Proof>();
    init_air(c, *P, p, tr);
    return P.release();
}

Proof* proof_new_air_trace(Ctx& c, const z0p_air_params& p, const uint32_t* col_a, uint64_t len_a,
                          const uint32_t* col_b, uint64_t len_b, Transcript* tr) {
    auto P = std::make_unique<Proof>();
    init_air_trace(c, *P, p, col_a, len_a, col_b, len_b, tr);
    return P.release();
}

Proof* proof_new_air_zk(Ctx& c, const z0p_air_params& p, uint64_t n_rand, ChaCha20& rng, Transcript* tr) {
    auto P = std::make_unique<Proof>();
    init_air_zk(c, *P, p, n_rand, rng, tr);
    return P.release();
}

Proof* proof_new_range(Ctx& c, const z0p_range_params& p, Transcript* tr) {
    auto P = std::make_unique<Proof>();
    init_range(c, *P, p, tr);
    return P.release();
}

Proof* proof_new_range_cols(Ctx& c, const z0p_range_params& p, const uint32_t* cols, uint64_t ncols, uint64_t len,
                            Transcript* tr) {
    auto P = std::make_unique<Proof>();
    init_range_cols(c, *P, p, cols, ncols, len, tr);
    return P.release();
}

Proof* proof_new_range_zk(Ctx& c, const z0p_range_params& p, uint64_t n_rand, ChaCha20& rng, Transcript* tr) {
    auto P = std::make_unique<Proof>();
    init_range_zk(c, *P, p, n_rand, rng, tr);
    return P.release();
}

Proof* proof_new_table(Ctx& c, const z0p_table_params& p, const uint32_t* table, uint64_t table_len,
                       Transcript* tr) {
    auto P = std::make_unique<Proof>();
    init_table(c, *P, p, table, table_len, tr);
    return P.release();
}

Proof* proof_new_table_zk(Ctx& c, const z0p_table_params& p, const uint32_t* table, uint64_t table_len,
                          uint64_t n_rand, ChaCha20& rng, Transcript* tr) {
    auto P = std::make_unique<Proof>();
    init_table_zk(c, *P, p, table, table_len, n_rand, rng, tr);
    return P.release();
}

Proof* proof_new_fri(Ctx& c, const uint32_t* src, uint64_t count, int kind, uint64_t degree_bound,
                     const z0p_fri_params& fp, Transcript* tr) {
    auto P = std::make_unique<Proof>();
    init_fri(c, *P, src, count, kind, degree_bound, fp, tr);
    return P.release();
}

Proof* proof_new_fold_fri(Ctx& c, const FoldStepDesc* steps, uint32_t nsteps, uint64_t seg_steps, uint64_t blowup,
                          const z0p_fri_params& fp, Transcript* tr) {
    auto P = std::make_unique<Proof>();
    init_fold_fri(c, *P, steps, nsteps, seg_steps, blowup, fp, tr);
    return P.release();
}

const std::vector<uint64_t>& proof_queries(const Proof* p) { return p->queries; }
void proof_free(Proof* p) { delete p; }
int proof_status(const Proof* p) { return p->status; }
bool proof_done(const Proof* p) { return p->phase == PH_DONE && p->status == Z0P_OK; }
Bytes& proof_bytes(Proof* p) { return p->proof; }
Bytes& proof_commitment(Proof* p) { return p->commitment; }
size_t proof_mem_bytes(const Proof* p) { return p->mem_bytes; }
void proof_times(const Proof* p, double& submit, double& admit, double& done) {
    submit = p->t_submit;
    admit = p->t_admit;
    done = p->t_done;
}
void proof_set_tag(Proof* p, uint64_t tag) { p->tag = tag; }
uint64_t proof_tag(const Proof* p) { return p->tag; }

void run_proof(Ctx& ctx, Proof* p) {
    Engine e(ctx, 0, 1);
    e.set_pipelined(true);
    e.submit(p);
    e.run();
    if (p->status != Z0P_OK) throw Status(p->status);
}

namespace {
Bytes run_one(Ctx& ctx, Proof* raw) {
    std::unique_ptr<Proof> P(raw);
    run_proof(ctx, P.get());
    return std::move(P->proof);
}
}  // namespace

Bytes air_prove(Ctx& ctx, const z0p_air_params& p, Transcript& tr) {
    return run_one(ctx, proof_new_air(ctx, p, &tr));
}

Bytes air_prove_with_trace(Ctx& ctx, const z0p_air_params& p, const uint32_t* col_a, uint64_t len_a,
                                          const uint32_t* col_b, uint64_t len_b, Transcript& tr) {
    return run_one(ctx, proof_new_air_trace(ctx, p, col_a, len_a, col_b, len_b, &tr));
}

Bytes air_prove_zk(Ctx& ctx, const z0p_air_params& p, uint64_t n_rand, ChaCha20& rng, Transcript& tr) {
    return run_one(ctx, proof_new_air_zk(ctx, p, n_rand, rng, &tr));
}

Bytes range_prove(Ctx& ctx, const z0p_range_params& p, Transcript& tr) {
    return run_one(ctx, proof_new_range(ctx, p, &tr));
}

Bytes range_prove_with_columns(Ctx& ctx, const z0p_range_params& p, const uint32_t* cols, uint64_t ncols,
                               uint64_t len, Transcript& tr) {
    return run_one(ctx, proof_new_range_cols(ctx, p, cols, ncols, len, &tr));
}

Bytes range_prove_zk(Ctx& ctx, const z0p_range_params& p, uint64_t n_rand, ChaCha20& rng, Transcript& tr) {
    return run_one(ctx, proof_new_range_zk(ctx, p, n_rand, rng, &tr));
}

Bytes table_prove(Ctx& ctx, const z0p_table_params& p, const uint32_t* table, uint64_t table_len, Transcript& tr) {
    return run_one(ctx, proof_new_table(ctx, p, table, table_len, &tr));
}

Bytes table_prove_zk(Ctx& ctx, const z0p_table_params& p, const uint32_t* table, uint64_t table_len, uint64_t n_rand,
                     ChaCha20& rng, Transcript& tr) {
    return run_one(ctx, proof_new_table_zk(ctx, p, table, table_len, n_rand, rng, &tr));
}

void fri_commit_open(Ctx& ctx, const uint32_t* src, uint64_t count, int kind, uint64_t degree_bound,
                     const z0p_fri_params& fp, Transcript& tr, Bytes& commitment, Bytes& proof,
                     std::vector<uint64_t>& indices) {
    std::unique_ptr<Proof> P(proof_new_fri(ctx, src, count, kind, degree_bound, fp, &tr));
    Engine e(ctx, 0, 1);
    e.set_pipelined(true);
    e.submit(P.get());
    e.run();
    if (P->status != Z0P_OK) throw Status(P->status);
    commitment = std::move(P->commitment);
    proof = std::move(P->proof);
    indices = P->queries;
}

// ---------------------------------------------------------------------------
// fri::commit / fri::open as two calls (fri.rs:408-666): the state keeps its
// layers and trees in device memory of its own between them.

FriState* fri_commit(Ctx& ctx, const uint32_t* src, uint64_t count, int kind, uint64_t degree_bound,
                     const z0p_fri_params& fp, Transcript& tr) {
    std::unique_ptr<Proof> P(proof_new_fri(ctx, src, count, kind, degree_bound, fp, &tr));
    P->split_commit = true;
    hip_check(hipMalloc(&P->mem, P->mem_bytes));
    P->own_mem = true;
    try {
        run_proof(ctx, P.get());
    } catch (...) {
        (void)hipStreamSynchronize(ctx.stream);
        (void)hipFree(P->mem);
        P->mem = nullptr;
        throw;
    }
    P->tr = nullptr;
    auto* s = new FriState;
    s->ctx = &ctx;
    s->proof = P.release();
    return s;
}

const Bytes& fri_state_commitment(const FriState& s) { return s.proof->commitment; }
uint64_t fri_state_queries(const FriState& s) { return s.proof->fp.num_queries; }

Bytes fri_open(FriState& s, Transcript& tr, std::vector<uint64_t>& indices) {
    Proof& P = *s.proof;
    P.tr = &tr;
    P.resume = true;
    P.status = Z0P_OK;
    P.proof = Bytes();
    P.queries.clear();
    begin_open(P);
    Engine e(*s.ctx, 0, 1);
    e.submit(&P);
    e.run();
    P.tr = nullptr;
    if (P.status != Z0P_OK) throw Status(P.status);
    indices = P.queries;
    return std::move(P.proof);
}

void fri_state_free(FriState* s) {
    if (!s) return;
    if (s->proof) {
        (void)hipStreamSynchronize(s->ctx->stream);
        if (s->proof->mem) (void)hipFree(s->proof->mem);
        delete s->proof;
    }
    delete s;
}

// ---------------------------------------------------------------------------
// Transcript::grind (transcript.rs:206-228) on the device.

uint64_t grind_search(Ctx& c, uint8_t* dev, const GrindHead& h, GrindJob j, uint32_t bits, bool profile) {
    GrindHead* dh = reinterpret_cast<GrindHead*>(dev);
    auto* res = reinterpret_cast<unsigned long long*>(dev + GRIND_BUF_RESULT);
    hip_check(hipMemcpyAsync(dh, &h, sizeof(h), hipMemcpyHostToDevice, c.stream));
    j.head = dh;
    j.bits = bits;
    j.nblocks = uint32_t(std::min(2048, std::max(16, c.sched.grind_blocks)));
    j.result = res;
    uint64_t found = ~uint64_t(0);
    for (uint64_t window = 0; found == ~uint64_t(0); ++window) {
        // Window w scans [2^(20+4w), 2^(24+4w)) after the windows before it.
        j.start = window == 0 ? 0 : uint64_t(1) << std::min<uint64_t>(63, 20 + 4 * window);
        j.limit = uint64_t(1) << std::min<uint64_t>(63, 24 + 4 * window);
        Launch L = c.launch();
        if (!profile) L.prof = nullptr;  // no event pairs
        launch_grind(L, {j});
        hip_check(hipMemcpyAsync(&found, res, 8, hipMemcpyDeviceToHost, c.stream));
        hip_check(hipStreamSynchronize(c.stream));
        if (profile && c.prof.enabled) c.prof.resolve();
    }
    return found;
}

uint64_t transcript_grind(Ctx& c, Transcript& tr, const char* label, uint32_t bits, uint8_t* dev) {
    uint8_t nb[8];
    if (bits > 32) {
        // The protocol caps difficulty at 32: fail closed with nonce 0.
        store_le64(nb, 0);
        tr.absorb(label, nb, 8);
        return 0;
    }
    std::vector<uint8_t> head = tr.pow_head(label);
    GrindHead h;
    GrindJob j{};
    require(grind_prepare(head.data(), head.size(), h, j), Z0P_ERR_MALFORMED_INPUT);
    const bool own = dev == nullptr;
    if (own) {
        if (!c.grind_buf) hip_check(hipMalloc(&c.grind_buf, GRIND_BUF_BYTES));
        dev = c.grind_buf;
    }
    const uint64_t nonce = grind_search(c, dev, h, j, bits, own);
    store_le64(nb, nonce);
    tr.absorb(label, nb, 8);
    return nonce;
}

}  // namespace z0p
