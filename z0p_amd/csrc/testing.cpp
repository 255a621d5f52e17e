// Kernel-level diagnostic entry points (include/z0p_testing.h): each runs
// one device stage on host data so a parity failure can be pinned to one
// kernel. Natural-order inputs are permuted to the device's bit-reversed
// layout on the host and back.
#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/z0p_testing.h"
#include "params.hpp"
#include "prover.hpp"


namespace z0p {
namespace {

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return Z0P_OK;
    } catch (const Status& s) {
        return s.code;
    } catch (...) {
        return Z0P_ERR_DEVICE;
    }
}

template <typename T>
T* to_dev(Ctx& c, const T* host, size_t count) {
    T* d = c.scratch_alloc<T>(count);
    hip_check(hipMemcpyAsync(d, host, count * sizeof(T), hipMemcpyHostToDevice, c.stream));
    return d;
}
template <typename T>
void to_host(Ctx& c, T* host, const T* dev, size_t count) {
    hip_check(hipMemcpyAsync(host, dev, count * sizeof(T), hipMemcpyDeviceToHost, c.stream));
    c.sync();
}

// Zeroed staged-top counters for a tree of geometry g (scratch memory).
uint32_t* zeroed_counters(Ctx& c, const TreeGeom& g) {
    const uint64_t n = std::max<uint64_t>(1, tree_stage_counters(g.leaves, 0, g.target));
    uint32_t* d = c.scratch_alloc<uint32_t>(n);
    hip_check(hipMemsetAsync(d, 0, n * 4, c.stream));
    return d;
}

// Runs one chain of steps (a single job per launch: no tables).
void run_chain(Ctx& c, const std::vector<Step>& chain) {
    std::vector<const std::vector<Step>*> chains{&chain};
    c.enqueue([&](const Launch& L) { launch_chains(L, chains, {}); });
}

// natural index i = j + B t  <->  position brev(j) n + brev(t)
uint64_t br_pos(uint64_t i, int logn, int logb) {
    uint64_t B = uint64_t(1) << logb;
    return (uint64_t(brev_host(i & (B - 1), logb)) << logn) | brev_host(i >> logb, logn);
}

void check_geometry(size_t n, size_t blowup) {
    require(pow2(n) && n >= 2 && pow2(blowup) && blowup >= 2 && n * blowup <= MAX_DOMAIN, Z0P_ERR_INVALID_DOMAIN_SIZE);
}

}  // namespace
}  // namespace z0p

using namespace z0p;

extern "C" {

int z0p_test_intt(z0p_ctx* ctx, const uint32_t* evals, size_t n, uint32_t* coeffs_br) {
    if (!ctx || !evals || !coeffs_br) return Z0P_ERR_INVALID_ARGUMENT;
    return guarded([&] {
        check_geometry(n, 16);
        Ctx& c = *ctx->c;
        int logn = int(ilog2(n));
        const GeometryTables& gt = c.geometry(logn, 4);
        c.reserve_scratch(n * 12 + (1 << 20));
        uint32_t* din = to_dev(c, evals, n);
        uint32_t* dout = c.scratch_alloc<uint32_t>(n);
        uint32_t* scratch = c.scratch_alloc<uint32_t>(n);
        NttPlan plan = plan_intt(din, dout, scratch, logn, c.tw4096_inv, gt.winv, 4, to_mont(inv_c(uint32_t(n % P))));
        std::vector<Step> chain(plan.pass1.begin(), plan.pass1.end());
        chain.insert(chain.end(), plan.pass2.begin(), plan.pass2.end());
        run_chain(c, chain);
        to_host(c, coeffs_br, dout, n);
    });
}

int z0p_test_lde(z0p_ctx* ctx, const uint32_t* coeffs, size_t n, size_t blowup, uint32_t* lde_br) {
    if (!ctx || !coeffs || !lde_br) return Z0P_ERR_INVALID_ARGUMENT;
    return guarded([&] {
        check_geometry(n, blowup);
        Ctx& c = *ctx->c;
        int logn = int(ilog2(n)), logb = int(ilog2(blowup));
        size_t m = n * blowup;
        const GeometryTables& gt = c.geometry(logn, logb);
        c.reserve_scratch(n * 8 + m * 8 + (1 << 20));
        uint32_t* din = to_dev(c, coeffs, n);
        uint32_t* dbr = c.scratch_alloc<uint32_t>(n);
        uint32_t* scratch = c.scratch_alloc<uint32_t>(m);
        uint32_t* dout = c.scratch_alloc<uint32_t>(m);
        NttPlan plan = plan_lde(dbr, dout, scratch, logn, logb, c.tw4096, gt.w, gt.coset);
        std::vector<Step> chain{BitrevJob{din, dbr, logn, 0}};
        chain.insert(chain.end(), plan.pass1.begin(), plan.pass1.end());
        chain.insert(chain.end(), plan.pass2.begin(), plan.pass2.end());
        run_chain(c, chain);
        to_host(c, lde_br, dout, m);
    });
}

int z0p_test_merkle64(z0p_ctx* ctx, const uint8_t* flat, size_t leaves, uint8_t* levels) {
    if (!ctx || !flat || !levels) return Z0P_ERR_INVALID_ARGUMENT;
    return guarded([&] {
        require(pow2(leaves), Z0P_ERR_INVALID_DOMAIN_SIZE);
        Ctx& c = *ctx->c;
        TreeGeom g = tree_geom(leaves, 0);
        c.reserve_scratch(leaves * 64 + g.words() * 4 + (1 << 20));
        const uint32_t* din = to_dev(c, reinterpret_cast<const uint32_t*>(flat), leaves * 16);
        uint32_t* tree = c.scratch_alloc<uint32_t>(g.words());
        LeavesJob lj{};
        lj.kind = LEAF_RAW;
        lj.src = din;
        std::vector<Step> chain;
        tree_steps(chain, tree, g, lj, c.sched, zeroed_counters(c, g));
        run_chain(c, chain);
        to_host(c, reinterpret_cast<uint32_t*>(levels), tree, g.words());
    });
}

int z0p_test_trace_tree(z0p_ctx* ctx, const uint32_t* a, const uint32_t* b, size_t n, size_t blowup,
                        uint8_t* levels) {
    if (!ctx || !a || !b || !levels) return Z0P_ERR_INVALID_ARGUMENT;
    return guarded([&] {
        check_geometry(n, blowup);
        require(blowup >= 4, Z0P_ERR_INVALID_BLOWUP);
        Ctx& c = *ctx->c;
        int logn = int(ilog2(n)), logb = int(ilog2(blowup));
        size_t m = n * blowup;
        std::vector<uint32_t> abr(m), bbr(m);
        for (size_t i = 0; i < m; ++i) {
            uint64_t p = br_pos(i, logn, logb);
            abr[p] = a[i];
            bbr[p] = b[i];
        }
        TreeGeom g = tree_geom(m / 8, 0);
        c.reserve_scratch(m * 8 + g.words() * 4 + (1 << 20));
        uint32_t* da = to_dev(c, abr.data(), m);
        uint32_t* db = to_dev(c, bbr.data(), m);
        uint32_t* tree = c.scratch_alloc<uint32_t>(g.words());
        LeavesJob lj{};
        lj.kind = LEAF_TRACE;
        lj.src = da;
        lj.src_b = db;
        lj.logn = logn;
        lj.logb = logb;
        std::vector<Step> chain;
        tree_steps(chain, tree, g, lj, c.sched, zeroed_counters(c, g));
        run_chain(c, chain);
        to_host(c, reinterpret_cast<uint32_t*>(levels), tree, g.words());
    });
}

int z0p_test_linked(z0p_ctx* ctx, const uint32_t* a, size_t n, size_t blowup, uint32_t actual_b_last, uint32_t a0,
                    uint32_t* b) {
    if (!ctx || !a || !b) return Z0P_ERR_INVALID_ARGUMENT;
    return guarded([&] {
        check_geometry(n, blowup);
        Ctx& c = *ctx->c;
        int logn = int(ilog2(n)), logb = int(ilog2(blowup));
        size_t m = n * blowup;
        const GeometryTables& gt = c.geometry(logn, logb);
        std::vector<uint32_t> abr(m), bbr(m);
        for (size_t i = 0; i < m; ++i) abr[br_pos(i, logn, logb)] = a[i];
        const uint32_t w_last = pow_c(root_of_unity(uint32_t(logn)), n - 1);
        const uint32_t coeff = mul(mul(sub(actual_b_last % P, a0 % P), w_last), inv_c(uint32_t(n % P)));
        c.reserve_scratch(m * 8 + (1 << 20));
        uint32_t* da = to_dev(c, abr.data(), m);
        uint32_t* db = c.scratch_alloc<uint32_t>(m);
        LinkedJob lj{};
        lj.a = da;
        lj.b = db;
        lj.iw = c.point_inverses(logn, logb).iw;
        lj.zvan_m = gt.zvan;
        lj.coeff = coeff;
        lj.logn = logn;
        lj.logb = logb;
        run_chain(c, {lj});
        to_host(c, bbr.data(), db, m);
        for (size_t i = 0; i < m; ++i) b[i] = bbr[br_pos(i, logn, logb)];
    });
}

int z0p_test_compose(z0p_ctx* ctx, const uint32_t* a, const uint32_t* b, size_t n, size_t blowup, uint32_t a0,
                     uint32_t b0, uint32_t b_last, const uint32_t* alphas, uint32_t* out, uint8_t* levels01) {
    if (!ctx || !a || !b || !alphas || !out) return Z0P_ERR_INVALID_ARGUMENT;
    return guarded([&] {
        check_geometry(n, blowup);
        Ctx& c = *ctx->c;
        int logn = int(ilog2(n)), logb = int(ilog2(blowup));
        size_t m = n * blowup;
        std::vector<uint32_t> abr(m), bbr(m), o(4 * m);
        for (size_t i = 0; i < m; ++i) {
            uint64_t p = br_pos(i, logn, logb);
            abr[p] = a[i];
            bbr[p] = b[i];
        }
        ComposeConsts cc{};
        for (int k = 0; k < 5; ++k)
            for (int q = 0; q < 4; ++q) cc.alpha_m[k][q] = to_mont(alphas[4 * k + q] % P);
        cc.w_last_m = to_mont(pow_c(root_of_unity(uint32_t(logn)), n - 1));
        cc.a0 = a0 % P;
        cc.b0 = b0 % P;
        cc.b_last = b_last % P;
        c.reserve_scratch(m * 36 + (1 << 20));
        uint32_t* da = to_dev(c, abr.data(), m);
        uint32_t* db = to_dev(c, bbr.data(), m);
        uint32_t* dl = c.scratch_alloc<uint32_t>(4 * m);
        ComposeJob cj{};
        cj.a = da;
        cj.b = db;
        cj.layer0 = dl;
        cj.pinv = c.point_inverses(logn, logb);
        cj.c = to_dev(c, &cc, 1);
        cj.logn = logn;
        cj.logb = logb;
        uint32_t* tree = nullptr;
        if (levels01) {
            tree = c.scratch_alloc<uint32_t>(8 * (m / 4 + m / 8));
            cj.tree = tree;
            cj.with_parent = 1;
        }
        run_chain(c, {cj});
        to_host(c, o.data(), dl, 4 * m);
        for (size_t i = 0; i < m; ++i) std::memcpy(out + 4 * i, &o[4 * br_pos(i, logn, logb)], 16);
        if (levels01) to_host(c, reinterpret_cast<uint32_t*>(levels01), tree, 8 * (m / 4 + m / 8));
    });
}

int z0p_test_fold(z0p_ctx* ctx, const uint32_t* layer, size_t len, size_t lde_len, size_t depth,
                  const uint32_t* challenges, uint32_t* out, uint8_t* levels01) {
    if (!ctx || !layer || !challenges || !out) return Z0P_ERR_INVALID_ARGUMENT;
    return guarded([&] {
        require(pow2(len) && len >= 4 && pow2(lde_len) && lde_len <= MAX_DOMAIN && (len << (2 * depth)) == lde_len,
                Z0P_ERR_INVALID_DOMAIN_SIZE);
        require(lde_len >= 16, Z0P_ERR_INVALID_DOMAIN_SIZE);
        Ctx& c = *ctx->c;
        int logm = int(ilog2(lde_len));
        const GeometryTables& gt = c.geometry(logm - 4, 4);
        E4 r[3];
        for (int k = 0; k < 3; ++k)
            for (int q = 0; q < 4; ++q) r[k].v[q] = challenges[4 * k + q] % P;
        const FoldConsts fc = make_fold_consts(r, int(ilog2(len)), int(depth));
        c.reserve_scratch(len * 24 + (1 << 20));
        uint32_t* dl = to_dev(c, layer, 4 * len);
        uint32_t* dn = c.scratch_alloc<uint32_t>(len);
        FoldJob fj{};
        fj.layer = dl;
        fj.next = dn;
        fj.finv = gt.finv;
        fj.c = to_dev(c, &fc, 1);
        fj.log_len = fc.log_len;
        fj.base_field = 0;
        uint32_t* tree = nullptr;
        const size_t leaves = len / 16;  // quads of the len/4 outputs
        if (levels01) {
            require(len >= 32, Z0P_ERR_INVALID_DOMAIN_SIZE);
            tree = c.scratch_alloc<uint32_t>(8 * (leaves + leaves / 2));
            fj.tree = tree;
            fj.with_parent = 1;
        }
        run_chain(c, {fj});
        to_host(c, out, dn, len);
        if (levels01) to_host(c, reinterpret_cast<uint32_t*>(levels01), tree, 8 * (leaves + leaves / 2));
    });
}

int z0p_test_copy(z0p_ctx* ctx, size_t bytes, float* ms) {
    if (!ctx || !ms) return Z0P_ERR_INVALID_ARGUMENT;
    return guarded([&] {
        require(bytes % 16 == 0 && bytes > 0, Z0P_ERR_MALFORMED_INPUT);
        Ctx& c = *ctx->c;
        c.reserve_scratch(2 * bytes + (1 << 20));
        uint32_t* a = c.scratch_alloc<uint32_t>(bytes / 4);
        uint32_t* b = c.scratch_alloc<uint32_t>(bytes / 4);
        hip_check(hipMemsetAsync(a, 1, bytes, c.stream));
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0));
        hip_check(hipEventCreate(&e1));
        hip_check(hipEventRecord(e0, c.stream));
        launch_copy(c.launch(), a, b, bytes / 4);
        hip_check(hipEventRecord(e1, c.stream));
        c.sync();
        hip_check(hipEventElapsedTime(ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    });
}

int z0p_test_set_gate_spin(z0p_ctx* ctx, uint32_t polls) {
    if (!ctx) return Z0P_ERR_INVALID_ARGUMENT;
    ctx->c->gate_spin = polls;
    return Z0P_OK;
}

int z0p_test_grind(z0p_ctx* ctx, const uint8_t* head, size_t head_len, uint32_t bits, uint64_t* nonce) {
    if (!ctx || !head || !nonce) return Z0P_ERR_INVALID_ARGUMENT;
    return guarded([&] {
        require(bits <= 32, Z0P_ERR_MALFORMED_INPUT);
        Ctx& c = *ctx->c;
        GrindHead h;
        GrindJob g{};
        require(grind_prepare(head, head_len, h, g), Z0P_ERR_MALFORMED_INPUT);
        c.reserve_scratch(1 << 20);
        g.head = to_dev(c, &h, 1);
        g.bits = bits;
        g.nblocks = 2048;
        g.result = c.scratch_alloc<unsigned long long>(1);
        uint64_t found = ~uint64_t(0);
        for (uint64_t window = 0; found == ~uint64_t(0); ++window) {
            g.start = window == 0 ? 0 : uint64_t(1) << std::min<uint64_t>(63, 20 + 4 * window);
            g.limit = uint64_t(1) << std::min<uint64_t>(63, 24 + 4 * window);
            run_chain(c, {g});
            to_host(c, &found, reinterpret_cast<const uint64_t*>(g.result), 1);
        }
        *nonce = found;
    });
}

}  // extern "C"
