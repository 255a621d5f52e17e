// IVC accumulation over Fibonacci segments (fold.rs:361-570) on the device.
//
// add_segment commits one segment the way the reference does: both columns
// interpolated and extended on the 31-coset, one Merkle leaf per LDE row
// (BLAKE3(0x00 || A || B)), the cap absorbed under "fold/trace" and six
// alphas drawn. The segment's LDE columns and tree stay in device memory;
// its composition is not materialised. finalize draws one gamma per segment
// and runs one FRI (the scheduler's FRI_FOLD_COMBINE proof) whose layer 0 is
// sum_k gamma_k C_k, computed by one kernel from every segment's columns;
// then one kernel gathers every segment's openings at the FRI queries.
#include <array>
#include <cstring>
#include <memory>
#include <vector>

#include "params.hpp"
#include "prover.hpp"

namespace z0p {

namespace {

constexpr int FOLD_ALPHAS = 6;  // fold.rs:38
using Digest = std::array<uint8_t, 32>;

// fold.rs:729-741: 56 bytes, grinding_bits as u32.
void absorb_fold_params(const z0p_fold_params& p, Transcript& tr) {
    uint8_t pb[56];
    store_le64(pb, p.seg_steps);
    store_le64(pb + 8, p.blowup);
    store_le64(pb + 16, p.num_queries);
    store_le64(pb + 24, p.final_size);
    store_le32(pb + 32, p.grinding_bits);
    store_le64(pb + 36, p.cap_height);
    store_le32(pb + 44, p.in_a % P);
    store_le32(pb + 48, p.in_b % P);
    store_le32(pb + 52, p.out_b % P);
    tr.absorb("fold/params", pb, 56);
}

// fold.rs:743-751
void absorb_fold_step(const uint32_t c[4], Transcript& tr) {
    uint8_t b[16];
    for (int i = 0; i < 4; ++i) store_le32(b + 4 * i, c[i]);
    tr.absorb("fold/step", b, 16);
}

// A device buffer from the context's spare list, returned there on release
// once the context's stream has drained (an error path may leave work queued).
struct DevBuf {
    Ctx& c;
    size_t bytes;
    void* p = nullptr;
    DevBuf(Ctx& ctx, size_t n) : c(ctx), bytes(n ? n : 4) { p = c.take_buffer(bytes); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        (void)hipStreamSynchronize(c.stream);
        c.give_buffer(bytes, p);
    }
    template <typename T>
    T* as() const { return static_cast<T*>(p); }
};

}  // namespace

struct FoldAcc {
    struct Segment {
        uint32_t claims[4];  // in_a, in_b, out_a, out_b
        E4 alphas[FOLD_ALPHAS];
        std::vector<Digest> cap;
        std::unique_ptr<DevBuf> mem;  // A, B (m words each), then the row tree
    };
    Ctx& c;
    z0p_fold_params p;
    std::vector<Segment> segs;
    // The segments' evaluations (2s), coefficients (2s) and NTT scratch (m),
    // kept across add_segment calls.
    std::unique_ptr<DevBuf> work;
    FoldAcc(Ctx& ctx, const z0p_fold_params& params) : c(ctx), p(params) {}

    uint64_t m() const { return p.seg_steps * p.blowup; }

    // fold.rs:407-492
    void add_segment(const uint32_t* col_a, uint64_t len_a, const uint32_t* col_b, uint64_t len_b, Transcript& tr) {
        validate_fold(p);
        const uint64_t s = p.seg_steps;
        require(len_a == s && len_b == s, Z0P_ERR_MALFORMED_INPUT);
        for (uint64_t i = 0; i < s; ++i) require(col_a[i] < P && col_b[i] < P, Z0P_ERR_MALFORMED_INPUT);
        Segment g;
        g.claims[0] = col_a[0];
        g.claims[1] = col_b[0];
        g.claims[2] = col_a[s - 1];
        g.claims[3] = col_b[s - 1];
        const int logn = int(ilog2(s)), logb = int(ilog2(p.blowup));
        const uint64_t mm = m();
        const TreeGeom tg = tree_geom(mm, p.cap_height);
        const GeometryTables& gt = c.geometry(logn, logb);
        g.mem = std::make_unique<DevBuf>(c, (2 * mm + tg.words()) * 4);
        uint32_t* A = g.mem->as<uint32_t>();
        uint32_t* tree = A + 2 * mm;
        if (!work) work = std::make_unique<DevBuf>(c, (4 * s + mm) * 4);
        uint32_t *evals = work->as<uint32_t>(), *coeff = evals + 2 * s, *scratch = coeff + 2 * s;
        hip_check(hipMemcpyAsync(evals, col_a, s * 4, hipMemcpyHostToDevice, c.stream));
        hip_check(hipMemcpyAsync(evals + s, col_b, s * 4, hipMemcpyHostToDevice, c.stream));
        std::vector<Step> chain;
        const uint32_t ninv = to_mont(inv_c(uint32_t(s % P)));
        for (int k = 0; k < 2; ++k) {
            NttPlan ip = plan_intt(evals + k * s, coeff + k * s, scratch, logn, c.tw4096_inv, gt.winv, logb, ninv);
            for (auto& j : ip.pass1) chain.push_back(j);
            for (auto& j : ip.pass2) chain.push_back(j);
            NttPlan lp = plan_lde(coeff + k * s, A + k * mm, scratch, logn, logb, c.tw4096, gt.w, gt.coset);
            for (auto& j : lp.pass1) chain.push_back(j);
            for (auto& j : lp.pass2) chain.push_back(j);
        }
        LeavesJob lj{};
        lj.kind = LEAF_ROWS;
        lj.src = A;
        lj.ncols = 2;
        lj.col_stride = mm;
        lj.logn = logn;
        lj.logb = logb;
        tree_steps(chain, tree, tg, lj, c.sched);
        const uint32_t cap_words = uint32_t(8u << tg.drop);
        c.reserve_stage(cap_words);
        std::vector<CollectItem> items{CollectItem{tree + tg.level_offset_words(tg.target), c.stage_dev, cap_words, 0}};
        const std::vector<const std::vector<Step>*> chains{&chain};
        const uint64_t seq = ++c.seq;
        c.enqueue([&](const Launch& L) { launch_chains(L, chains, items, c.flag_dev, seq); });
        c.wait_flag(seq);
        g.cap.resize(size_t(1) << tg.drop);
        for (size_t i = 0; i < g.cap.size(); ++i) std::memcpy(g.cap[i].data(), c.stage_host + 8 * i, 32);
        // transcript (fold.rs:418-478)
        if (segs.empty()) absorb_fold_params(p, tr);
        absorb_fold_step(g.claims, tr);
        std::vector<uint8_t> cb(g.cap.size() * 32);
        for (size_t i = 0; i < g.cap.size(); ++i) std::memcpy(cb.data() + 32 * i, g.cap[i].data(), 32);
        tr.absorb("fold/trace", cb);
        for (auto& a : g.alphas) a = tr.challenge_ext("fold/alpha");
        segs.push_back(std::move(g));
    }

    // fold.rs:494-570
    Bytes finalize(Transcript& tr) {
        require(!segs.empty(), Z0P_ERR_EMPTY_INPUT);
        const uint64_t s = p.seg_steps, mm = m();
        const int logn = int(ilog2(s)), logb = int(ilog2(p.blowup));
        std::vector<E4> gammas(segs.size());
        for (auto& gm : gammas) gm = tr.challenge_ext("fold/gamma");
        const TreeGeom tg = tree_geom(mm, p.cap_height);
        std::vector<FoldStepDesc> desc(segs.size());
        for (size_t k = 0; k < segs.size(); ++k) {
            FoldStepDesc& d = desc[k];
            uint32_t* A = segs[k].mem->as<uint32_t>();
            d.a = A;
            d.b = A + mm;
            d.tree = A + 2 * mm;
            d.in_a = segs[k].claims[0];
            d.in_b = segs[k].claims[1];
            d.out_a = segs[k].claims[2];
            d.out_b = segs[k].claims[3];
            for (int t = 0; t < FOLD_ALPHAS; ++t) {
                const E4 beta = e4_mul(gammas[k], segs[k].alphas[t]);
                for (int q = 0; q < 4; ++q) d.beta_m[t][q] = to_mont(beta.v[q]);
            }
        }
        DevBuf desc_dev(c, desc.size() * sizeof(FoldStepDesc));
        hip_check(hipMemcpyAsync(desc_dev.p, desc.data(), desc.size() * sizeof(FoldStepDesc), hipMemcpyHostToDevice,
                                 c.stream));
        z0p_fri_params fp = fold_fri_params(p);
        std::unique_ptr<Proof, void (*)(Proof*)> fri(
            proof_new_fold_fri(c, desc_dev.as<FoldStepDesc>(), uint32_t(desc.size()), s, p.blowup, fp, &tr),
            proof_free);
        run_proof(c, fri.get());
        const std::vector<uint64_t>& queries = proof_queries(fri.get());
        // Step openings: A, B at the index and the next row, two paths.
        const uint32_t wpq = 4 + 16 * uint32_t(tg.target);
        const size_t nq = queries.size(), words = segs.size() * nq * size_t(wpq);
        DevBuf qdev(c, nq * 8), odev(c, words * 4);
        hip_check(hipMemcpyAsync(qdev.p, queries.data(), nq * 8, hipMemcpyHostToDevice, c.stream));
        FoldOpenJob oj{};
        oj.steps = desc_dev.as<FoldStepDesc>();
        oj.nsteps = uint32_t(segs.size());
        oj.nq = uint32_t(nq);
        oj.queries = qdev.as<uint64_t>();
        oj.logn = logn;
        oj.logb = logb;
        oj.target = tg.target;
        oj.words_per_query = wpq;
        oj.out = odev.as<uint32_t>();
        launch_fold_open(c.launch(), oj);
        std::vector<uint32_t> ow(words);
        hip_check(hipMemcpyAsync(ow.data(), odev.p, words * 4, hipMemcpyDeviceToHost, c.stream));
        hip_check(hipStreamSynchronize(c.stream));
        // IvcProof::to_bytes (fold.rs:281-310)
        const Bytes& cm = proof_commitment(fri.get());
        const Bytes& pf = proof_bytes(fri.get());
        Writer w;
        w.reserve(words * 5 + cm.size() + pf.size() + segs.size() * (64 + 32 * (size_t(1) << tg.drop)));
        w.byte(WIRE_VERSION);
        w.u32(uint32_t(segs.size()));
        const int logm = logn + logb;
        for (size_t k = 0; k < segs.size(); ++k) {
            const Segment& g = segs[k];
            for (int i = 0; i < 4; ++i) w.u32(g.claims[i]);
            w.u32(uint32_t(g.cap.size()));
            for (auto& d : g.cap) w.raw(d.data(), 32);
            w.u32(uint32_t(nq));
            for (size_t q = 0; q < nq; ++q) {
                const uint32_t* o = ow.data() + (k * nq + q) * wpq;
                const uint64_t index = brev_host(queries[q], logm);
                const uint64_t shifted = (index + p.blowup) % mm;
                w.u64(index);
                w.words(o, 4);
                write_row_path(w, o + 4, tg.target, index);
                write_row_path(w, o + 4 + 8 * tg.target, tg.target, shifted);
            }
        }
        // the FRI commitment and proof bodies, without their version bytes
        require(cm.size() >= 1 && pf.size() >= 1, Z0P_ERR_DEVICE);
        w.raw(cm.data() + 1, cm.size() - 1);
        w.raw(pf.data() + 1, pf.size() - 1);
        segs.clear();  // finalize consumes the accumulator (fold.rs:494)
        return w.finish();
    }

    // Authentication path from `count` digests of 8 words; one direction
    // byte per step, 1 when the opened node is the left child.
    static void write_row_path(Writer& w, const uint32_t* words, int count, uint64_t index) {
        w.u32(uint32_t(count));
        uint8_t* o = w.room(33 * size_t(count));
        for (int l = 0; l < count; ++l, o += 33) {
            std::memcpy(o, words + 8 * l, 32);
            o[32] = ((index >> l) & 1) == 0 ? 1 : 0;
        }
    }
};

FoldAcc* fold_acc_new(Ctx& c, const z0p_fold_params& p) { return new FoldAcc(c, p); }
void fold_acc_free(FoldAcc* a) { delete a; }
size_t fold_acc_len(const FoldAcc* a) { return a->segs.size(); }
void fold_acc_add_segment(FoldAcc* a, const uint32_t* col_a, uint64_t len_a, const uint32_t* col_b, uint64_t len_b,
                          Transcript& tr) {
    a->add_segment(col_a, len_a, col_b, len_b, tr);
}
Bytes fold_acc_finalize(FoldAcc* a, Transcript& tr) { return a->finalize(tr); }

}  // namespace z0p
