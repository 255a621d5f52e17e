// HIP kernels for gfx950 (MI355X): the proving hot path of the Fibonacci
// AIR / FRI commitment. All arithmetic is exact 32-bit integer work; no
// MFMA. Layout conventions:
//  - LDE columns and FRI layers live in bit-reversed order over the whole
//    domain of size m = B*n: natural index i = j + B*t (coset j, step t)
//    sits at brev_{log B}(j)*n + brev_{log n}(t). Coset j is therefore one
//    contiguous block, FRI quads are adjacent, and the next trace row
//    (i + B) stays inside its block.
//  - Digests are 8 little-endian u32 words; tree levels are stored
//    contiguously, leaf level first.
//  - Every kernel takes (inline descriptor, table, njobs): njobs == 0 runs
//    the inline job over the whole grid, otherwise each workgroup looks its
//    job up by block index (device_common.hpp: job_of).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <type_traits>

#include "device_common.hpp"
#include "profiler.hpp"
#include "status.hpp"

namespace z0p {

using namespace dev;

namespace {

// ---------------------------------------------------------------------------
// Fibonacci column coefficients in closed form (see FibCoeffJob): eight
// positions per thread, strided by the workgroup (coalesced stores), with one
// batched inversion of their 16 denominators.
constexpr int FIB_ELT = 8;

__global__ __launch_bounds__(WG) void k_fib_coeffs(FibCoeffJob inl, const FibCoeffJob* tab, uint32_t njobs) {
    uint32_t local;
    const FibCoeffJob J = job_of(inl, tab, njobs, local);
    const uint32_t n = 1u << J.logn;
    const uint32_t base = local * (WG * FIB_ELT) + threadIdx.x;
    uint32_t den[2 * FIB_ELT];
#pragma unroll
    for (int e = 0; e < FIB_ELT; ++e) {
        const uint32_t P = base + e * WG;
        const uint32_t k = P < n ? brev_bits(P, J.logn) : 0u;
        const uint32_t w = pow_lookup(J.winv, k << J.logb);
        den[2 * e] = sub(R1, mont_mul(J.phi_m, w));
        den[2 * e + 1] = sub(R1, mont_mul(J.psi_m, w));
    }
    batch_inv_m<2 * FIB_ELT>(den);
#pragma unroll
    for (int e = 0; e < FIB_ELT; ++e) {
        const uint32_t P = base + e * WG;
        if (P >= n) break;
        const uint32_t c = mont_mul2(J.k1_m, den[2 * e], J.k2_m, den[2 * e + 1]);
        gm(J.out)[P] = mont_mul(c, 1u);  // out of Montgomery form
    }
}

// ---------------------------------------------------------------------------
// NTT passes. n = n1 * n2; "rows" are contiguous length-n2 segments,
// "columns" strided length-n1 transforms handled `cols` at a time.
//
// INTT (natural x[r*n2 + c]):
//   cols (over r): Y[u*n2 + c] = w_n^{-uc} * sum_r x[r n2 + c] w_n1^{-ur}
//   rows (over c): X[u + n1 v] at brev(u)*n2 + brev(v), times 1/n
// LDE (bit-reversed coeffs, k = r*n1 + c sits at brev(c)*n2 + brev(r)):
//   rows (over r): Y[c*n2 + u] = s_j^c w_n^{uc} * sum_r c_k (s_j^{n1})^r w_n2^{ur}
//   cols (over c): X[u + n2 v] at block_j*n + brev(u)*n1 + brev(v)
constexpr int COL_LDS_WORDS = 8192;

// ---- 4096-point transforms: 256 threads x 16 registers -------------------
// Three radix-16 passes of four radix-2 stages each, with two LDS
// exchanges instead of twelve barriers. Element index i of the transform:
//   pass 0: thread q holds i = q + 256 t          (stage distances 256..2048)
//   pass 1: thread (blk = q >> 4, b = q & 15) holds i = 256 blk + b + 16 t
//   pass 2: thread q holds i = 16 q + t           (stage distances 1..8)
// A stage of distance h pairs (i, i + h) with twiddle w_4096^(j * 2048 / h),
// j = i mod h; DIF runs passes 0, 1, 2 with h falling (natural in,
// bit-reversed out), DIT runs 2, 1, 0 with h rising (bit-reversed in,
// natural out). LDS rows are padded by one word per 16.
constexpr int N4K = 4096;
constexpr int N4K_LDS = N4K + N4K / 16;

__device__ __forceinline__ int pad16(int i) { return i + (i >> 4); }

// DIT: b may be unreduced (< 2p), a must be canonical; outputs canonical.
// DIF: inputs and outputs canonical (the difference the product consumes
// stays unreduced).
template <bool DIT>
__device__ __forceinline__ void bfly(uint32_t& a, uint32_t& b, uint32_t w) {
    if (DIT) {
        const uint32_t v = mont_mul(b, w);
        b = sub(a, v);
        a = add(a, v);
    } else {
        const uint32_t u = a;
        a = add(u, b);
        b = mont_mul(sub_lazy(u, b), w);
    }
}

// In a DIT pass of the register radix-16 transform, the output at register
// index i of the stage of distance D is the multiplied operand of the next
// stage (distance 2D, same pass) when bit 2D of i is set; it may then stay
// unreduced, unless that next butterfly has the trivial twiddle (pass 2 with
// i & (2D - 1) == 0), which adds it as it is. Operands that are added next
// (bit 2D clear) and the last stage's outputs stay canonical.
template <int P, int D>
constexpr bool dit_lazy_out(int i) {
    return D < 8 && (i & (2 * D)) != 0 && !(P == 2 && (i & (2 * D - 1)) == 0);
}

// LDS twiddle tables of a 4096-point row (twl, TWL_WORDS): the 2048 powers
// w^e (pass 0 reads w^((q + 256 r) 8/D): consecutive lanes read words 8/D
// apart), then for each stage D of pass 1 the 16D powers w^(k 128/D) (pass 1
// reads entry (q mod 16) + 16 r: consecutive words, or one word).
constexpr int TWL_WORDS = 2048 + 16 * (1 + 2 + 4 + 8);
__device__ __forceinline__ constexpr int twl1_base(int D) { return 2048 + 16 * (D - 1); }

__device__ __forceinline__ void load_twl(uint32_t* twl, gcptr tw) {
    for (int i = int(threadIdx.x); i < TWL_WORDS; i += WG) {
        int e = i;
        if (i >= 2048) {
            const int j = i - 2048;
            const int D = j < 16 ? 1 : j < 48 ? 2 : j < 112 ? 4 : 8;
            e = (j - 16 * (D - 1)) * (128 / D);
        }
        twl[i] = tw[e];
    }
}

// One stage of register distance D within pass P; q is the thread index.
// With twl (LDS tables, load_twl) passes 0 and 1 read their twiddles there;
// pass 2's are the same for every thread and come from tw.
template <bool DIT, int P, int D, bool TWL = false>
__device__ __forceinline__ void stage16(uint32_t (&x)[16], gcptr tw, uint32_t q, const uint32_t* twl = nullptr) {
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        if (t & D) continue;
        const uint32_t r = uint32_t(t & (D - 1));
        if (P == 2 && r == 0) {  // twiddle w^0 = 1: both inputs canonical here
            const uint32_t u = x[t];
            if (DIT) {
                x[t] = dit_lazy_out<P, D>(t) ? add_lazy(u, x[t + D]) : add(u, x[t + D]);
                x[t + D] = dit_lazy_out<P, D>(t + D) ? sub_lazy(u, x[t + D]) : sub(u, x[t + D]);
            } else {
                x[t] = add(u, x[t + D]);
                x[t + D] = sub(u, x[t + D]);
            }
            continue;
        }
        uint32_t w;
        if (P == 0) w = TWL ? twl[(q + 256u * r) * (8u / D)] : tw[(q + 256u * r) * (8u / D)];
        else if (P == 1) w = TWL ? twl[twl1_base(D) + (q & 15u) + 16u * r] : tw[((q & 15u) + 16u * r) * (128u / D)];
        else w = tw[r * (2048u / D)];
        if (DIT) {
            const uint32_t v = mont_mul(x[t + D], w);
            x[t + D] = dit_lazy_out<P, D>(t + D) ? sub_lazy(x[t], v) : sub(x[t], v);
            x[t] = dit_lazy_out<P, D>(t) ? add_lazy(x[t], v) : add(x[t], v);
        } else {
            bfly<false>(x[t], x[t + D], w);
        }
    }
}

template <bool DIT, int P, bool TWL = false>
__device__ __forceinline__ void pass16(uint32_t (&x)[16], gcptr tw, uint32_t q, const uint32_t* twl = nullptr) {
    if (DIT) {
        stage16<DIT, P, 1, TWL>(x, tw, q, twl);
        stage16<DIT, P, 2, TWL>(x, tw, q, twl);
        stage16<DIT, P, 4, TWL>(x, tw, q, twl);
        stage16<DIT, P, 8, TWL>(x, tw, q, twl);
    } else {
        stage16<DIT, P, 8, TWL>(x, tw, q, twl);
        stage16<DIT, P, 4, TWL>(x, tw, q, twl);
        stage16<DIT, P, 2, TWL>(x, tw, q, twl);
        stage16<DIT, P, 1, TWL>(x, tw, q, twl);
    }
}

template <int P>
__device__ __forceinline__ int idx16(uint32_t q, int t) {
    if (P == 0) return int(q) + 256 * t;
    if (P == 1) return int(q >> 4) * 256 + int(q & 15) + 16 * t;
    return 16 * int(q) + t;
}

template <int P>
__device__ __forceinline__ void lds_put(uint32_t* s, const uint32_t (&x)[16], uint32_t q) {
#pragma unroll
    for (int t = 0; t < 16; ++t) s[pad16(idx16<P>(q, t))] = x[t];
}
template <int P>
__device__ __forceinline__ void lds_get(const uint32_t* s, uint32_t (&x)[16], uint32_t q) {
#pragma unroll
    for (int t = 0; t < 16; ++t) x[t] = s[pad16(idx16<P>(q, t))];
}

// DIF over x in pass-0 layout; returns with x in pass-2 layout (bit-reversed
// positions 16 q + t).
__device__ __forceinline__ void dif4096(uint32_t (&x)[16], uint32_t* s, gcptr tw) {
    const uint32_t q = threadIdx.x;
    pass16<false, 0>(x, tw, q);
    lds_put<0>(s, x, q);
    __syncthreads();
    lds_get<1>(s, x, q);
    pass16<false, 1>(x, tw, q);
    lds_put<1>(s, x, q);
    __syncthreads();
    lds_get<2>(s, x, q);
    pass16<false, 2>(x, tw, q);
}

// DIT over x in pass-2 layout (bit-reversed input positions 16 q + t);
// returns with x in pass-0 layout (natural i = q + 256 t). With twl, the LDS
// twiddle tables (load_twl, complete by the first barrier) serve passes 1, 0.
template <bool TWL = false>
__device__ __forceinline__ void dit4096(uint32_t (&x)[16], uint32_t* s, gcptr tw, const uint32_t* twl = nullptr) {
    const uint32_t q = threadIdx.x;
    pass16<true, 2>(x, tw, q);
    lds_put<2>(s, x, q);
    __syncthreads();
    lds_get<1>(s, x, q);
    pass16<true, 1, TWL>(x, tw, q, twl);
    lds_put<1>(s, x, q);
    __syncthreads();
    lds_get<0>(s, x, q);
    pass16<true, 0, TWL>(x, tw, q, twl);
}

// Inverse of one 4096-point row (natural evaluations at `in`), scaled by
// 1/n, bit-reversed coefficients to `out`.
__device__ void intt4096_row(gcptr in, gptr out, gcptr tw, uint32_t ninv_m, uint32_t* s) {
    const uint32_t q = threadIdx.x;
    uint32_t x[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) x[t] = in[q + 256 * t];
    dif4096(x, s, tw);
    __syncthreads();  // everyone has read pass-2 inputs before they are overwritten
#pragma unroll
    for (int t = 0; t < 16; ++t) s[pad16(16 * int(q) + t)] = mont_mul(x[t], ninv_m);
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 16; ++t) out[q + 256 * t] = s[pad16(int(q) + 256 * t)];
}

// ---- column transforms: C = 4096 / N1 columns of N1 = 2^L points --------
// The tile sits in LDS row-major with pitch C + 1 (element i of column c at
// i * (C + 1) + c): global rows load and store coalesced, and a read down a
// column has odd stride. Thread (c = tid mod C, ti = tid / C) owns 16/G
// groups of G = 2^R elements i = blk * S * G + base + S * t' for a pass of R
// stages with distances h = S * 2^k; twiddle exponent j * 2048 / h in
// w_4096 units with j = base + S * (t' mod 2^k).
// lS and logC are ints, or std::integral_constant when the caller knows them
// at compile time (then every index folds to a per-thread base + constants).
template <bool DIT, int R, typename LS, typename LOGC>
__device__ __forceinline__ void cols_pass(uint32_t* s, LS lS_, LOGC logC_, gcptr tw) {
    constexpr int G = 1 << R, NG = 16 / G;
    const int lS = int(lS_), logC = int(logC_);
    const int tid = threadIdx.x;
    const int C = 1 << logC, pitch = C + 1;
    const int c = tid & (C - 1), ti = tid >> logC;
    const int S = 1 << lS;
    uint32_t x[16];
    int i0[NG];
    uint32_t base[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        const int gid = ti * NG + g;
        base[g] = uint32_t(gid & (S - 1));
        i0[g] = ((gid >> lS) << (lS + R)) + int(base[g]);
#pragma unroll
        for (int t = 0; t < G; ++t) x[g * G + t] = s[(i0[g] + (t << lS)) * pitch + c];
    }
#pragma unroll
    for (int kk = 0; kk < R; ++kk) {
        const int k = DIT ? kk : R - 1 - kk;
        const int D = 1 << k;
#pragma unroll
        for (int g = 0; g < NG; ++g) {
#pragma unroll
            for (int t = 0; t < G; ++t) {
                if (t & D) continue;
                const uint32_t j = base[g] + (uint32_t(t & (D - 1)) << lS);
                bfly<DIT>(x[g * G + t], x[g * G + t + D], tw[j << (11 - lS - k)]);
            }
        }
    }
#pragma unroll
    for (int g = 0; g < NG; ++g)
#pragma unroll
        for (int t = 0; t < G; ++t) s[(i0[g] + (t << lS)) * pitch + c] = x[g * G + t];
}

template <bool DIT>
__device__ __forceinline__ void cols_partial(uint32_t* s, int r0, int logC, gcptr tw) {
    if (r0 == 1) cols_pass<DIT, 1>(s, 0, logC, tw);
    else if (r0 == 2) cols_pass<DIT, 2>(s, 0, logC, tw);
    else cols_pass<DIT, 3>(s, 0, logC, tw);
}

// Full transform of every column of the tile (4 <= L <= 12); the tile must
// be complete in LDS on entry (barrier done) and is complete on exit.
template <bool DIT>
__device__ __forceinline__ void cols_ntt(uint32_t* s, int L, int logC, gcptr tw) {
    const int r0 = L & 3;
    if (DIT) {
        if (r0) {
            cols_partial<true>(s, r0, logC, tw);
            __syncthreads();
        }
        for (int lS = r0; lS + 4 <= L; lS += 4) {
            cols_pass<true, 4>(s, lS, logC, tw);
            __syncthreads();
        }
    } else {
        for (int lS = L - 4; lS >= r0; lS -= 4) {
            cols_pass<false, 4>(s, lS, logC, tw);
            __syncthreads();
        }
        if (r0) {
            cols_partial<false>(s, r0, logC, tw);
            __syncthreads();
        }
    }
}

// The DIF column transform for a column length 2^L known at compile time
// (the LDE's columns pass): the same passes as cols_ntt<false>.
template <int V>
using ic = std::integral_constant<int, V>;

template <int L>
__device__ __forceinline__ void cols_dif_t(uint32_t* s, gcptr tw) {
    constexpr int r0 = L & 3, LOGC = 12 - L;
    if constexpr (L - 4 >= r0) {
        cols_pass<false, 4>(s, ic<L - 4>{}, ic<LOGC>{}, tw);
        __syncthreads();
    }
    if constexpr (L - 8 >= r0) {
        cols_pass<false, 4>(s, ic<L - 8>{}, ic<LOGC>{}, tw);
        __syncthreads();
    }
    if constexpr (L - 12 >= r0) {
        cols_pass<false, 4>(s, ic<L - 12>{}, ic<LOGC>{}, tw);
        __syncthreads();
    }
    if constexpr (r0 != 0) {
        cols_pass<false, r0>(s, ic<0>{}, ic<LOGC>{}, tw);
        __syncthreads();
    }
}

__device__ __forceinline__ bool cols_in_regs(int log1) { return log1 >= 4 && log1 <= 12; }

__device__ void ntt_intt_small(const NttJob& J, uint32_t* lds) {
    if (J.logn == 12) {
        intt4096_row(gc(J.in), gm(J.out), gc(J.tw4096), J.ninv_m, lds);
        return;
    }
    const int n = 1 << J.logn;
    uint32_t* s = lds;
    uint32_t* tw = lds + n;
    load_tw(tw, J.tw4096, J.logn);
    for (int k = threadIdx.x; k < n; k += blockDim.x) s[k] = J.in[k];
    __syncthreads();
    lds_dif(s, J.logn, 1, n, tw);
    for (int q = threadIdx.x; q < n; q += blockDim.x) J.out[q] = mont_mul(s[q], J.ninv_m);
}

// Coefficient k of the masked polynomial on coset j: the mask term
// (x^n - 1) R(x) is (z_j - 1) sum_b z_j^b R_b(x) there, R_b being the b-th
// n-coefficient block of R.
template <bool MASK>
__device__ __forceinline__ uint32_t lde_coeff(const NttJob& J, uint32_t c, uint32_t k, uint32_t j) {
    if (!MASK || k >= J.mask_len) return c;
    uint32_t f = gc(J.zvan)[j];
    const uint32_t z = add(f, R1);
    for (uint32_t i = k; i < J.mask_len; i += 1u << J.logn) {
        c = add(c, mont_mul(gc(J.mask)[i], f));
        f = mont_mul(f, z);
    }
    return c;
}

template <bool MASK>
__device__ void ntt_lde_small(const NttJob& J, uint32_t* lds, uint32_t j) {
    const int n = 1 << J.logn;
    if (J.logn == 12) {
        // natural k = q + 256 t: x = c_k s_j^k (c_k at bit-reversed position)
        const uint32_t q = threadIdx.x;
        const gcptr sc = gc(J.tab_a) + size_t(j) * n;
        const gcptr in = gc(J.in);
        uint32_t x[16];
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            const uint32_t k = q + 256u * t;
            x[t] = mont_mul(lde_coeff<MASK>(J, in[brev_bits(k, 12)], k, j), sc[k]);
        }
        dif4096(x, lds, gc(J.tw4096));
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 16; ++t) lds[pad16(16 * int(q) + t)] = x[t];
        __syncthreads();
        const gptr o = gm(J.out) + (size_t(brev_bits(j, J.logb)) << J.logn);
#pragma unroll
        for (int t = 0; t < 16; ++t) o[q + 256 * t] = lds[pad16(int(q) + 256 * t)];
        return;
    }
    uint32_t* s = lds;
    uint32_t* tw = lds + n;
    load_tw(tw, J.tw4096, J.logn);
    const uint32_t* sc = J.tab_a + size_t(j) * n;
    for (int P = threadIdx.x; P < n; P += blockDim.x) {
        uint32_t k = brev_bits(P, J.logn);
        s[k] = mont_mul(lde_coeff<MASK>(J, J.in[P], k, j), sc[k]);
    }
    __syncthreads();
    lds_dif(s, J.logn, 1, n, tw);
    uint32_t* o = J.out + (size_t(brev_bits(j, J.logb)) << J.logn);
    for (int q = threadIdx.x; q < n; q += blockDim.x) o[q] = s[q];
}

// Column transforms use a padded LDS stride (n1 + 1) so the transposing
// loads and stores hit distinct banks.
__device__ void ntt_intt_cols(const NttJob& J, uint32_t* lds, uint32_t bx) {
    if (cols_in_regs(J.log1)) {
        const int L = J.log1, logC = 12 - L, C = 1 << logC, pitch = C + 1;
        const int n2 = 1 << J.log2;
        const int c0 = int(bx) * C;
        const int tid = threadIdx.x;
        const gcptr in = gc(J.in);
        const gptr out = gm(J.out);
        uint32_t v[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int e = tid + 256 * k, r = e >> logC, c = e & (C - 1);
            v[k] = in[size_t(r) * n2 + c0 + c];
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int e = tid + 256 * k, r = e >> logC, c = e & (C - 1);
            lds[r * pitch + c] = v[k];
        }
        __syncthreads();
        cols_ntt<false>(lds, L, logC, gc(J.tw4096));
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int e = tid + 256 * k, qq = e >> logC, c = e & (C - 1);
            const uint32_t u = brev_bits(uint32_t(qq), L);
            const uint32_t cc = uint32_t(c0 + c);
            out[size_t(u) * n2 + cc] = mont_mul(lds[qq * pitch + c], pow_lookup(J.pt, (u * cc) << J.logb));
        }
        return;
    }
    const int n1 = 1 << J.log1, n2 = 1 << J.log2, cols = J.cols;
    const int st = n1 + 1;
    uint32_t* s = lds;
    uint32_t* tw = lds + cols * st;
    load_tw(tw, J.tw4096, J.log1);
    const int c0 = bx * cols;
    for (int e = threadIdx.x; e < cols * n1; e += blockDim.x) {
        int r = e / cols, c = e % cols;
        s[c * st + r] = J.in[size_t(r) * n2 + c0 + c];
    }
    __syncthreads();
    lds_dif(s, J.log1, cols, st, tw);
    for (int e = threadIdx.x; e < cols * n1; e += blockDim.x) {
        int q = e / cols, c = e % cols;
        uint32_t u = brev_bits(q, J.log1);
        uint32_t cc = uint32_t(c0 + c);
        uint32_t tw_uc = pow_lookup(J.pt, (u * cc) << J.logb);
        J.out[size_t(u) * n2 + cc] = mont_mul(s[c * st + q], tw_uc);
    }
}

__device__ void ntt_intt_rows(const NttJob& J, uint32_t* lds, uint32_t u) {
    const int n2 = 1 << J.log2;
    if (J.log2 == 12) {
        intt4096_row(gc(J.in) + size_t(u) * n2, gm(J.out) + size_t(brev_bits(u, J.log1)) * n2, gc(J.tw4096), J.ninv_m,
                     lds);
        return;
    }
    uint32_t* s = lds;
    uint32_t* tw = lds + n2;
    load_tw(tw, J.tw4096, J.log2);
    const uint32_t* row = J.in + size_t(u) * n2;
    for (int c = threadIdx.x; c < n2; c += blockDim.x) s[c] = row[c];
    __syncthreads();
    lds_dif(s, J.log2, 1, n2, tw);
    uint32_t* o = J.out + size_t(brev_bits(u, J.log1)) * n2;
    for (int q = threadIdx.x; q < n2; q += blockDim.x) o[q] = mont_mul(s[q], J.ninv_m);
}

// The coefficient row is already bit-reversed, so a DIT transform consumes
// it in place; the row table is stored bit-reversed to match.
// Rows pass of the coset LDE for 4096-point rows (k_lde_rows): thread q
// loads coefficients 16q .. 16q + 15 of row R (four 16-byte loads; the row
// is bit-reversed, so this is exactly the DIT's first-pass register layout)
// with their row-table scales, runs the DIT, and writes
// y[u] = x * s_j^c * w_n^(u c) for u = q + 256 t, c = brev(R).
template <bool MASK>
__device__ void lde_rows_4096(const NttJob& J, uint32_t* lds, uint32_t R, uint32_t j) {
    const uint32_t q = threadIdx.x;
    uint32_t* twl = lds + N4K_LDS;  // the launch sizes LDS for N4K_LDS + TWL_WORDS words
    load_twl(twl, gc(J.tw4096));
    const int logn = J.log1 + 12;
    const uint32_t c = brev_bits(R, J.log1);
    const uint32_t* rt = J.tab_a + size_t(j) * N4K + 16 * q;
    const uint32_t* in = J.in + size_t(R) * N4K + 16 * q;
    uint32_t x[16];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const u32x4 a = ld4(in, v), sc = ld4(rt, v);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t C = 16 * q + 4 * v + k;
            x[4 * v + k] = mont_mul(lde_coeff<MASK>(J, a[k], brev_bits(R * uint32_t(N4K) + C, logn), j), sc[k]);
        }
    }
    dit4096<true>(x, lds, gc(J.tw4096), twl);
    // y[u] = x * s_j^c * w_n^(u c), u = q + 256 t: geometric in t.
    uint32_t f = mont_mul(pow_lookup(J.pt, (q * c) << J.logb), gc(J.tab_b)[(size_t(j) << J.log1) + c]);
    const uint32_t step = pow_lookup(J.pt, (256u * c) << J.logb);
    const gptr yo = gm(J.out) + (size_t(j) << logn) + size_t(c) * N4K;
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        yo[q + 256 * t] = mont_mul(x[t], f);
        f = mont_mul(f, step);
    }
}

// Rows of any other length (k_ntt): DIT in LDS from the bit-reversed row.
template <bool MASK>
__device__ void ntt_lde_rows(const NttJob& J, uint32_t* lds, uint32_t R, uint32_t j) {
    const int n1 = 1 << J.log1, n2 = 1 << J.log2;
    const int logn = J.log1 + J.log2;
    uint32_t* s = lds;
    uint32_t* tw = lds + n2;
    load_tw(tw, J.tw4096, J.log2);
    const uint32_t c = brev_bits(R, J.log1);
    const uint32_t* rt = J.tab_a + size_t(j) * n2;
    const uint32_t* in = J.in + size_t(R) * n2;
    for (int Cc = threadIdx.x; Cc < n2; Cc += blockDim.x)
        s[Cc] = mont_mul(lde_coeff<MASK>(J, in[Cc], brev_bits(R * uint32_t(n2) + uint32_t(Cc), logn), j), rt[Cc]);
    __syncthreads();
    lds_dit(s, J.log2, 1, n2, tw);
    const uint32_t colscale = J.tab_b[size_t(j) * n1 + c];
    uint32_t* yo = J.out + (size_t(j) << logn) + size_t(c) * n2;
    for (int u = threadIdx.x; u < n2; u += blockDim.x) {
        uint32_t f = mont_mul(pow_lookup(J.pt, (uint32_t(u) * c) << J.logb), colscale);
        yo[u] = mont_mul(s[u], f);
    }
}

// Columns pass of the coset LDE (k_lde_cols) for 2^4..2^12-point columns:
// a 4096-element tile of C = 4096 / n1 columns through LDS (pitch C + 1),
// the DIF transform in registers, bit-reversed rows out.
template <int L>
__device__ __forceinline__ void lde_cols_tile_t(const NttJob& J, uint32_t* lds, uint32_t bx, uint32_t j) {
    constexpr int n1 = 1 << L, n2 = N4K;
    constexpr int logn = L + 12;
    constexpr int logC = 12 - L, C = 1 << logC, pitch = C + 1;
    const int u0 = int(bx) * C;
    const int tid = threadIdx.x;
    const gcptr yj = gc(J.in) + (size_t(j) << logn);
    uint32_t v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int e = tid + 256 * k, r = e >> logC, c = e & (C - 1);
        v[k] = yj[size_t(r) * n2 + u0 + c];
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int e = tid + 256 * k, r = e >> logC, c = e & (C - 1);
        lds[r * pitch + c] = v[k];
    }
    __syncthreads();
    cols_dif_t<L>(lds, gc(J.tw4096));
    const gptr oj = gm(J.out) + (size_t(brev_bits(j, J.logb)) << logn);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int e = tid + 256 * k, qq = e & (n1 - 1), uu = e >> L;
        oj[size_t(brev_bits(uint32_t(u0 + uu), 12)) * n1 + qq] = lds[qq * pitch + uu];
    }
}

// Columns pass for 2^4..2^12-point columns of 4096-point rows, one
// instantiation per column length.
__device__ __forceinline__ void lde_cols_tile(const NttJob& J, uint32_t* lds, uint32_t bx, uint32_t j) {
    switch (J.log1) {
        case 4: lde_cols_tile_t<4>(J, lds, bx, j); break;
        case 5: lde_cols_tile_t<5>(J, lds, bx, j); break;
        case 6: lde_cols_tile_t<6>(J, lds, bx, j); break;
        case 7: lde_cols_tile_t<7>(J, lds, bx, j); break;
        case 8: lde_cols_tile_t<8>(J, lds, bx, j); break;
        case 9: lde_cols_tile_t<9>(J, lds, bx, j); break;
        case 10: lde_cols_tile_t<10>(J, lds, bx, j); break;
        case 11: lde_cols_tile_t<11>(J, lds, bx, j); break;
        case 12: lde_cols_tile_t<12>(J, lds, bx, j); break;
        default: break;
    }
}

// Columns of other lengths (k_ntt).
__device__ void ntt_lde_cols(const NttJob& J, uint32_t* lds, uint32_t bx, uint32_t j) {
    const int n1 = 1 << J.log1, n2 = 1 << J.log2, cols = J.cols;
    const int logn = J.log1 + J.log2;
    const int st = n1 + 1;
    uint32_t* s = lds;
    uint32_t* tw = lds + cols * st;
    load_tw(tw, J.tw4096, J.log1);
    const int u0 = bx * cols;
    const uint32_t* yj = J.in + (size_t(j) << logn);
    for (int e = threadIdx.x; e < cols * n1; e += blockDim.x) {
        int c = e / cols, uu = e % cols;
        s[uu * st + c] = yj[size_t(c) * n2 + u0 + uu];
    }
    __syncthreads();
    lds_dif(s, J.log1, cols, st, tw);
    uint32_t* oj = J.out + (size_t(brev_bits(j, J.logb)) << logn);
    for (int e = threadIdx.x; e < cols * n1; e += blockDim.x) {
        int uu = e / n1, q = e % n1;
        uint32_t u = uint32_t(u0 + uu);
        oj[size_t(brev_bits(u, J.log2)) * n1 + q] = s[uu * st + q];
    }
}

__global__ __launch_bounds__(WG) void k_ntt(NttJob inl, const NttJob* tab, uint32_t njobs) {
    extern __shared__ uint32_t lds[];
    uint32_t local;
    const NttJob J = job_of(inl, tab, njobs, local);
    switch (J.kind) {
        case NTT_INTT_SMALL:
            ntt_intt_small(J, lds);
            break;
        case NTT_INTT_COLS:
            ntt_intt_cols(J, lds, local);
            break;
        case NTT_INTT_ROWS:
            ntt_intt_rows(J, lds, local);
            break;
        case NTT_LDE_SMALL:
            if (J.mask)
                ntt_lde_small<true>(J, lds, local);
            else
                ntt_lde_small<false>(J, lds, local);
            break;
        case NTT_LDE_ROWS: {
            const uint32_t n1 = 1u << J.log1;
            if (J.log2 == 12) {
                if (J.mask)
                    lde_rows_4096<true>(J, lds, local % n1, local / n1);
                else
                    lde_rows_4096<false>(J, lds, local % n1, local / n1);
            } else if (J.mask) {
                ntt_lde_rows<true>(J, lds, local % n1, local / n1);
            } else {
                ntt_lde_rows<false>(J, lds, local % n1, local / n1);
            }
            break;
        }
        case NTT_LDE_COLS: {
            const uint32_t per = (1u << J.log2) / uint32_t(J.cols);
            if (cols_in_regs(J.log1) && J.log2 == 12)
                lde_cols_tile(J, lds, local % per, local / per);
            else
                ntt_lde_cols(J, lds, local % per, local / per);
            break;
        }
        default:
            break;
    }
}

// The two passes of the 2^16..2^24-point coset LDEs (4096-point rows, 2^4..
// 2^12-point columns) in kernels of their own: register use and LDS are
// those of one path, not of every NTT kind.
__global__ __launch_bounds__(WG) void k_lde_rows(NttJob inl, const NttJob* tab, uint32_t njobs) {
    extern __shared__ uint32_t lds[];
    uint32_t local;
    const NttJob J = job_of(inl, tab, njobs, local);
    const uint32_t n1 = 1u << J.log1;
    if (J.mask)
        lde_rows_4096<true>(J, lds, local % n1, local / n1);
    else
        lde_rows_4096<false>(J, lds, local % n1, local / n1);
}

__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(7))) void k_lde_cols(NttJob inl, const NttJob* tab, uint32_t njobs) {
    extern __shared__ uint32_t lds[];
    uint32_t local;
    const NttJob J = job_of(inl, tab, njobs, local);
    const uint32_t per = (1u << J.log2) / uint32_t(J.cols);
    // A tile reads 4 * 2^(12 - log1) bytes of each row: from log1 = 8 on,
    // G = 2^(log1 - 7) neighbouring tiles share every 128-byte line. Blocks
    // go to the 8 XCDs round-robin, so the G tiles of a line are given to
    // blocks 8 apart (one XCD, one L2): the line is fetched once. A
    // bijection on each group of 8G blocks (per is a multiple of 8G).
    uint32_t t = local;
    if (J.log1 > 7) {
        const int gl = J.log1 - 7;
        const uint32_t grp = 8u << gl, r = t & (grp - 1);
        t = (t & ~(grp - 1)) | ((r & 7u) << gl) | (r >> 3);
    }
    lde_cols_tile(J, lds, t % per, t / per);
}

__global__ __launch_bounds__(WG) void k_bitrev(BitrevJob inl, const BitrevJob* tab, uint32_t njobs) {
    uint32_t local;
    const BitrevJob J = job_of(inl, tab, njobs, local);
    uint64_t i = uint64_t(local) * WG + threadIdx.x;
    if (i >> J.logn) return;
    const uint64_t src = uint64_t(brev_bits(uint32_t(i), J.logn)) * uint64_t(J.in_stride);
    for (int k = 0; k < J.width; ++k) J.out[i * uint64_t(J.width) + k] = gc(J.in)[src + k] % P;
}

__global__ __launch_bounds__(WG) void k_interleave(InterleaveJob inl, const InterleaveJob* tab, uint32_t njobs) {
    uint32_t local;
    const InterleaveJob J = job_of(inl, tab, njobs, local);
    const uint64_t i = uint64_t(local) * WG + threadIdx.x;
    if (i >= J.count) return;
    const gcptr pl = gc(J.planes);
    st4(J.out, i, u32x4{pl[i], pl[J.count + i], pl[2 * J.count + i], pl[3 * J.count + i]});
}

// ---------------------------------------------------------------------------
// Pointwise kernels over the bit-reversed LDE: ELT elements per thread,
// strided by the workgroup size so every access is coalesced.
constexpr int ELT = 16;
// Composition keeps four values per element live across its inversion; a
// smaller batch keeps its register use (and occupancy) in check.
constexpr int CELT = 8;

// B[i] = A[i + B] + coeff*(x^n - 1)/(x - w_last) (air.rs:806-844);
// x^n - 1 = s_j^n - 1 is constant per coset (table zvan_m), 1/(x - w_last)
// comes from the geometry's point table.
__global__ __launch_bounds__(WG) void k_linked(LinkedJob inl, const LinkedJob* tab, uint32_t njobs) {
    uint32_t local;
    const LinkedJob J = job_of(inl, tab, njobs, local);
    const uint32_t m = 1u << (J.logn + J.logb);
    const uint32_t base = local * (WG * ELT) + threadIdx.x;
    const gcptr A = gc(J.a);
    uint32_t an[ELT], iw[ELT];
    // The next-row loads are scattered inside the coset block: all in
    // flight before the arithmetic.
#pragma unroll
    for (int e = 0; e < ELT; ++e) {
        const uint32_t p = base + e * WG;
        if (p < m) {
            const PosInfo pi = pos_info(p, J.logn, J.logb);
            an[e] = A[pi.pnext];
            iw[e] = gc(J.iw)[p];
        }
    }
#pragma unroll
    for (int e = 0; e < ELT; ++e) {
        const uint32_t p = base + e * WG;
        if (p < m) {
            const uint32_t j = brev_bits(p >> J.logn, J.logb);
            gm(J.b)[p] = add(an[e], mont_mul(iw[e], mont_mul(J.coeff, gc(J.zvan_m)[j])));
        }
    }
}

// Per-point inverses of a geometry (PointInverses): 8 positions per thread,
// strided by the workgroup, with one batched inversion of (x - 1)(x - w).
__global__ __launch_bounds__(WG) void k_point_inv(PointInvJob J) {
    const uint32_t m = 1u << (J.logn + J.logb);
    const uint32_t base = blockIdx.x * (WG * CELT) + threadIdx.x;
    const uint32_t wl1 = sub(J.w_last_m, R1);
    uint32_t d1[CELT], pr[CELT];
#pragma unroll
    for (int e = 0; e < CELT; ++e) {
        const uint32_t p = base + e * WG;
        if (p < m) {
            const uint32_t j = brev_bits(p >> J.logn, J.logb);
            const uint32_t x = mont_mul(gc(J.xn)[p & ((1u << J.logn) - 1)], gc(J.shift_m)[j]);
            d1[e] = sub(x, R1);
            pr[e] = mont_mul(d1[e], sub(d1[e], wl1));
        } else {
            d1[e] = pr[e] = R1;
        }
    }
    batch_inv_m<CELT>(pr);
#pragma unroll
    for (int e = 0; e < CELT; ++e) {
        const uint32_t p = base + e * WG;
        if (p >= m) continue;
        const uint32_t j = brev_bits(p >> J.logn, J.logb);
        const uint32_t dw = sub(d1[e], wl1);  // x - w_last
        gm(J.i1)[p] = mont_mul(dw, pr[e]);
        gm(J.iw)[p] = mont_mul(d1[e], pr[e]);
        gm(J.iz)[p] = mont_mul(dw, gc(J.zinv_m)[j]);
    }
}

// One composition value (air.rs:847-929) from the trace values at the row
// and the next row and the point's inverses:
//   a0*(A'-B)/Z + a1*(B'-A-B)/Z + a2*(A-a0)/(x-1) + a3*(B-b0)/(x-1) + a4*(B-bl)/(x-w)
__device__ __forceinline__ void compose_value(const ComposeConsts& C, uint32_t va, uint32_t vb, uint32_t na,
                                              uint32_t nb, uint32_t i1, uint32_t iw, uint32_t iz, uint32_t (&o)[4]) {
    const uint32_t s0 = mont_mul(sub_lazy(na, vb), iz);
    const uint32_t s1 = mont_mul(sub_lazy(nb, add(va, vb)), iz);
    const uint32_t s2 = mont_mul(sub_lazy(va, C.a0), i1);
    const uint32_t s3 = mont_mul(sub_lazy(vb, C.b0), i1);
    const uint32_t s4 = mont_mul(sub_lazy(vb, C.b_last), iw);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t p01 = mont_mul2(s0, C.alpha_m[0][k], s1, C.alpha_m[1][k]);
        const uint32_t p23 = mont_mul2(s2, C.alpha_m[2][k], s3, C.alpha_m[3][k]);
        o[k] = add(add(p01, p23), mont_mul(s4, C.alpha_m[4][k]));
    }
}

// Composition per point (air.rs:847-929), written as FRI layer 0 (every
// load of an element in flight before its arithmetic).
__global__ __launch_bounds__(WG) void k_compose(ComposeJob inl, const ComposeJob* tab, uint32_t njobs) {
    uint32_t local;
    const ComposeJob J = job_of(inl, tab, njobs, local);
    const uint32_t m = 1u << (J.logn + J.logb);
    const uint32_t base = local * (WG * CELT) + threadIdx.x;
    const gcptr A = gc(J.a), Bv = gc(J.b);
    const ComposeConsts C = *J.c;
    uint32_t va[CELT], vb[CELT], na[CELT], nb[CELT], i1[CELT], iw[CELT], iz[CELT];
#pragma unroll
    for (int e = 0; e < CELT; ++e) {
        const uint32_t p = base + e * WG;
        if (p < m) {
            const PosInfo pi = pos_info(p, J.logn, J.logb);
            na[e] = A[pi.pnext];
            nb[e] = Bv[pi.pnext];
            va[e] = A[p];
            vb[e] = Bv[p];
            i1[e] = gc(J.pinv.i1)[p];
            iw[e] = gc(J.pinv.iw)[p];
            iz[e] = gc(J.pinv.iz)[p];
        }
    }
#pragma unroll
    for (int e = 0; e < CELT; ++e) {
        const uint32_t p = base + e * WG;
        if (p >= m) continue;
        uint32_t o[4];
        compose_value(C, va[e], vb[e], na[e], nb[e], i1[e], iw[e], iz[e], o);
        st4(J.layer0, p, u32x4{o[0], o[1], o[2], o[3]});
    }
}

// Range-check composition per point (range.rs:767-823), as FRI layer 0.
// 1/Z = 1/(s_j^n - 1) is constant per coset; 1/(x - 1) is batch-inverted.
// The booleanity terms share 1/Z: sum_c a_c [C_c (C_c - 1) / R] is summed in
// pairs, then multiplied once by 1/Z in the R^2 domain.
__global__ __launch_bounds__(WG) void k_range_compose(RangeComposeJob inl, const RangeComposeJob* tab,
                                                      uint32_t njobs) {
    uint32_t local;
    const RangeComposeJob J = job_of(inl, tab, njobs, local);
    const uint32_t m = 1u << (J.logn + J.logb);
    const uint32_t base = local * (WG * CELT) + threadIdx.x;
    const gcptr cols = gc(J.cols);
    const RangeConsts* C = J.c;
    const uint32_t nb = C->nbits, value = C->value;
    uint32_t d1[CELT];
#pragma unroll
    for (int e = 0; e < CELT; ++e) {
        const uint32_t p = base + e * WG;
        if (p < m) {
            const uint32_t j = brev_bits(p >> J.logn, J.logb);
            const uint32_t x = mont_mul(gc(J.xn)[p & ((1u << J.logn) - 1)], gc(J.shift_m)[j]);
            d1[e] = sub(x, R1);
        } else {
            d1[e] = R1;
        }
    }
    batch_inv_m<CELT>(d1);
#pragma unroll 1
    for (int e = 0; e < CELT; ++e) {
        const uint32_t p = base + e * WG;
        if (p >= m) break;
        const uint32_t j = brev_bits(p >> J.logn, J.logb);
        const uint32_t iz = gc(J.zinv_m)[j];
        const uint32_t v = cols[p];
        // every bit column's load in flight before the arithmetic
        uint32_t bits[RANGE_MAX_BITS];
#pragma unroll
        for (int c = 0; c < RANGE_MAX_BITS; ++c)
            bits[c] = uint32_t(c) < nb ? cols[uint64_t(c + 1) * J.col_stride + p] : 0u;
        uint32_t rec = 0, acc[4] = {0, 0, 0, 0};
#pragma unroll
        for (int c = RANGE_MAX_BITS - 1; c >= 0; --c) {
            if (uint32_t(c) >= nb) continue;
            const uint32_t b = bits[c];
            rec = add(dbl(rec), b);
            const uint32_t q = mont_mul(b, sub(b, 1u));  // C (C - 1) / R
#pragma unroll
            for (int k = 0; k < 4; ++k) acc[k] = add(acc[k], mont_mul(q, C->alpha_m[1 + c][k]));
        }
        const uint32_t izz = to_mont(iz);  // 1/Z in the R^2 domain
        const uint32_t sd = mont_mul(sub(v, rec), iz);
        const uint32_t sb = mont_mul(sub(v, value), d1[e]);
        uint32_t o[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            o[k] = add(mont_mul2(sd, C->alpha_m[0][k], sb, C->alpha_m[nb + 1][k]), mont_mul(acc[k], izz));
        st4(J.layer0, p, u32x4{o[0], o[1], o[2], o[3]});
    }
}

// Table-membership composition per point (table.rs:597-644), as FRI layer 0.
// The vanishing product runs in the Montgomery domain over the table.
__global__ __launch_bounds__(WG) void k_table_compose(TableComposeJob inl, const TableComposeJob* tab,
                                                      uint32_t njobs) {
    uint32_t local;
    const TableComposeJob J = job_of(inl, tab, njobs, local);
    const uint32_t m = 1u << (J.logn + J.logb);
    const uint32_t base = local * (WG * CELT) + threadIdx.x;
    const TableConsts C = *J.c;
    const gcptr T = gc(J.table_m);
    uint32_t d1[CELT];
#pragma unroll
    for (int e = 0; e < CELT; ++e) {
        const uint32_t p = base + e * WG;
        if (p < m) {
            const uint32_t j = brev_bits(p >> J.logn, J.logb);
            const uint32_t x = mont_mul(gc(J.xn)[p & ((1u << J.logn) - 1)], gc(J.shift_m)[j]);
            d1[e] = sub(x, R1);
        } else {
            d1[e] = R1;
        }
    }
    batch_inv_m<CELT>(d1);
    // The CELT vanishing products advance together: independent chains.
    uint32_t vv[CELT], vm[CELT], pm[CELT];
#pragma unroll
    for (int e = 0; e < CELT; ++e) {
        const uint32_t p = base + e * WG;
        vv[e] = p < m ? gc(J.v)[p] : 0u;
        vm[e] = to_mont(vv[e]);
        pm[e] = R1;
    }
    for (uint32_t k = 0; k < J.table_len; ++k) {
        const uint32_t t = T[k];
#pragma unroll
        for (int e = 0; e < CELT; ++e) pm[e] = mont_mul(pm[e], sub(vm[e], t));
    }
#pragma unroll
    for (int e = 0; e < CELT; ++e) {
        const uint32_t p = base + e * WG;
        if (p >= m) continue;
        const uint32_t j = brev_bits(p >> J.logn, J.logb);
        const uint32_t v = vv[e];
        const uint32_t sm = mont_mul(from_mont(pm[e]), gc(J.zinv_m)[j]);  // prod / Z
        const uint32_t sb = mont_mul(sub(v, C.value), d1[e]);            // (V - value) / (x - 1)
        uint32_t o[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) o[k] = mont_mul2(sm, C.alpha_m[0][k], sb, C.alpha_m[1][k]);
        st4(J.layer0, p, u32x4{o[0], o[1], o[2], o[3]});
    }
}

// IVC batched composition (fold.rs:785-854 per step, 494-521 combined):
// the points' inverses are shared by every step; each step adds its six
// terms, scaled by beta = gamma * alpha, to the Ext4 accumulator.
//   t1 = (A' - B)/Z, t2 = (B' - A - B)/Z with 1/Z = (x - w)/(x^n - 1),
//   (A - in_a)/(x - 1), (B - in_b)/(x - 1), (A - out_a)/(x - w), (B - out_b)/(x - w).
__global__ __launch_bounds__(WG) void k_fold_combine(FoldCombineJob inl, const FoldCombineJob* tab,
                                                     uint32_t njobs) {
    uint32_t local;
    const FoldCombineJob J = job_of(inl, tab, njobs, local);
    const uint32_t m = 1u << (J.logn + J.logb);
    const uint32_t base = local * (WG * CELT) + threadIdx.x;
    const uint32_t wl1 = sub(J.w_last_m, R1);
    uint32_t d1[CELT], pr[CELT];
#pragma unroll
    for (int e = 0; e < CELT; ++e) {
        const uint32_t p = base + e * WG;
        if (p < m) {
            const uint32_t j = brev_bits(p >> J.logn, J.logb);
            const uint32_t x = mont_mul(gc(J.xn)[p & ((1u << J.logn) - 1)], gc(J.shift_m)[j]);
            d1[e] = sub(x, R1);
            pr[e] = mont_mul(d1[e], sub(d1[e], wl1));
        } else {
            d1[e] = pr[e] = R1;
        }
    }
    batch_inv_m<CELT>(pr);
#pragma unroll 1
    for (int e = 0; e < CELT; ++e) {
        const uint32_t p = base + e * WG;
        if (p >= m) break;
        const PosInfo pi = pos_info(p, J.logn, J.logb);
        const uint32_t dw = sub(d1[e], wl1);            // x - w_last
        const uint32_t i1 = mont_mul(dw, pr[e]);        // 1/(x - 1)
        const uint32_t iw = mont_mul(d1[e], pr[e]);     // 1/(x - w_last)
        const uint32_t iz = mont_mul(dw, gc(J.zinv_m)[pi.j]);  // 1/Z
        uint32_t acc[4] = {0, 0, 0, 0};
        for (uint32_t k = 0; k < J.nsteps; ++k) {
            const FoldStepDesc& S = J.steps[k];
            const uint32_t a = gc(S.a)[p], b = gc(S.b)[p];
            const uint32_t an = gc(S.a)[pi.pnext], bn = gc(S.b)[pi.pnext];
            const uint32_t t1 = mont_mul(sub(an, b), iz);
            const uint32_t t2 = mont_mul(sub(bn, add(a, b)), iz);
            const uint32_t ia = mont_mul(sub(a, S.in_a), i1);
            const uint32_t ib = mont_mul(sub(b, S.in_b), i1);
            const uint32_t oa = mont_mul(sub(a, S.out_a), iw);
            const uint32_t ob = mont_mul(sub(b, S.out_b), iw);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t s01 = mont_mul2(t1, S.beta_m[0][q], t2, S.beta_m[1][q]);
                const uint32_t s23 = mont_mul2(ia, S.beta_m[2][q], ib, S.beta_m[3][q]);
                const uint32_t s45 = mont_mul2(oa, S.beta_m[4][q], ob, S.beta_m[5][q]);
                acc[q] = add(acc[q], add(add(s01, s23), s45));
            }
        }
        st4(J.layer0, p, u32x4{acc[0], acc[1], acc[2], acc[3]});
    }
}

// ---------------------------------------------------------------------------
// Subtree reduction in LDS (defined after the quad-lane BLAKE3 below): a
// workgroup's `cnt` digests of level `lvl` (cnt <= SUB_CNT, a power of two)
// sit in `a` as a[k * SUB_CNT + i] = word k of digest i, the workgroup's
// first digest being digest `first` of its level; `b` is laid out the same
// way (8 * SUB_CNT words). Builds levels lvl + 1 .. lvl + nlev (cnt >> nlev
// >= 1) of the level-contiguous tree whose level 0 at `tree` has `leaves`
// digests, writing every node. Every thread of the workgroup must call it.
constexpr int SUB_CNT = WG;
constexpr int SUB_LDS_WORDS = 2 * 8 * SUB_CNT;
__device__ void subtree_reduce(uint32_t* a, uint32_t* b, int cnt, int nlev, uint32_t* tree, uint64_t leaves, int lvl,
                               uint64_t first);

// ---------------------------------------------------------------------------
// Merkle leaves: thread g hashes leaves 2g and 2g+1 and, when asked, their
// parent (one node of level 1); contiguous kinds then reduce their
// workgroup's 256 parents further in LDS (sub_levels). Each leaf kind has
// its own fully unrolled body (static register indexing).
template <int KIND>
__device__ __forceinline__ void leaf_message(const LeavesJob& J, uint64_t leaf, uint32_t (&m)[16]) {
    if (KIND == LEAF_BASE) {
        const u32x4 v = ld4(J.src, leaf);
#pragma unroll
        for (int k = 0; k < 16; ++k) m[k] = 0;
        m[0] = v.x;
        m[4] = v.y;
        m[8] = v.z;
        m[12] = v.w;
    } else if (KIND == LEAF_TRACE) {
        // Leaf L packs natural rows 8L..8L+7 as A||B LE words; row
        // i = j + B t lives at brev(j)*n + brev(t).
        const int logn = J.logn, logb = J.logb;
        const uint32_t bmask = (1u << logb) - 1;
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const uint64_t i = leaf * 8 + r;
            const uint32_t j = uint32_t(i) & bmask;
            const uint32_t t = uint32_t(i >> logb);
            const uint64_t pos = (uint64_t(brev_bits(j, logb)) << logn) | brev_bits(t, logn);
            m[2 * r] = gc(J.src)[pos];
            m[2 * r + 1] = gc(J.src_b)[pos];
        }
    } else {  // LEAF_EXT, LEAF_RAW: 64 contiguous bytes
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const u32x4 v = ld4(J.src, leaf * 4 + k);
            m[4 * k] = v.x;
            m[4 * k + 1] = v.y;
            m[4 * k + 2] = v.z;
            m[4 * k + 3] = v.w;
        }
    }
}

// Trace leaves: thread tau is mapped to its leaf pair so that lanes read
// consecutive positions of each coset block.
__device__ __forceinline__ uint64_t trace_pair_of(const LeavesJob& J, uint64_t tau) {
    const int logn = J.logn, logb = J.logb;
    if (logb >= 4) {
        const int logh = logb - 4;  // pairs per step
        uint64_t h = tau >> logn, q = tau & ((uint64_t(1) << logn) - 1);
        return (uint64_t(brev_bits(uint32_t(q), logn)) << logh) | h;
    }
    const int logs = 4 - logb;  // steps per pair
    return brev_bits(uint32_t(tau), logn - logs);
}

__device__ __forceinline__ void load_digest(const uint32_t* src, uint32_t d[8]) {
    const u32x4 x = ld4(src, 0), y = ld4(src, 1);
    d[0] = x.x; d[1] = x.y; d[2] = x.z; d[3] = x.w;
    d[4] = y.x; d[5] = y.y; d[6] = y.z; d[7] = y.w;
}

// The level-1 digest this thread produced (thread t of the workgroup owns
// level-1 digest first + t) goes to LDS, then the workgroup reduces
// sub_levels more levels. Called by every thread of the workgroup.
__device__ __forceinline__ void subtree_after_parent(uint32_t* sa, uint32_t* sb, const uint32_t (&p)[8], bool live,
                                                     uint32_t* tree, uint64_t leaves, uint64_t first, int sub_levels) {
    if (live) {
#pragma unroll
        for (int k = 0; k < 8; ++k) sa[k * SUB_CNT + threadIdx.x] = p[k];
    }
    __syncthreads();
    const uint64_t parents = leaves >> 1;
    const int cnt = int(parents - first < uint64_t(SUB_CNT) ? parents - first : uint64_t(SUB_CNT));
    subtree_reduce(sa, sb, cnt, sub_levels, tree, leaves, 1, first);
}

template <int KIND>
__device__ __forceinline__ void leaves_body(const LeavesJob& J, uint64_t tau, uint32_t* sa, uint32_t* sb) {
    uint32_t m[16], d0[8], d1[8], p[8];
    if (J.leaves == 1) {
        if (tau != 0) return;
        leaf_message<KIND>(J, 0, m);
        leaf64_words(m, d0);
        store_digest(J.levels, d0);
        return;
    }
    const uint64_t pairs = J.leaves >> 1;
    const bool live = tau < pairs;
    if (!live && J.sub_levels == 0) return;
    if (live) {
        const uint64_t g = KIND == LEAF_TRACE ? trace_pair_of(J, tau) : tau;
        // Both halves' loads are issued before the first compression.
        if (KIND == LEAF_NODES) {
            uint32_t s0[8], s1[8], s2[8], s3[8];
            load_digest(J.src + (4 * g) * 8, s0);
            load_digest(J.src + (4 * g + 1) * 8, s1);
            load_digest(J.src + (4 * g + 2) * 8, s2);
            load_digest(J.src + (4 * g + 3) * 8, s3);
            node_words(s0, s1, d0);
            node_words(s2, s3, d1);
        } else {
            uint32_t m1[16];
            leaf_message<KIND>(J, 2 * g, m);
            leaf_message<KIND>(J, 2 * g + 1, m1);
            leaf64_words(m, d0);
            leaf64_words(m1, d1);
        }
        store_digest(J.levels + (2 * g) * 8, d0);
        store_digest(J.levels + (2 * g + 1) * 8, d1);
        if (J.with_parent) {
            node_words(d0, d1, p);
            store_digest(J.levels + (J.leaves + g) * 8, p);
        }
    }
    if (KIND != LEAF_TRACE && J.sub_levels > 0)
        subtree_after_parent(sa, sb, p, live, J.levels, J.leaves, tau - threadIdx.x, J.sub_levels);
}

// Trace leaves with the linked column computed in the same pass: a thread's
// two leaves are 16 natural rows i = j + B t (scattered over the coset
// blocks); for each it loads A at the row and at the next row (same coset,
// t + 1), forms B with one 16-way batched inversion of x - w_last, stores B
// and hashes the rows as A||B words.
__device__ __forceinline__ void leaves_linked(const LeavesJob& J, uint64_t tau) {
    const uint64_t pairs = J.leaves >> 1;
    if (tau >= pairs) return;
    const uint64_t g = trace_pair_of(J, tau);
    const int logn = J.logn, logb = J.logb;
    const uint32_t bmask = (1u << logb) - 1, nmask = (1u << logn) - 1;
    const gcptr A = gc(J.src);
    // One leaf (8 rows) at a time: its loads, its B, its hash; fewer values
    // live at once than with all 16 rows loaded up front, so more waves fit.
    uint32_t d0[8], d1[8];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        uint32_t a[8], an[8], iw[8], pos[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const uint64_t i = g * 16 + 8 * h + r;
            const uint32_t j = uint32_t(i) & bmask;
            const uint32_t t = uint32_t(i >> logb);
            const uint32_t blk = brev_bits(j, logb) << logn;
            pos[r] = blk | brev_bits(t, logn);
            a[r] = A[pos[r]];
            an[r] = A[blk | brev_bits((t + 1) & nmask, logn)];
            iw[r] = gc(J.iw)[pos[r]];
        }
        uint32_t m[16];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const uint32_t j = uint32_t(g * 16 + 8 * h + r) & bmask;
            const uint32_t b = add(an[r], mont_mul(iw[r], mont_mul(J.coeff, gc(J.zvan_m)[j])));
            gm(J.b_out)[pos[r]] = b;
            m[2 * r] = a[r];
            m[2 * r + 1] = b;
        }
        leaf64_words(m, h == 0 ? d0 : d1);
        __builtin_amdgcn_sched_barrier(0);  // the next leaf's loads stay after this hash
    }
    store_digest(J.levels + (2 * g) * 8, d0);
    store_digest(J.levels + (2 * g + 1) * 8, d1);
    if (J.with_parent) {
        uint32_t p[8];
        node_words(d0, d1, p);
        store_digest(J.levels + (J.leaves + g) * 8, p);
    }
}

// Range AIR row leaf: BLAKE3(0x00 || w_0 || .. || w_{C-1}) over the C column
// words at LDE position p, 4C + 1 bytes (one block for C <= 15, two above).
// The 0x00 prefix shifts every word by one byte: message word k is
// (w_{k-1} >> 24) | (w_k << 8).
constexpr int ROW_MAX_COLS = RANGE_MAX_BITS + 1;
__device__ __forceinline__ void row_leaf(const LeavesJob& J, uint64_t p, uint32_t (&d)[8]) {
    const int C = J.ncols;
    uint32_t w[ROW_MAX_COLS];
#pragma unroll
    for (int c = 0; c < ROW_MAX_COLS; ++c) w[c] = c < C ? gc(J.src)[uint64_t(c) * J.col_stride + p] : 0u;
    uint32_t m[32];
#pragma unroll
    for (int k = 0; k < 32; ++k) {
        const uint32_t lo = k >= 1 && k - 1 < ROW_MAX_COLS ? w[k - 1] >> 24 : 0u;
        const uint32_t hi = k < ROW_MAX_COLS ? w[k] << 8 : 0u;
        m[k] = lo | hi;
    }
    const uint32_t len = 4u * uint32_t(C) + 1u;
    uint32_t iv[8];
    b3_iv(iv);
    if (len <= 64) {
        compress8(iv, m, 0, 0, len, b3::CHUNK_START | b3::CHUNK_END | b3::ROOT, d);
    } else {
        uint32_t cv[8];
        compress8(iv, m, 0, 0, 64, b3::CHUNK_START, cv);
        compress8(cv, m + 16, 0, 0, len - 64, b3::CHUNK_END | b3::ROOT, d);
    }
}

// Thread tau < m/2 hashes the rows at positions tau and tau + m/2: with
// p = brev(j) n + brev(t) and B >= 2 cosets, they are natural rows
// i = j + B t (j even) and i + 1, a leaf pair, so the thread also forms
// their parent. Column loads are coalesced; digests land at natural indices.
__device__ __forceinline__ void leaves_rows(const LeavesJob& J, uint64_t tau) {
    const uint64_t half = J.leaves >> 1;
    if (tau >= half) return;
    const int logn = J.logn, logb = J.logb;
    const uint32_t nmask = (1u << logn) - 1;
    const uint32_t j = brev_bits(uint32_t(tau >> logn), logb);
    const uint32_t t = brev_bits(uint32_t(tau) & nmask, logn);
    const uint64_t i = uint64_t(j) + (uint64_t(t) << logb);
    uint32_t d0[8], d1[8];
    row_leaf(J, tau, d0);
    row_leaf(J, tau + half, d1);
    store_digest(J.levels + i * 8, d0);
    store_digest(J.levels + (i + 1) * 8, d1);
    if (J.with_parent) {
        uint32_t p[8];
        node_words(d0, d1, p);
        store_digest(J.levels + (J.leaves + i / 2) * 8, p);
    }
}

// Three kernels, so the light leaf kinds do not inherit the register
// footprint (and occupancy) of the trace rows with the linked column.
__global__ __launch_bounds__(WG) void k_leaves(LeavesJob inl, const LeavesJob* tab, uint32_t njobs) {
    __shared__ uint32_t sub[SUB_LDS_WORDS];
    uint32_t local;
    const LeavesJob J = job_of(inl, tab, njobs, local);
    const uint64_t tau = uint64_t(local) * WG + threadIdx.x;
    uint32_t* sa = sub;
    uint32_t* sb = sub + 8 * SUB_CNT;
    if (J.kind == LEAF_TRACE) leaves_body<LEAF_TRACE>(J, tau, sa, sb);
    else if (J.kind == LEAF_BASE) leaves_body<LEAF_BASE>(J, tau, sa, sb);
    else if (J.kind == LEAF_NODES) leaves_body<LEAF_NODES>(J, tau, sa, sb);
    else leaves_body<LEAF_EXT>(J, tau, sa, sb);
}

__global__ __launch_bounds__(WG) void k_leaves_linked(LeavesJob inl, const LeavesJob* tab, uint32_t njobs) {
    uint32_t local;
    const LeavesJob J = job_of(inl, tab, njobs, local);
    leaves_linked(J, uint64_t(local) * WG + threadIdx.x);
}

__global__ __launch_bounds__(WG) void k_leaves_rows(LeavesJob inl, const LeavesJob* tab, uint32_t njobs) {
    uint32_t local;
    const LeavesJob J = job_of(inl, tab, njobs, local);
    leaves_rows(J, uint64_t(local) * WG + threadIdx.x);
}

// Wide levels: thread g reduces digests 2g, 2g+1 (fuse = 1) or 4g..4g+3
// (fuse = 2: two nodes and their parent).
__global__ __launch_bounds__(WG) void k_levels(LevelsJob inl, const LevelsJob* tab, uint32_t njobs) {
    uint32_t local;
    const LevelsJob J = job_of(inl, tab, njobs, local);
    const uint64_t g = uint64_t(local) * WG + threadIdx.x;
    const uint64_t count = J.count;
    uint32_t* next = J.level + count * 8;
    uint32_t a[8], b[8];
    if (J.fuse == 2) {
        if (g >= count / 4) return;
        // All four digests' loads before the first compression.
        uint32_t c[8], d[8], p0[8], p1[8], top[8];
        load_digest(J.level + (4 * g) * 8, a);
        load_digest(J.level + (4 * g + 1) * 8, b);
        load_digest(J.level + (4 * g + 2) * 8, c);
        load_digest(J.level + (4 * g + 3) * 8, d);
        node_words(a, b, p0);
        node_words(c, d, p1);
        store_digest(next + (2 * g) * 8, p0);
        store_digest(next + (2 * g + 1) * 8, p1);
        node_words(p0, p1, top);
        store_digest(next + (count / 2 + g) * 8, top);
        return;
    }
    if (g >= count / 2) return;
    uint32_t p[8];
    load_digest(J.level + (2 * g) * 8, a);
    load_digest(J.level + (2 * g + 1) * 8, b);
    node_words(a, b, p);
    store_digest(next + g * 8, p);
}

// Layer values staged in LDS for hashing: 2048 Ext4 values per workgroup,
// one padding word per 8 values so that a thread's 8 consecutive values (two
// quad leaves, 32 words) start 33 words after its neighbour's: conflict-free
// across lanes.
constexpr int FUSE_VALUES = 8 * WG;
constexpr int FUSE_LDS = FUSE_VALUES * 4 + FUSE_VALUES / 8;
static_assert(FUSE_LDS >= SUB_LDS_WORDS, "staged values' LDS holds the subtree buffers");
__device__ __forceinline__ int fuse_at(int q, int k) { return q * 4 + k + (q >> 3); }

// Thread t hashes staged values 8t..8t+7 (leaves 2t, 2t+1 of the workgroup's
// first leaf `leaf0`) and, with_parent, their parent into the next level.
__device__ __forceinline__ bool hash_staged(const uint32_t* lds, uint32_t* tree, uint64_t leaves, uint64_t leaf0,
                                            int with_parent, int nvals, uint32_t (&par)[8]) {
    const int t = threadIdx.x;
    if (8 * t >= nvals) return false;
    uint32_t m0[16], m1[16];
#pragma unroll
    for (int w = 0; w < 16; ++w) {
        m0[w] = lds[fuse_at(8 * t + (w >> 2), w & 3)];
        m1[w] = lds[fuse_at(8 * t + 4 + (w >> 2), w & 3)];
    }
    uint32_t dg0[8], dg1[8];
    leaf64_words(m0, dg0);
    leaf64_words(m1, dg1);
    const uint64_t leaf = leaf0 + 2 * uint64_t(t);
    store_digest(tree + leaf * 8, dg0);
    store_digest(tree + (leaf + 1) * 8, dg1);
    if (with_parent) {
        node_words(dg0, dg1, par);
        store_digest(tree + (leaves + leaf / 2) * 8, par);
    }
    return true;
}

// After hash_staged: the staged values are no longer needed, so the same
// LDS holds the subtree buffers (a barrier first: other threads may still
// read their staged values).
__device__ __forceinline__ void staged_subtree(uint32_t* lds, const uint32_t (&par)[8], bool live, uint32_t* tree,
                                               uint64_t leaves, uint64_t leaf0, int sub_levels) {
    __syncthreads();
    subtree_after_parent(lds, lds + 8 * SUB_CNT, par, live, tree, leaves, leaf0 / 2, sub_levels);
}

// Composition fused with layer 0's Merkle leaves: the workgroup computes
// 2048 consecutive layer positions as k_compose does (coalesced), stages
// them in LDS, then hashes them as 512 quad leaves (and their parents).
// The composition of this workgroup's 2048 positions (CELT per thread),
// staged in LDS and written as layer 0. FULL: every position exists (all
// but a layer's last, partial workgroup), so no per-lane guard splits the
// loads from the arithmetic.
template <bool FULL>
__device__ __forceinline__ void compose_stage(const ComposeJob& J, const ComposeConsts& C, uint32_t* lds, uint32_t base,
                                              uint32_t m) {
    const gcptr A = gc(J.a), Bv = gc(J.b);
    uint32_t va[CELT], vb[CELT], na[CELT], nb[CELT], i1[CELT], iw[CELT], iz[CELT];
#pragma unroll
    for (int e = 0; e < CELT; ++e) {
        const uint32_t p = base + e * WG;
        if (FULL || p < m) {
            const PosInfo pi = pos_info(p, J.logn, J.logb);
            na[e] = A[pi.pnext];
            nb[e] = Bv[pi.pnext];
            va[e] = A[p];
            vb[e] = Bv[p];
            i1[e] = gc(J.pinv.i1)[p];
            iw[e] = gc(J.pinv.iw)[p];
            iz[e] = gc(J.pinv.iz)[p];
        }
    }
    static_for<0, CELT>([&](auto ec) {
        constexpr int e = decltype(ec)::value;
        const uint32_t p = base + e * WG;
        if (!FULL && p >= m) return;
        uint32_t o[4];
        compose_value(C, va[e], vb[e], na[e], nb[e], i1[e], iw[e], iz[e], o);
#pragma unroll
        for (int k = 0; k < 4; ++k) lds[fuse_at(int(threadIdx.x) + e * WG, k)] = o[k];
        st4(J.layer0, p, u32x4{o[0], o[1], o[2], o[3]});
    });
}
__global__ __launch_bounds__(WG) void k_compose_leaves(ComposeJob inl, const ComposeJob* tab, uint32_t njobs) {
    __shared__ uint32_t lds[FUSE_LDS];
    uint32_t local;
    const ComposeJob J = job_of(inl, tab, njobs, local);
    const uint32_t m = 1u << (J.logn + J.logb);
    const uint32_t first = local * FUSE_VALUES;
    const uint32_t base = first + threadIdx.x;
    const ComposeConsts C = *J.c;
    if (first + uint32_t(FUSE_VALUES) <= m) compose_stage<true>(J, C, lds, base, m);
    else compose_stage<false>(J, C, lds, base, m);
    __syncthreads();
    const int nvals = int(m - first < uint32_t(FUSE_VALUES) ? m - first : uint32_t(FUSE_VALUES));
    uint32_t par[8];
    const bool live = hash_staged(lds, J.tree, m >> 2, first >> 2, J.with_parent, nvals, par);
    if (J.sub_levels > 0) staged_subtree(lds, par, live, J.tree, m >> 2, first >> 2, J.sub_levels);
}

// Batched verification: lane i walks one Merkle path from its 64-byte leaf
// (one path per lane; paths are at most 63 levels, checked by the host).
__global__ __launch_bounds__(WG) void k_verify_paths(const uint32_t* checks, const uint32_t* siblings,
                                                     uint32_t count, uint8_t* ok) {
    const uint32_t i = blockIdx.x * WG + threadIdx.x;
    if (i >= count) return;
    const gcptr c = gc(checks) + size_t(i) * PATH_CHECK_WORDS;
    uint32_t m[16], h[8];
#pragma unroll
    for (int k = 0; k < 16; ++k) m[k] = c[k];
    leaf64_words(m, h);
    const uint32_t steps = c[24], first = c[25];
    const uint64_t dirs = uint64_t(c[26]) | (uint64_t(c[27]) << 32);
    for (uint32_t l = 0; l < steps; ++l) {
        const gcptr sp = gc(siblings) + (size_t(first) + l) * 8;
        uint32_t sw[8], o[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) sw[k] = sp[k];
        if ((dirs >> l) & 1)
            node_words(h, sw, o);
        else
            node_words(sw, h, o);
#pragma unroll
        for (int k = 0; k < 8; ++k) h[k] = o[k];
    }
    bool eq = true;
#pragma unroll
    for (int k = 0; k < 8; ++k) eq = eq && h[k] == c[16 + k];
    ok[i] = eq ? 1 : 0;
}

// ---- quad-lane BLAKE3 for the narrow tree levels ------------------------
// Four lanes of a quad compute one compression: lane q holds state column q
// (v[q], v[4+q], v[8+q], v[12+q]) and runs the column G on it; for the
// diagonal G it borrows v[4+(q+1)%4], v[8+(q+2)%4], v[12+(q+3)%4] from its
// neighbours by DPP quad permutations and gives them back afterwards. The
// chaining value enters and leaves as words (q, 4+q) of lane q. A narrow
// level is latency-bound, and this cuts a compression's dependent
// instruction chain to under half. Control flow must be uniform per quad.
template <int CTRL>
__device__ __forceinline__ uint32_t quad_dpp(uint32_t v) {
    return uint32_t(__builtin_amdgcn_mov_dpp(int(v), CTRL, 0xF, 0xF, true));
}
constexpr int QP_NEXT1 = 0x39;  // lane q reads lane (q + 1) & 3
constexpr int QP_NEXT2 = 0x4E;  // lane q reads lane (q + 2) & 3
constexpr int QP_NEXT3 = 0x93;  // lane q reads lane (q + 3) & 3

__device__ __forceinline__ uint32_t sel4(bool q_odd, bool q_hi, uint32_t v0, uint32_t v1, uint32_t v2, uint32_t v3) {
    const uint32_t lo = q_odd ? v1 : v0, hi = q_odd ? v3 : v2;
    return q_hi ? hi : lo;
}

// Message word of slot (column: k = 0, 1 -> 2q + k; diagonal: k = 2, 3 ->
// 8 + 2q + k - 2) in round r, picked from the 16 words every lane holds.
template <int RND, int K>
__device__ __forceinline__ uint32_t quad_msg(const uint32_t (&m)[16], bool q_odd, bool q_hi) {
    constexpr B3Schedule S;
    constexpr int s0 = K < 2 ? K : 8 + K - 2;
    return sel4(q_odd, q_hi, m[S.idx[RND][s0]], m[S.idx[RND][s0 + 2]], m[S.idx[RND][s0 + 4]], m[S.idx[RND][s0 + 6]]);
}

#define Z0P_Q4_G(x, y)         \
    a = a + b + (x);           \
    d = rotr(d ^ a, 16);       \
    c = c + d;                 \
    b = rotr(b ^ c, 12);       \
    a = a + b + (y);           \
    d = rotr(d ^ a, 8);        \
    c = c + d;                 \
    b = rotr(b ^ c, 7);

template <int RND>
__device__ __forceinline__ void quad_round(uint32_t& a, uint32_t& b, uint32_t& c, uint32_t& d, const uint32_t (&m)[16],
                                           bool q_odd, bool q_hi) {
    const uint32_t m0 = quad_msg<RND, 0>(m, q_odd, q_hi), m1 = quad_msg<RND, 1>(m, q_odd, q_hi);
    const uint32_t m2 = quad_msg<RND, 2>(m, q_odd, q_hi), m3 = quad_msg<RND, 3>(m, q_odd, q_hi);
    Z0P_Q4_G(m0, m1);
    b = quad_dpp<QP_NEXT1>(b);
    c = quad_dpp<QP_NEXT2>(c);
    d = quad_dpp<QP_NEXT3>(d);
    Z0P_Q4_G(m2, m3);
    b = quad_dpp<QP_NEXT3>(b);
    c = quad_dpp<QP_NEXT2>(c);
    d = quad_dpp<QP_NEXT1>(d);
}

// One compression; (ha, hb) = chaining-value words (q, 4 + q) in and out.
__device__ __forceinline__ void quad_compress(uint32_t& ha, uint32_t& hb, const uint32_t (&m)[16], uint32_t q,
                                              uint32_t block_len, uint32_t flags) {
    const bool q_odd = (q & 1) != 0, q_hi = (q & 2) != 0;
    uint32_t a = ha, b = hb;
    uint32_t c = sel4(q_odd, q_hi, b3::IV0, b3::IV1, b3::IV2, b3::IV3);
    uint32_t d = sel4(q_odd, q_hi, 0u, 0u, block_len, flags);  // counter 0
    quad_round<0>(a, b, c, d, m, q_odd, q_hi);
    quad_round<1>(a, b, c, d, m, q_odd, q_hi);
    quad_round<2>(a, b, c, d, m, q_odd, q_hi);
    quad_round<3>(a, b, c, d, m, q_odd, q_hi);
    quad_round<4>(a, b, c, d, m, q_odd, q_hi);
    quad_round<5>(a, b, c, d, m, q_odd, q_hi);
    quad_round<6>(a, b, c, d, m, q_odd, q_hi);
    ha = a ^ c;
    hb = b ^ d;
}

// Node digest BLAKE3(0x01 || l || r) by a quad (node_words, four lanes):
// returns words (q, 4 + q) of the digest in (oa, ob).
__device__ __forceinline__ void quad_node(const uint32_t (&l)[8], const uint32_t (&r)[8], uint32_t q, uint32_t& oa,
                                          uint32_t& ob) {
    uint32_t m[16];
    m[0] = 0x01u | (l[0] << 8);
#pragma unroll
    for (int k = 1; k < 8; ++k) m[k] = (l[k - 1] >> 24) | (l[k] << 8);
    m[8] = (l[7] >> 24) | (r[0] << 8);
#pragma unroll
    for (int k = 9; k < 16; ++k) m[k] = (r[k - 9] >> 24) | (r[k - 8] << 8);
    const bool q_odd = (q & 1) != 0, q_hi = (q & 2) != 0;
    uint32_t ha = sel4(q_odd, q_hi, b3::IV0, b3::IV1, b3::IV2, b3::IV3);
    uint32_t hb = sel4(q_odd, q_hi, b3::IV4, b3::IV5, b3::IV6, b3::IV7);
    quad_compress(ha, hb, m, q, 64, b3::CHUNK_START);
    uint32_t m2[16];
    m2[0] = r[7] >> 24;
#pragma unroll
    for (int k = 1; k < 16; ++k) m2[k] = 0;
    quad_compress(ha, hb, m2, q, 1, b3::CHUNK_END | b3::ROOT);
    oa = ha;
    ob = hb;
}

__device__ void subtree_reduce(uint32_t* a, uint32_t* b, int cnt, int nlev, uint32_t* tree, uint64_t leaves, int lvl,
                               uint64_t first) {
    const int t = threadIdx.x;
    uint32_t* src = a;
    uint32_t* dst = b;
    uint64_t f = first;
    for (int r = 1; r <= nlev; ++r) {
        const int nodes = cnt >> r;
        f >>= 1;
        uint32_t* out = tree + tree_level_words(leaves, lvl + r) + f * 8;
        if (nodes > WG / 8) {
            // Wide: one lane per node (the fewest instructions per node).
            for (int i = t; i < nodes; i += WG) {
                uint32_t L[8], R[8], o[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    L[k] = src[k * SUB_CNT + 2 * i];
                    R[k] = src[k * SUB_CNT + 2 * i + 1];
                }
                node_words(L, R, o);
#pragma unroll
                for (int k = 0; k < 8; ++k) dst[k * SUB_CNT + i] = o[k];
                store_digest(out + uint64_t(i) * 8, o);
            }
        } else {
            // Narrow: a quad of lanes per node cuts the dependent chain.
            const int g = t >> 2;
            const uint32_t q = uint32_t(t & 3);
            if (g < nodes) {  // quad-uniform
                uint32_t L[8], R[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    L[k] = src[k * SUB_CNT + 2 * g];
                    R[k] = src[k * SUB_CNT + 2 * g + 1];
                }
                uint32_t oa, ob;
                quad_node(L, R, q, oa, ob);
                dst[q * SUB_CNT + g] = oa;
                dst[(4 + q) * SUB_CNT + g] = ob;
                gm(out)[uint64_t(g) * 8 + q] = oa;
                gm(out)[uint64_t(g) * 8 + 4 + q] = ob;
            }
        }
        __syncthreads();
        uint32_t* tmp = src;
        src = dst;
        dst = tmp;
    }
}

// Narrow levels: up to group_log levels per 2^group_log-digest group in LDS,
// one quad per parent (2^(group_log+1) threads: every parent of the first
// level at once). Active quads stay packed in the lowest waves, so a narrow
// level occupies as few waves as it can; small groups spread a level over
// many compute units, large ones need fewer launches.
constexpr int TAIL_MAX_SUB = 1 << TAIL_GROUP_MAX_LOG;

__global__ __launch_bounds__(2 * TAIL_MAX_SUB) void k_tail(TailJob inl, const TailJob* tab, uint32_t njobs) {
    __shared__ uint32_t bufA[8][TAIL_MAX_SUB];
    __shared__ uint32_t bufB[8][TAIL_MAX_SUB];
    uint32_t local;
    const TailJob J = job_of(inl, tab, njobs, local);
    const uint64_t sub = uint64_t(1) << J.group_log;
    const uint64_t first = uint64_t(local) * sub;
    const int here = int(J.count - first < sub ? J.count - first : sub);
    const int tid = threadIdx.x;
    const gcptr level_in = gc(J.level);
    for (int w = tid; w < here * 8; w += int(blockDim.x)) bufA[w & 7][w >> 3] = level_in[(first * 8) + w];
    __syncthreads();
    uint32_t(*cur)[TAIL_MAX_SUB] = bufA;
    uint32_t(*nxt)[TAIL_MAX_SUB] = bufB;
    const uint32_t q = uint32_t(tid & 3);
    const int g = tid >> 2;
    uint32_t* lvl = J.level;
    uint64_t lc = J.count, f = first;
    int count = here;
    for (int l = 1; l <= J.nlevels && count > 1; ++l) {
        uint32_t* next_lvl = lvl + lc * 8;
        const int pc = count >> 1;
        if (g < pc) {  // quad-uniform
            uint32_t L[8], R[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                L[k] = cur[k][2 * g];
                R[k] = cur[k][2 * g + 1];
            }
            uint32_t oa, ob;
            quad_node(L, R, q, oa, ob);
            nxt[q][g] = oa;
            nxt[4 + q][g] = ob;
            const gptr dst = gm(next_lvl) + (f / 2 + uint64_t(g)) * 8;
            dst[q] = oa;
            dst[4 + q] = ob;
        }
        __syncthreads();
        auto tmp = cur;
        cur = nxt;
        nxt = tmp;
        count = pc;
        lvl = next_lvl;
        lc >>= 1;
        f >>= 1;
    }
}

// ---------------------------------------------------------------------------
// Arity-4 fold of a bit-reversed layer (fri.rs:1273-1331). Quad i sits at
// (x, -x, j0 x, -j0 x), x the point of position 4i;
// x^{-1} = 31^{-4^d} * Omega_m^{-(brev(i) << 2d)}.
// The operands of output i of the fold of `layer` (length 2^log_len): its
// quad and its entry of the inverse-point table.
struct FoldIn {
    E4 q[4];
    uint32_t fi;
};
template <bool BASE>
__device__ __forceinline__ FoldIn fold_load(const uint32_t* layer, const uint32_t* finv, uint64_t i) {
    FoldIn r;
    r.fi = gc(finv)[i];
    if (BASE) {
        const u32x4 v = ld4(layer, i);
        r.q[0] = e4_base(v.x); r.q[1] = e4_base(v.y); r.q[2] = e4_base(v.z); r.q[3] = e4_base(v.w);
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const u32x4 v = ld4(layer, i * 4 + k);
            r.q[k] = E4{{v.x, v.y, v.z, v.w}};
        }
    }
    return r;
}
__device__ __forceinline__ E4 fold_compute(const FoldIn& in, const FoldConsts& C) {
    const uint32_t inv_x = mont_mul(in.fi, C.ginv_m);
    const E4* q = in.q;
    const E4 s01 = e4_add(q[0], q[1]), s23 = e4_add(q[2], q[3]);
    const E4 dodd = e4_sub(q[0], q[1]);
    const E4 dskew = e4_scale_m(e4_sub(q[2], q[3]), C.unit_m);
    const E4 M = e4_add(s01, s23), O = e4_sub(dodd, dskew), Ev = e4_sub(s01, s23), S = e4_add(dodd, dskew);
    E4 acc = e4_scale_m(e4_mul_mat(S, C.cm[2]), inv_x);
    acc = e4_scale_m(e4_add(acc, e4_mul_mat(Ev, C.cm[1])), inv_x);
    acc = e4_add(acc, e4_mul_mat(O, C.cm[0]));
    return e4_add(e4_scale_m(M, C.inv4_m), e4_scale_m(acc, inv_x));
}
__device__ __forceinline__ E4 fold_one(const uint32_t* layer, int base_field, const uint32_t* finv,
                                        const FoldConsts& C, uint64_t i) {
    return base_field ? fold_compute(fold_load<true>(layer, finv, i), C)
                      : fold_compute(fold_load<false>(layer, finv, i), C);
}

// Two outputs per thread, i and i + WG (i < quads): both outputs' loads are
// issued before the first one's arithmetic. The second's index is clamped
// into the layer (tiny layers), so no branch separates the loads from the
// arithmetic, and its result is stored only when it exists.
constexpr int FOLD_PER = 2;
template <bool BASE>
__device__ __forceinline__ void fold_two(const FoldJob& J, const FoldConsts& C, uint64_t i, uint64_t quads) {
    const uint64_t i2 = i + WG;
    const FoldIn a = fold_load<BASE>(J.layer, J.finv, i);
    const FoldIn b = fold_load<BASE>(J.layer, J.finv, i2 < quads ? i2 : i);
    const E4 x = fold_compute(a, C);
    st4(J.next, i, u32x4{x.v[0], x.v[1], x.v[2], x.v[3]});
    const E4 y = fold_compute(b, C);
    if (i2 < quads) st4(J.next, i2, u32x4{y.v[0], y.v[1], y.v[2], y.v[3]});
}

__global__ __launch_bounds__(WG) void k_fold(FoldJob inl, const FoldJob* tab, uint32_t njobs) {
    uint32_t local;
    const FoldJob J = job_of(inl, tab, njobs, local);
    const uint64_t quads = uint64_t(1) << (J.log_len - 2);
    const uint64_t i = uint64_t(local) * (WG * FOLD_PER) + threadIdx.x;
    if (i >= quads) return;
    const FoldConsts C = *J.c;
    if (J.base_field) fold_two<true>(J, C, i, quads);
    else fold_two<false>(J, C, i, quads);
}

// Fold fused with the next layer's Merkle leaves: the workgroup folds 2048
// consecutive outputs as k_fold does (coalesced quads), stages them in LDS,
// then hashes them as 512 quad leaves of layer d+1 (and their parents). The
// next output's loads are issued before each output's arithmetic (indices
// clamped into the layer, results kept only for outputs that exist).
template <bool BASE, bool FULL>
__device__ __forceinline__ void fold_stage(const FoldJob& J, const FoldConsts& C, uint32_t* lds, uint64_t first,
                                           uint64_t outs) {
    const uint64_t i0 = first + threadIdx.x;
    FoldIn cur = fold_load<BASE>(J.layer, J.finv, i0 < outs ? i0 : outs - 1);
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const uint64_t i = i0 + uint64_t(e) * WG;
        FoldIn nxt;
        if (e < 7) {
            const uint64_t in = i + WG;
            nxt = fold_load<BASE>(J.layer, J.finv, in < outs ? in : outs - 1);
        }
        const E4 v = fold_compute(cur, C);
        if (FULL || i < outs) {
#pragma unroll
            for (int k = 0; k < 4; ++k) lds[fuse_at(int(threadIdx.x) + e * WG, k)] = v.v[k];
            st4(J.next, i, u32x4{v.v[0], v.v[1], v.v[2], v.v[3]});
        }
        if (e < 7) cur = nxt;
    }
}
__global__ __launch_bounds__(WG) void k_fold_leaves(FoldJob inl, const FoldJob* tab, uint32_t njobs) {
    __shared__ uint32_t lds[FUSE_LDS];
    uint32_t local;
    const FoldJob J = job_of(inl, tab, njobs, local);
    const uint64_t outs = uint64_t(1) << (J.log_len - 2);
    const uint64_t first = uint64_t(local) * FUSE_VALUES;
    const FoldConsts C = *J.c;
    // FULL: all 2048 outputs of this workgroup exist (every workgroup but a
    // small layer's last).
    if (first + FUSE_VALUES <= outs) {
        if (J.base_field) fold_stage<true, true>(J, C, lds, first, outs);
        else fold_stage<false, true>(J, C, lds, first, outs);
    } else {
        if (J.base_field) fold_stage<true, false>(J, C, lds, first, outs);
        else fold_stage<false, false>(J, C, lds, first, outs);
    }
    __syncthreads();
    const int nvals = int(outs - first < uint64_t(FUSE_VALUES) ? outs - first : uint64_t(FUSE_VALUES));
    uint32_t par[8];
    const bool live = hash_staged(lds, J.tree, outs >> 2, first >> 2, J.with_parent, nvals, par);
    if (J.sub_levels > 0) staged_subtree(lds, par, live, J.tree, outs >> 2, first >> 2, J.sub_levels);
}

// `nlev` levels above `count` digests held in LDS as cur[k][i] (word k of
// digest i; the first is digest `first` of level `lvl` of the tree whose
// level 0 has `leaves` digests), one quad of lanes per node, every level
// written to the tree; the digests of the last level end in cur again when
// nlev is even, else in nxt. Called by every thread of the workgroup.
__device__ __forceinline__ void quad_levels(uint32_t (*cur)[64], uint32_t (*nxt)[64], int count, int nlev,
                                            uint32_t* tree, uint64_t leaves, int lvl, uint64_t first) {
    const uint32_t q = threadIdx.x & 3u;
    for (int l = 0; l < nlev; ++l) {
        const int pc = count >> 1;
        uint32_t* out = tree + tree_level_words(leaves, lvl + l + 1) + (first >> (l + 1)) * 8;
        for (int g = int(threadIdx.x >> 2); g < pc; g += WG / 4) {  // quad-uniform
            uint32_t L[8], R[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                L[k] = cur[k][2 * g];
                R[k] = cur[k][2 * g + 1];
            }
            uint32_t oa, ob;
            quad_node(L, R, q, oa, ob);
            nxt[q][g] = oa;
            nxt[4 + q][g] = ob;
            gm(out)[uint64_t(g) * 8 + q] = oa;
            gm(out)[uint64_t(g) * 8 + 4 + q] = ob;
        }
        __syncthreads();
        auto tmp = cur;
        cur = nxt;
        nxt = tmp;
        count = pc;
    }
}

// Levels lvl+1 .. target of a tree reduced in stages of six levels: this
// workgroup holds digest `idx` of level lvl (count digests there) in LDS
// (word k at (in_b ? b : a)[k][0]); it publishes it, and the last workgroup
// to arrive of each group of 64 loads the group's digests and reduces them
// up to six levels, until the target level. `ctr` holds one zeroed counter
// per group per stage (tree_stage_counters) and is left zero. The digest a
// stage hands on is published with agent-scope atomic stores (sc1), waited
// for (s_waitcnt: acknowledged) before the same wave's counter update, and
// read with agent-scope atomic loads (sc1) after the counter says all have
// arrived: no workgroup writes back or invalidates whole caches, which
// thousands of device-scope fences would (buffer_wbl2 per workgroup
// measured ~3x slower on a 4096-workgroup layer). Called by every thread of
// the workgroup.
__device__ void tree_stages(uint32_t (*a)[64], uint32_t (*b)[64], bool in_b, uint32_t* tree, uint64_t leaves,
                            int lvl, uint64_t idx, uint64_t count, int target, uint32_t* ctr) {
    __shared__ uint32_t last;
    const uint32_t tid = threadIdx.x;
    while (lvl < target) {
        // Threads 0..7 and 0 are in wave 0: the counter update follows the
        // completed stores.
        const uint64_t group = idx >> FOLD_COMMIT_GROUP_LOG;
        const uint64_t gsize = count - group * 64 < 64 ? count - group * 64 : 64;
        uint32_t* dig = tree + tree_level_words(leaves, lvl) + idx * 8;
        if (tid < 8) __hip_atomic_store(dig + tid, (in_b ? b : a)[tid][0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // The stores complete (acknowledged at agent scope: they are sc1
        // stores) before the counter update below, in the same wave.
        if (tid < 64) __builtin_amdgcn_s_waitcnt(0);
        if (tid == 0)
            last = __hip_atomic_fetch_add(ctr + group * STAGE_COUNTER_WORDS, 1u, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT) ==
                           uint32_t(gsize - 1)
                       ? 1u
                       : 0u;
        __syncthreads();
        if (!last) return;  // workgroup-uniform
        uint32_t* src = tree + tree_level_words(leaves, lvl) + group * 64 * 8;
        for (int w = int(tid); w < int(gsize) * 8; w += WG)
            a[w & 7][w >> 3] = __hip_atomic_load(src + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const int nl = target - lvl < FOLD_COMMIT_GROUP_LOG ? target - lvl : FOLD_COMMIT_GROUP_LOG;
        quad_levels(a, b, int(gsize), nl, tree, leaves, lvl, group * 64);
        in_b = (nl & 1) != 0;
        if (tid == 0) ctr[group * STAGE_COUNTER_WORDS] = 0;  // left zero for the next use
        ctr += (count + 63) / 64 * STAGE_COUNTER_WORDS;
        lvl += nl;
        idx = group;
        count >>= nl;
    }
}

// Levels above `lvl` of a tree up to its target, in one launch: workgroup w
// loads digests 64w.. of level lvl (written by an earlier launch), reduces
// them up to six levels in LDS, then the staged top (tree_stages). Replaces
// the quad-lane tail launches of a tree's narrow levels.
__global__ __launch_bounds__(WG) void k_tree_top(TreeTopJob inl, const TreeTopJob* tab, uint32_t njobs) {
    __shared__ uint32_t bufA[8][64], bufB[8][64];
    uint32_t local;
    const TreeTopJob J = job_of(inl, tab, njobs, local);
    const uint64_t count = J.leaves >> J.lvl;
    const uint64_t first = uint64_t(local) * 64;
    const uint64_t here = count - first < 64 ? count - first : 64;
    const gcptr src = gc(J.tree + tree_level_words(J.leaves, J.lvl)) + first * 8;
    for (int w = int(threadIdx.x); w < int(here) * 8; w += WG) bufA[w & 7][w >> 3] = src[w];
    __syncthreads();
    int own = J.target - J.lvl < FOLD_COMMIT_GROUP_LOG ? J.target - J.lvl : FOLD_COMMIT_GROUP_LOG;
    quad_levels(bufA, bufB, int(here), own, J.tree, J.leaves, J.lvl, first);
    tree_stages(bufA, bufB, own & 1, J.tree, J.leaves, J.lvl + own, local, count >> own, J.target, J.counters);
}

// Lane q of a quad gets the value of lane K of its quad.
template <int K>
__device__ __forceinline__ uint32_t quad_bcast(uint32_t v) {
    return uint32_t(__builtin_amdgcn_mov_dpp(int(v), K * 0x55, 0xF, 0xF, true));
}

// One FRI layer's whole commitment in one launch (a FoldJob with counters):
// thread t folds output t of its workgroup's 256 (fri.rs:1273-1331); each quad
// of lanes then hashes its four outputs as one 64-byte leaf (lane q holds
// output q; DPP broadcasts give every lane the 16 message words), and the
// workgroup reduces its 64 leaves by six levels in LDS, a quad per node. The
// levels above are reduced in stages: the last workgroup to finish of each
// group of 64 (device-scope counter) takes the group's 64 digests six levels
// further, up to the cap level. The groups at one stage are independent, so
// wide trees keep many workgroups busy; a narrow layer costs one launch and
// ~1 us per level instead of a fold, a leaves and several tree launches.
// The digest a stage hands to the next one is published with agent-scope
// atomic stores, completed (workgroup-scope release: the stores are
// acknowledged) before the same wave's counter update, and read with
// agent-scope atomic loads: no workgroup writes back or invalidates whole
// caches, which thousands of device-scope fences would.
__global__ __launch_bounds__(WG) void k_fold_commit(FoldJob inl, const FoldJob* tab, uint32_t njobs) {
    __shared__ uint32_t bufA[8][64], bufB[8][64];
    uint32_t local;
    const FoldJob J = job_of(inl, tab, njobs, local);
    const uint64_t outs = uint64_t(1) << (J.log_len - 2);
    const uint64_t leaves = outs >> 2;
    const uint32_t tid = threadIdx.x, q = tid & 3u, g = tid >> 2;
    const uint64_t i = uint64_t(local) * WG + tid;  // outs is a multiple of 256
    const E4 v = fold_one(J.layer, J.base_field, J.finv, *J.c, i);
    st4(J.next, i, u32x4{v.v[0], v.v[1], v.v[2], v.v[3]});
    uint32_t m[16];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        m[c] = quad_bcast<0>(v.v[c]);
        m[4 + c] = quad_bcast<1>(v.v[c]);
        m[8 + c] = quad_bcast<2>(v.v[c]);
        m[12 + c] = quad_bcast<3>(v.v[c]);
    }
    const bool q_odd = (q & 1) != 0, q_hi = (q & 2) != 0;
    uint32_t ha = sel4(q_odd, q_hi, b3::IV0, b3::IV1, b3::IV2, b3::IV3);
    uint32_t hb = sel4(q_odd, q_hi, b3::IV4, b3::IV5, b3::IV6, b3::IV7);
    quad_compress(ha, hb, m, q, 64, b3::CHUNK_START | b3::CHUNK_END | b3::ROOT);  // hash.rs:388-397
    const uint64_t leaf = uint64_t(local) * 64 + g;
    gm(J.tree)[leaf * 8 + q] = ha;
    gm(J.tree)[leaf * 8 + 4 + q] = hb;
    bufA[q][g] = ha;
    bufA[4 + q][g] = hb;
    __syncthreads();
    const int own = J.target < FOLD_COMMIT_GROUP_LOG ? J.target : FOLD_COMMIT_GROUP_LOG;
    quad_levels(bufA, bufB, 64, own, J.tree, leaves, 0, uint64_t(local) * 64);
    tree_stages(bufA, bufB, own & 1, J.tree, leaves, own, local, leaves >> own, J.target, J.counters);
}

// ---------------------------------------------------------------------------
// A small FRI layer in one workgroup: fold layer d into layer d+1 (at most
// SMALL_MAX_VALUES values), hash its quad leaves, and build its tree up to
// the cap level in LDS. Thread t folds outputs 4t..4t+3, which are exactly
// leaf t, so the leaf message never leaves registers. Levels above the
// leaves use a quad of lanes per node. One dispatch replaces fold, leaves
// and tree kernels whose fixed costs dominate at these sizes.
constexpr int SMALL_WG = 1024;

__global__ __launch_bounds__(SMALL_WG) void k_small_commit(SmallCommitJob inl, const SmallCommitJob* tab,
                                                           uint32_t njobs) {
    __shared__ uint32_t bufA[8][SMALL_MAX_VALUES / 4];
    __shared__ uint32_t bufB[8][SMALL_MAX_VALUES / 8];
    uint32_t local;
    const SmallCommitJob J = job_of(inl, tab, njobs, local);
    const FoldConsts C = *J.c;
    const int leaves = int(J.leaves);
    const int t = threadIdx.x;
    if (t < leaves) {
        uint32_t m[16];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint64_t i = uint64_t(4 * t + k);
            const E4 out = fold_one(J.layer, J.base_field, J.finv, C, i);
            st4(J.next, i, u32x4{out.v[0], out.v[1], out.v[2], out.v[3]});
            m[4 * k] = out.v[0];
            m[4 * k + 1] = out.v[1];
            m[4 * k + 2] = out.v[2];
            m[4 * k + 3] = out.v[3];
        }
        uint32_t d[8];
        leaf64_words(m, d);
        store_digest(J.tree + uint64_t(t) * 8, d);
#pragma unroll
        for (int k = 0; k < 8; ++k) bufA[k][t] = d[k];
    }
    __syncthreads();
    // Parent levels: quads of lanes, SMALL_WG / 4 nodes per round.
    uint32_t* lvl = J.tree;
    int count = leaves;
    const uint32_t q = uint32_t(t & 3);
    const int g0 = t >> 2;
    for (int l = 1; l <= J.target; ++l) {
        uint32_t* next_lvl = lvl + uint64_t(count) * 8;
        const int pc = count >> 1;
        // Even levels read bufA and write bufB, odd levels the reverse.
        uint32_t(*cur)[SMALL_MAX_VALUES / 4] = bufA;
        for (int g = g0; g < pc; g += SMALL_WG / 4) {  // quad-uniform
            uint32_t Lw[8], Rw[8];
            if (l & 1) {
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    Lw[k] = cur[k][2 * g];
                    Rw[k] = cur[k][2 * g + 1];
                }
            } else {
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    Lw[k] = bufB[k][2 * g];
                    Rw[k] = bufB[k][2 * g + 1];
                }
            }
            uint32_t oa, ob;
            quad_node(Lw, Rw, q, oa, ob);
            if (l & 1) {
                bufB[q][g] = oa;
                bufB[4 + q][g] = ob;
            } else {
                bufA[q][g] = oa;
                bufA[4 + q][g] = ob;
            }
            const gptr dst = gm(next_lvl) + uint64_t(g) * 8;
            dst[q] = oa;
            dst[4 + q] = ob;
        }
        __syncthreads();
        count = pc;
        lvl = next_lvl;
    }
}

// ---------------------------------------------------------------------------
// Proof-of-work scan over head || u64 nonce (transcript.rs:182-351). A job's
// workgroups walk its chunks of [start, limit) in increasing order and stop
// once a passing nonce below their next chunk is known, so the reported
// minimum is the exact smallest passing nonce in the range: every chunk at
// or below it is scanned.
constexpr int GRIND_PER_THREAD = 4;

// One nonce of a job whose nonce starts in tail word W (byte offset 4W + s/8;
// the word is a template argument so the message stays in registers): its
// digest is compared with the difficulty and a pass lowers the job's minimum.
template <int W>
__device__ __forceinline__ void grind_try(const GrindJob& J, const uint32_t (&tw)[32], const uint32_t (&cv)[8],
                                          uint32_t s, bool two, uint64_t nonce) {
    const uint32_t lo = uint32_t(nonce), hi = uint32_t(nonce >> 32);
    // The nonce's 8 LE bytes spread over words W..W+2.
    const uint32_t c0 = lo << s;
    const uint32_t c1 = s ? (lo >> (32u - s)) | (hi << s) : hi;
    const uint32_t c2 = s ? hi >> (32u - s) : 0u;
    uint32_t m[32];
#pragma unroll
    for (int w = 0; w < 32; ++w) m[w] = tw[w];
    m[W] |= c0;
    m[W + 1] |= c1;
    if (W + 2 < 32) m[W + 2] |= c2;
    uint32_t d[8];
    compress8(cv, m, 0, 0, J.len0, J.flags0, d);
    if (W >= 14 && two) compress8(d, m + 16, 0, 0, J.len1, b3::CHUNK_END | b3::ROOT, d);
    const uint32_t be = __builtin_bswap32(d[0]);
    const uint32_t lz = be == 0 ? 32u : uint32_t(__builtin_clz(be));
    if (lz >= J.bits) atomicMin(J.result, (unsigned long long)nonce);
}

template <int W>
__device__ __forceinline__ void grind_load_head(const GrindJob& J, uint32_t (&tw)[32], uint32_t (&cv)[8]) {
#pragma unroll
    for (int w = 0; w < 32; ++w) tw[w] = gc(J.head->tail)[w];
#pragma unroll
    for (int w = 0; w < 8; ++w) cv[w] = gc(J.head->cv)[w];
}

// Nonce scan of one job by its own workgroups (strided chunks).
template <int W>
__device__ __forceinline__ void grind_scan(const GrindJob& J, uint32_t local) {
    uint32_t tw[32], cv[8];
    grind_load_head<W>(J, tw, cv);
    const uint32_t s = 8u * (J.nonce_off & 3u);
    const bool two = J.len1 != 0;
    const uint64_t chunk = uint64_t(WG) * GRIND_PER_THREAD;
    const uint64_t stride = uint64_t(J.nblocks) * chunk;
    for (uint64_t start = J.start + uint64_t(local) * chunk; start < J.limit; start += stride) {
        if (__hip_atomic_load(J.result, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < start) return;
        for (int k = 0; k < GRIND_PER_THREAD; ++k) {
            const uint64_t nonce = start + uint64_t(k) * WG + threadIdx.x;
            if (nonce >= J.limit) break;
            grind_try<W>(J, tw, cv, s, two, nonce);
        }
    }
}

// Pooled search over the jobs of one launch (batches): every workgroup takes
// GRIND_POOL_CHUNK-nonce chunks of its current job, in increasing order, from
// the job's counter (the u64 after its running minimum) and scans them. A job
// with no chunk left below its minimum or its limit is finished, and the
// workgroup moves to another job, so no job's tail holds the device while
// workgroups idle. Every chunk that starts below a job's final minimum is
// handed out and scanned, so the minimum is exactly the smallest passing
// nonce. A chunk is 16 nonces per thread: one counter update per 4096 nonces
// keeps the device-scope atomics on a job's counter rare.
constexpr int GRIND_POOL_PER_THREAD = 16;
constexpr uint64_t GRIND_POOL_CHUNK = uint64_t(WG) * GRIND_POOL_PER_THREAD;

__device__ __forceinline__ unsigned long long* grind_counter(const GrindJob& J) { return J.result + 1; }

template <int W>
__device__ __forceinline__ void grind_pool_job(const GrindJob& J, unsigned long long* slot) {
    uint32_t tw[32], cv[8];
    grind_load_head<W>(J, tw, cv);
    const uint32_t s = 8u * (J.nonce_off & 3u);
    const bool two = J.len1 != 0;
    for (;;) {
        __syncthreads();  // everyone has read the previous chunk's start
        if (threadIdx.x == 0) {
            // The first thread takes the chunk and decides for the workgroup.
            const unsigned long long c =
                __hip_atomic_fetch_add(grind_counter(J), 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const uint64_t st = J.start + uint64_t(c) * GRIND_POOL_CHUNK;
            const bool done =
                st >= J.limit || st >= __hip_atomic_load(J.result, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            *slot = done ? ~0ull : st;
        }
        __syncthreads();
        const uint64_t start = *slot;
        if (start == ~0ull) return;  // workgroup-uniform
        for (int k = 0; k < GRIND_POOL_PER_THREAD; ++k) {
            const uint64_t nonce = start + uint64_t(k) * WG + threadIdx.x;
            if (nonce < J.limit) grind_try<W>(J, tw, cv, s, two, nonce);
        }
    }
}

// True while the job still has a chunk that starts below its minimum.
__device__ __forceinline__ bool grind_job_live(const GrindJob& J) {
    const uint64_t next = __hip_atomic_load(grind_counter(J), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t start = J.start + next * GRIND_POOL_CHUNK;
    return start < J.limit && start < __hip_atomic_load(J.result, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Running minimum to "none", pooled chunk counter to 0.
__global__ void k_grind_init(GrindJob inl, const GrindJob* tab, uint32_t njobs) {
    uint32_t i = blockIdx.x * WG + threadIdx.x;
    unsigned long long* r = nullptr;
    if (njobs == 0) {
        if (i == 0) r = inl.result;
    } else if (i < njobs) {
        r = tab[i].result;
    }
    if (r) {
        r[0] = ~0ull;
        r[1] = 0;
    }
}

// Workgroup w starts on job (w mod njobs); a workgroup whose job is finished
// has its first wave look at the next jobs in order (64 at a time) and takes
// the first live one, and it exits when none is left.
__global__ __launch_bounds__(WG) void k_grind_pool(const GrindJob* tab, uint32_t njobs) {
    __shared__ unsigned long long slot;
    __shared__ uint32_t next_job;
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t j = blockIdx.x % njobs;
    for (;;) {
        const GrindJob J = tab[j];
        switch (J.nonce_off >> 2) {
            case 0: grind_pool_job<0>(J, &slot); break;
            case 1: grind_pool_job<1>(J, &slot); break;
            case 2: grind_pool_job<2>(J, &slot); break;
            case 3: grind_pool_job<3>(J, &slot); break;
            case 4: grind_pool_job<4>(J, &slot); break;
            case 5: grind_pool_job<5>(J, &slot); break;
            case 6: grind_pool_job<6>(J, &slot); break;
            case 7: grind_pool_job<7>(J, &slot); break;
            case 8: grind_pool_job<8>(J, &slot); break;
            case 9: grind_pool_job<9>(J, &slot); break;
            case 10: grind_pool_job<10>(J, &slot); break;
            case 11: grind_pool_job<11>(J, &slot); break;
            case 12: grind_pool_job<12>(J, &slot); break;
            case 13: grind_pool_job<13>(J, &slot); break;
            case 14: grind_pool_job<14>(J, &slot); break;
            case 15: grind_pool_job<15>(J, &slot); break;
            default: return;  // unreachable: the nonce starts in the first 64 tail bytes
        }
        // Offsets 1..njobs from j (j itself last); njobs means none is live.
        if (threadIdx.x < 64) {
            uint32_t pick = njobs;
            for (uint32_t base = 1; base <= njobs; base += 64) {
                const uint32_t k = base + lane;
                const bool live = k <= njobs && grind_job_live(tab[(j + k) % njobs]);
                const unsigned long long mask = __ballot(live);
                if (mask) {
                    pick = (j + base + uint32_t(__builtin_ctzll(mask))) % njobs;
                    break;
                }
            }
            if (threadIdx.x == 0) next_job = pick;
        }
        __syncthreads();
        j = next_job;
        __syncthreads();  // next_job is read by everyone before it is written again
        if (j == njobs) return;
    }
}

__global__ __launch_bounds__(WG) void k_grind(GrindJob inl, const GrindJob* tab, uint32_t njobs) {
    uint32_t local;
    const GrindJob J = job_of(inl, tab, njobs, local);
    switch (J.nonce_off >> 2) {
        case 0: grind_scan<0>(J, local); break;
        case 1: grind_scan<1>(J, local); break;
        case 2: grind_scan<2>(J, local); break;
        case 3: grind_scan<3>(J, local); break;
        case 4: grind_scan<4>(J, local); break;
        case 5: grind_scan<5>(J, local); break;
        case 6: grind_scan<6>(J, local); break;
        case 7: grind_scan<7>(J, local); break;
        case 8: grind_scan<8>(J, local); break;
        case 9: grind_scan<9>(J, local); break;
        case 10: grind_scan<10>(J, local); break;
        case 11: grind_scan<11>(J, local); break;
        case 12: grind_scan<12>(J, local); break;
        case 13: grind_scan<13>(J, local); break;
        case 14: grind_scan<14>(J, local); break;
        case 15: grind_scan<15>(J, local); break;
        default: break;
    }
}

// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
constexpr int COLLECT_WG_MAX = 1024;
constexpr uint32_t COLLECT_LDS_ITEMS = 8192;

__global__ __launch_bounds__(COLLECT_WG_MAX) void k_collect(CollectItem inl, const CollectItem* tab, uint32_t nitems,
                                                            uint32_t total, uint64_t* flag, uint64_t seq,
                                                            HostGate gate) {
    if (nitems == 1) {
        for (uint32_t k = threadIdx.x; k < inl.words; k += blockDim.x) inl.dst[k] = gc(inl.src)[k];
    } else if (nitems <= COLLECT_LDS_ITEMS) {
        // All items' words as one range: every thread has independent copies
        // in flight, whatever the items' sizes (a batch has hundreds). The
        // items' first words sit in LDS for the search.
        __shared__ uint32_t first[COLLECT_LDS_ITEMS];
        for (uint32_t i = threadIdx.x; i < nitems; i += blockDim.x) first[i] = tab[i].first;
        __syncthreads();
        for (uint32_t g = threadIdx.x; g < total; g += blockDim.x) {
            uint32_t lo = 0, hi = nitems;
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (first[mid] <= g) lo = mid;
                else hi = mid;
            }
            const CollectItem& J = tab[lo];
            J.dst[g - first[lo]] = gc(J.src)[g - first[lo]];
        }
    } else {
        for (uint32_t i = 0; i < nitems; ++i) {
            const CollectItem& J = tab[i];
            for (uint32_t k = threadIdx.x; k < J.words; k += blockDim.x) J.dst[k] = gc(J.src)[k];
        }
    }
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0 && flag) __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    if (gate.go == nullptr) return;
    // Gate: lane 0 polls the host counter (system scope, bounded), then the
    // workgroup copies the next phase's constants from host memory.
    __shared__ uint32_t open;
    if (threadIdx.x == 0) {
        uint32_t i = 0;
        for (; i < gate.spin_limit; ++i) {
            if (__hip_atomic_load(gate.go, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM) >= gate.want) break;
            __builtin_amdgcn_s_sleep(2);
        }
        open = i < gate.spin_limit ? 1u : 0u;
        if (!open) __hip_atomic_store(gate.timeout, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    __syncthreads();
    // A gate that gave up hands zeros to the next phase instead of whatever
    // the mailbox holds: the openings then read live = 0 and skip, and the
    // other phases compute in-bounds garbage that the host discards.
    for (uint32_t k = threadIdx.x; k < gate.words; k += blockDim.x)
        gate.dst[k] = open ? __hip_atomic_load(gate.src + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) : 0u;
}

__global__ __launch_bounds__(WG) void k_zero(ZeroJob inl, const ZeroJob* tab, uint32_t njobs) {
    uint32_t local;
    const ZeroJob J = job_of(inl, tab, njobs, local);
    const uint64_t i = uint64_t(local) * WG + threadIdx.x;
    if (i < J.words) J.dst[i] = 0;
}

// Plain streaming copy: the calibration reference for FETCH/WRITE_SIZE.
__global__ void k_copy(const uint4* __restrict__ in, uint4* __restrict__ out, uint64_t n16) {
    uint64_t i = uint64_t(blockIdx.x) * WG + threadIdx.x;
    const uint64_t stride = uint64_t(gridDim.x) * WG;
    for (; i < n16; i += stride) out[i] = in[i];
}

__global__ __launch_bounds__(WG) void k_pow_brev(PowTable pt, uint32_t* out, int bits, int shift) {
    const uint64_t i = uint64_t(blockIdx.x) * WG + threadIdx.x;
    if (i >> bits) return;
    gm(out)[i] = pow_lookup(pt, brev_bits(uint32_t(i), bits) << shift);
}

inline unsigned grid_for(uint64_t work, uint64_t per_block) { return unsigned((work + per_block - 1) / per_block); }

// Assign block_begin, return the total grid.
template <typename J, typename F>
unsigned assign_blocks(std::vector<J>& jobs, F blocks_of) {
    uint64_t total = 0;
    for (J& j : jobs) {
        j.block_begin = uint32_t(total);
        total += blocks_of(j);
    }
    return unsigned(total);
}

template <typename J>
const J* table_of(const Launch& L, const std::vector<J>& jobs) {
    if (jobs.size() <= 1) return nullptr;
    return static_cast<const J*>(L.sink->push(jobs.data(), jobs.size() * sizeof(J)));
}

template <typename J, typename K>
void run(const Launch& L, const char* name, K kernel, std::vector<J>& jobs, unsigned grid, size_t lds) {
    if (jobs.empty() || grid == 0) return;
    const J* table = table_of(L, jobs);
    if (L.dry) return;
    const uint32_t n = jobs.size() <= 1 ? 0u : uint32_t(jobs.size());
    ProfScope ps(L.prof, name, L.stream);
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(WG), lds, L.stream, jobs[0], table, n);
}

int ntt_cols_for(int log1, int log2) {
    if (log1 >= 4 && log1 <= 12) return N4K >> log1;  // register path: 4096-element tiles
    int cols = COL_LDS_WORDS >> log1;
    return std::min(cols, 1 << log2);
}

uint64_t ntt_blocks(const NttJob& j) {
    const uint64_t B = uint64_t(1) << j.logb;
    switch (j.kind) {
        case NTT_INTT_SMALL: return 1;
        case NTT_INTT_COLS: return (uint64_t(1) << j.log2) / j.cols;
        case NTT_INTT_ROWS: return uint64_t(1) << j.log1;
        case NTT_LDE_SMALL: return B;
        case NTT_LDE_ROWS: return (uint64_t(1) << j.log1) * B;
        case NTT_LDE_COLS: return ((uint64_t(1) << j.log2) / j.cols) * B;
        default: return 0;
    }
}

size_t ntt_lds(const NttJob& j) {
    switch (j.kind) {
        case NTT_INTT_SMALL:
        case NTT_LDE_SMALL: return j.logn == 12 ? size_t(N4K_LDS) * 4 : (size_t(1) << j.logn) * 6;
        case NTT_INTT_ROWS:
        case NTT_LDE_ROWS: return j.log2 == 12 ? size_t(N4K_LDS + TWL_WORDS) * 4 : (size_t(1) << j.log2) * 6;
        case NTT_INTT_COLS:
        case NTT_LDE_COLS:
            if (j.log1 >= 4 && j.log1 <= 12) return ((size_t(1) << j.log1) + N4K) * 4;  // pitch C + 1
            return size_t(j.cols) * ((size_t(1) << j.log1) + 1) * 4 + (size_t(1) << j.log1) * 2;
        default: return 0;
    }
}

}  // namespace

NttShape ntt_shape(int logn) {
    // Rows (contiguous, DIT/DIF in LDS) take up to 4096 points; the strided
    // column pass takes the remainder, COL_LDS_WORDS / n1 columns at a time.
    NttShape s;
    s.logn = logn;
    if (logn <= 12) {
        s.log1 = 0;
        s.log2 = logn;
    } else if (logn <= 25) {
        s.log2 = 12;
        s.log1 = logn - 12;
    } else {
        // 2^26 (FRI at blowup 2): 8192-point rows keep the strided length
        // at 2^13, one column per workgroup in LDS.
        s.log2 = 13;
        s.log1 = logn - 13;
    }
    return s;
}

NttPlan plan_intt(const uint32_t* evals, uint32_t* coeffs_br, uint32_t* scratch, int logn, const uint32_t* tw4096_inv,
                  PowTable winv_m, int logb, uint32_t ninv_m) {
    NttShape s = ntt_shape(logn);
    NttJob j{};
    j.tw4096 = tw4096_inv;
    j.pt = winv_m;
    j.ninv_m = ninv_m;
    j.logn = logn;
    j.log1 = s.log1;
    j.log2 = s.log2;
    j.logb = logb;
    NttPlan plan;
    if (s.log1 == 0) {
        j.kind = NTT_INTT_SMALL;
        j.in = evals;
        j.out = coeffs_br;
        plan.pass1.push_back(j);
        return plan;
    }
    j.cols = ntt_cols_for(s.log1, s.log2);
    j.kind = NTT_INTT_COLS;
    j.in = evals;
    j.out = scratch;
    plan.pass1.push_back(j);
    j.kind = NTT_INTT_ROWS;
    j.in = scratch;
    j.out = coeffs_br;
    plan.pass2.push_back(j);
    return plan;
}

NttPlan plan_lde(const uint32_t* coeffs_br, uint32_t* lde_br, uint32_t* scratch, int logn, int logb,
                 const uint32_t* tw4096, PowTable w_m, CosetTables ct) {
    NttShape s = ntt_shape(logn);
    NttJob j{};
    j.tw4096 = tw4096;
    j.pt = w_m;
    j.logn = logn;
    j.log1 = s.log1;
    j.log2 = s.log2;
    j.logb = logb;
    NttPlan plan;
    if (s.log1 == 0) {
        j.kind = NTT_LDE_SMALL;
        j.in = coeffs_br;
        j.out = lde_br;
        j.tab_a = ct.small;
        plan.pass1.push_back(j);
        return plan;
    }
    j.cols = ntt_cols_for(s.log1, s.log2);
    j.kind = NTT_LDE_ROWS;
    j.in = coeffs_br;
    j.out = scratch;
    j.tab_a = ct.row;
    j.tab_b = ct.col;
    plan.pass1.push_back(j);
    j.kind = NTT_LDE_COLS;
    j.in = scratch;
    j.out = lde_br;
    j.tab_a = j.tab_b = nullptr;
    plan.pass2.push_back(j);
    return plan;
}

void launch_fib_coeffs(const Launch& L, std::vector<FibCoeffJob> jobs) {
    unsigned g = assign_blocks(jobs, [](const FibCoeffJob& j) { return grid_for(uint64_t(1) << j.logn, WG * FIB_ELT); });
    run(L, "fib_coeffs", k_fib_coeffs, jobs, g, 0);
}

bool lde_rows_job(const NttJob& j) { return j.kind == NTT_LDE_ROWS && j.log2 == 12; }
bool lde_cols_job(const NttJob& j) { return j.kind == NTT_LDE_COLS && j.log2 == 12 && j.log1 >= 4 && j.log1 <= 12; }

void launch_ntt(const Launch& L, std::vector<NttJob> jobs) {
    std::vector<NttJob> rows, cols, rest;
    for (const NttJob& j : jobs)
        (!L.lde_split ? rest : lde_rows_job(j) ? rows : lde_cols_job(j) ? cols : rest).push_back(j);
    size_t lds = 0;
    bool intt = false;
    for (const NttJob& j : rest) {
        lds = std::max(lds, ntt_lds(j));
        intt = intt || j.kind <= NTT_INTT_ROWS;
    }
    unsigned g = assign_blocks(rest, [](const NttJob& j) { return ntt_blocks(j); });
    run(L, intt ? "intt" : "lde", k_ntt, rest, g, lds);
    g = assign_blocks(rows, [](const NttJob& j) { return ntt_blocks(j); });
    run(L, "lde_rows", k_lde_rows, rows, g, size_t(N4K_LDS + TWL_WORDS) * 4);
    lds = 0;
    for (const NttJob& j : cols) lds = std::max(lds, ntt_lds(j));
    g = assign_blocks(cols, [](const NttJob& j) { return ntt_blocks(j); });
    run(L, "lde_cols", k_lde_cols, cols, g, lds);
}

void launch_linked(const Launch& L, std::vector<LinkedJob> jobs) {
    unsigned g = assign_blocks(jobs, [](const LinkedJob& j) {
        return grid_for(uint64_t(1) << (j.logn + j.logb), WG * ELT);
    });
    run(L, "linked_b", k_linked, jobs, g, 0);
}

void launch_point_inv(const Launch& L, const PointInvJob& job) {
    if (L.dry) return;
    const uint64_t m = uint64_t(1) << (job.logn + job.logb);
    hipLaunchKernelGGL(k_point_inv, dim3(grid_for(m, WG * CELT)), dim3(WG), 0, L.stream, job);
}

void launch_compose(const Launch& L, std::vector<ComposeJob> all) {
    std::vector<ComposeJob> plain, fused;
    for (const ComposeJob& j : all) (j.tree ? fused : plain).push_back(j);
    if (!plain.empty()) {
        unsigned g = assign_blocks(plain, [](const ComposeJob& j) {
            return grid_for(uint64_t(1) << (j.logn + j.logb), WG * CELT);
        });
        run(L, "compose", k_compose, plain, g, 0);
    }
    if (!fused.empty()) {
        unsigned g = assign_blocks(fused, [](const ComposeJob& j) {
            return grid_for(uint64_t(1) << (j.logn + j.logb), FUSE_VALUES);
        });
        run(L, "compose_leaves", k_compose_leaves, fused, g, 0);
    }
}

void launch_range_compose(const Launch& L, std::vector<RangeComposeJob> jobs) {
    unsigned g = assign_blocks(jobs, [](const RangeComposeJob& j) {
        return grid_for(uint64_t(1) << (j.logn + j.logb), WG * CELT);
    });
    run(L, "range_compose", k_range_compose, jobs, g, 0);
}

void launch_table_compose(const Launch& L, std::vector<TableComposeJob> jobs) {
    unsigned g = assign_blocks(jobs, [](const TableComposeJob& j) {
        return grid_for(uint64_t(1) << (j.logn + j.logb), WG * CELT);
    });
    run(L, "table_compose", k_table_compose, jobs, g, 0);
}

void launch_fold_combine(const Launch& L, std::vector<FoldCombineJob> jobs) {
    unsigned g = assign_blocks(jobs, [](const FoldCombineJob& j) {
        return grid_for(uint64_t(1) << (j.logn + j.logb), WG * CELT);
    });
    run(L, "fold_combine", k_fold_combine, jobs, g, 0);
}

void launch_leaves(const Launch& L, std::vector<LeavesJob> all) {
    std::vector<LeavesJob> linked, rows, plain;
    bool trace = false, nodes_only = true;
    for (const LeavesJob& j : all) {
        if (j.kind == LEAF_TRACE_LINKED) linked.push_back(j);
        else if (j.kind == LEAF_ROWS) rows.push_back(j);
        else plain.push_back(j);
        trace = trace || j.kind == LEAF_TRACE;
        nodes_only = nodes_only && j.kind == LEAF_NODES;
    }
    auto blocks = [](const LeavesJob& j) { return j.leaves == 1 ? 1u : grid_for(j.leaves / 2, WG); };
    if (!linked.empty()) run(L, "trace_leaves", k_leaves_linked, linked, assign_blocks(linked, blocks), 0);
    if (!rows.empty()) run(L, "row_leaves", k_leaves_rows, rows, assign_blocks(rows, blocks), 0);
    if (!plain.empty())
        run(L, nodes_only ? "tree_nodes" : trace ? "trace_leaves" : "fri_leaves", k_leaves, plain,
            assign_blocks(plain, blocks), 0);
}

void launch_levels(const Launch& L, std::vector<LevelsJob> jobs) {
    unsigned g = assign_blocks(jobs, [](const LevelsJob& j) { return grid_for(j.count / (2 * j.fuse), WG); });
    run(L, "tree_levels", k_levels, jobs, g, 0);
}

// One launch per group size: the workgroup is two threads per digest.
void launch_zero(const Launch& L, std::vector<ZeroJob> jobs) {
    unsigned g = assign_blocks(jobs, [](const ZeroJob& j) { return grid_for(j.words, WG); });
    run(L, "zero", k_zero, jobs, g, 0);
}

void launch_tree_top(const Launch& L, std::vector<TreeTopJob> jobs) {
    unsigned g = assign_blocks(jobs, [](const TreeTopJob& j) { return grid_for(j.leaves >> j.lvl, 64); });
    run(L, "tree_top", k_tree_top, jobs, g, 0);
}

void launch_tail(const Launch& L, std::vector<TailJob> all) {
    for (int gl = TAIL_GROUP_MIN_LOG; gl <= TAIL_GROUP_MAX_LOG; ++gl) {
        std::vector<TailJob> jobs;
        for (const TailJob& j : all)
            if (j.group_log == gl) jobs.push_back(j);
        unsigned g = assign_blocks(jobs, [gl](const TailJob& j) { return grid_for(j.count, uint64_t(1) << gl); });
        if (jobs.empty() || g == 0) continue;
        const TailJob* table = table_of(L, jobs);
        if (L.dry) continue;
        const uint32_t n = jobs.size() <= 1 ? 0u : uint32_t(jobs.size());
        ProfScope ps(L.prof, "tree_tail", L.stream);
        hipLaunchKernelGGL(k_tail, dim3(g), dim3(2u << gl), 0, L.stream, jobs[0], table, n);
    }
}

void launch_small_commit(const Launch& L, std::vector<SmallCommitJob> jobs) {
    unsigned g = assign_blocks(jobs, [](const SmallCommitJob&) { return uint64_t(1); });
    if (jobs.empty()) return;
    const SmallCommitJob* table = table_of(L, jobs);
    if (L.dry) return;
    const uint32_t n = jobs.size() <= 1 ? 0u : uint32_t(jobs.size());
    ProfScope ps(L.prof, "small_commit", L.stream);
    hipLaunchKernelGGL(k_small_commit, dim3(g), dim3(SMALL_WG), 0, L.stream, jobs[0], table, n);
}

void launch_fold(const Launch& L, std::vector<FoldJob> all) {
    std::vector<FoldJob> plain, fused, commit;
    for (const FoldJob& j : all) (j.counters ? commit : j.tree ? fused : plain).push_back(j);
    if (!commit.empty()) {
        unsigned g = assign_blocks(commit, [](const FoldJob& j) { return grid_for(uint64_t(1) << (j.log_len - 2), WG); });
        run(L, "fold_commit", k_fold_commit, commit, g, 0);
    }
    if (!plain.empty()) {
        unsigned g = assign_blocks(plain, [](const FoldJob& j) {
            return grid_for(uint64_t(1) << (j.log_len - 2), WG * FOLD_PER);
        });
        run(L, "fold", k_fold, plain, g, 0);
    }
    if (!fused.empty()) {
        unsigned g = assign_blocks(fused, [](const FoldJob& j) {
            return grid_for(uint64_t(1) << (j.log_len - 2), FUSE_VALUES);
        });
        run(L, "fold_leaves", k_fold_leaves, fused, g, 0);
    }
}

void launch_grind(const Launch& L, std::vector<GrindJob> jobs) {
    if (jobs.empty()) return;
    unsigned g = assign_blocks(jobs, [](const GrindJob& j) { return uint64_t(j.nblocks); });
    const GrindJob* table = table_of(L, jobs);
    if (L.dry) return;
    const uint32_t n = jobs.size() <= 1 ? 0u : uint32_t(jobs.size());
    {
        ProfScope ps(L.prof, "grind_init", L.stream);
        hipLaunchKernelGGL(k_grind_init, dim3(grid_for(std::max<size_t>(jobs.size(), 1), WG)), dim3(WG), 0, L.stream,
                           jobs[0], table, n);
    }
    ProfScope ps(L.prof, "grind", L.stream);
    if (n >= 2 && L.grind_pool > 0) {
        // Several proofs grinding: one pool of workgroups over all of them.
        hipLaunchKernelGGL(k_grind_pool, dim3(L.grind_pool), dim3(WG), 0, L.stream, table, n);
    } else {
        hipLaunchKernelGGL(k_grind, dim3(g), dim3(WG), 0, L.stream, jobs[0], table, n);
    }
}

// ---------------------------------------------------------------------------
// The allocation-free grind's device memory: one grind head and the running
// minimum (GRIND_BUF_BYTES), a static array of which every device holds its
// own copy until the process ends. Host code never passes the array's name
// to a kernel or a copy: it uses the address hipGetSymbolAddress returns on
// each device, looked up once per device when a context is created there.
static __device__ __align__(256) uint8_t scratch[GRIND_BUF_BYTES];

namespace {
std::mutex noalloc_mu;
uint8_t* noalloc_ptr[NOALLOC_MAX_DEVICES] = {};
}  // namespace

void noalloc_scratch_init(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) throw Status(Z0P_ERR_NO_DEVICE);
    std::lock_guard<std::mutex> lk(noalloc_mu);
    if (noalloc_ptr[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    require(e == hipSuccess && p != nullptr, Z0P_ERR_DEVICE);
    noalloc_ptr[device] = static_cast<uint8_t*>(p);
}

uint8_t* noalloc_scratch(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(noalloc_mu);
    return noalloc_ptr[device];
}

bool grind_prepare(const uint8_t* head, size_t head_len, GrindHead& h, GrindJob& j) {
    const size_t total = head_len + 8;
    if (total > GRIND_MAX_PREIMAGE) return false;
    const size_t first = head_len / 64;            // block the nonce starts in
    const size_t blocks = (total + 63) / 64;       // blocks of head || nonce
    b3_iv(h.cv);
    for (size_t b = 0; b < first; ++b) {
        uint32_t m[16];
        for (int k = 0; k < 16; ++k) m[k] = load_le32(head + 64 * b + 4 * k);
        compress8(h.cv, m, 0, 0, 64, b == 0 ? b3::CHUNK_START : 0u, h.cv);
    }
    uint8_t tail[128] = {0};
    std::memcpy(tail, head + 64 * first, head_len - 64 * first);
    for (int k = 0; k < 32; ++k) h.tail[k] = load_le32(tail + 4 * k);
    const bool two = blocks - first == 2;
    j.nonce_off = uint32_t(head_len - 64 * first);
    j.len0 = two ? 64u : uint32_t(total - 64 * first);
    j.flags0 = (first == 0 ? b3::CHUNK_START : 0u) | (two ? 0u : b3::CHUNK_END | b3::ROOT);
    j.len1 = two ? uint32_t(total - 64 * (blocks - 1)) : 0u;
    return true;
}

void launch_bitrev(const Launch& L, std::vector<BitrevJob> jobs) {
    unsigned g = assign_blocks(jobs, [](const BitrevJob& j) { return grid_for(uint64_t(1) << j.logn, WG); });
    run(L, "bitrev", k_bitrev, jobs, g, 0);
}

void launch_interleave(const Launch& L, std::vector<InterleaveJob> jobs) {
    unsigned g = assign_blocks(jobs, [](const InterleaveJob& j) { return grid_for(j.count, WG); });
    run(L, "interleave", k_interleave, jobs, g, 0);
}

void launch_collect(const Launch& L, std::vector<CollectItem> items, uint64_t* flag, uint64_t seq,
                    const HostGate& gate) {
    uint64_t words = 0;
    for (CollectItem& it : items) {
        it.first = uint32_t(words);
        words += it.words;
    }
    const CollectItem* table = table_of(L, items);
    if (L.dry) return;
    // One wave for the few words of one proof; a full workgroup for a batch.
    const unsigned threads = words <= 4096 ? 64u : unsigned(COLLECT_WG_MAX);
    const CollectItem inl = items.empty() ? CollectItem{nullptr, nullptr, 0, 0} : items[0];
    ProfScope ps(L.prof, "collect", L.stream);
    hipLaunchKernelGGL(k_collect, dim3(1), dim3(threads), 0, L.stream, inl, table, uint32_t(items.size()),
                       uint32_t(words), flag, seq, gate);
}

void launch_pow_brev(const Launch& L, PowTable pt, uint32_t* out, int bits, int shift) {
    if (L.dry) return;
    hipLaunchKernelGGL(k_pow_brev, dim3(grid_for(uint64_t(1) << bits, WG)), dim3(WG), 0, L.stream, pt, out, bits,
                       shift);
}

void launch_verify_paths(hipStream_t s, const uint32_t* checks, const uint32_t* siblings, uint32_t count,
                         uint8_t* ok) {
    if (count == 0) return;
    hipLaunchKernelGGL(k_verify_paths, dim3((count + WG - 1) / WG), dim3(WG), 0, s, checks, siblings, count, ok);
}

void launch_copy(const Launch& L, const uint32_t* in, uint32_t* out, uint64_t words) {
    if (L.dry) return;
    ProfScope ps(L.prof, "copy", L.stream);
    hipLaunchKernelGGL(k_copy, dim3(4096), dim3(WG), 0, L.stream, reinterpret_cast<const uint4*>(in),
                       reinterpret_cast<uint4*>(out), words / 4);
}

}  // namespace z0p

namespace z0p {
namespace {

// Word offset of level l in a level-contiguous tree of L = 2^depth leaves
// (l <= depth): 8 * (L + L/2 + ... + L/2^(l-1)).
__device__ __forceinline__ uint64_t level_offset_words(uint64_t L, int l) { return 16 * (L - (L >> l)); }

// Source of output word w of one query's opening (layout in kernels.hpp).
__device__ const uint32_t* open_src(const OpenJob& J, uint64_t leaf, uint32_t w) {
    uint64_t index = leaf;
    for (int d = 0; d < J.nlayers; ++d) {
        const uint64_t base = index / 4 * 4;
        const uint32_t vals = J.base[d] ? 4u : 16u;
        if (w < vals) return J.base[d] ? J.layer[d] + base + w : J.layer[d] + base * 4 + w;
        w -= vals;
        const uint32_t pw = uint32_t(J.target[d]) * 8u;
        if (w < pw) {
            const int l = int(w >> 3);
            const uint64_t sib = ((base / 4) >> l) ^ 1;
            return J.tree[d] + level_offset_words(J.tree_leaves[d], l) + sib * 8 + (w & 7);
        }
        w -= pw;
        index /= 4;
    }
    const int logm = J.logn + J.logb;
    if (J.cols) {
        // Range rows: the columns at the query's position (natural row
        // brev(leaf)), then the row's path in the one-leaf-per-row tree.
        if (w < uint32_t(J.ncols)) return J.cols + uint64_t(w) * J.col_stride + leaf;
        w -= uint32_t(J.ncols);
        const uint64_t row = brev_bits(uint32_t(leaf), logm);
        const int l = int(w >> 3);
        return J.trace_tree + level_offset_words(J.trace_leaves, l) + (((row >> l) ^ 1) * 8) + (w & 7);
    }
    // AIR trace rows: leaves of natural rows idx and idx + B, then their paths.
    const uint64_t m = uint64_t(1) << logm, B = uint64_t(1) << J.logb;
    const uint64_t idx = brev_bits(uint32_t(leaf), logm);
    const uint64_t shifted = (idx + B) & (m - 1);
    if (w < 32) {
        const uint64_t row0 = ((w >> 4) ? shifted : idx) & ~uint64_t(7);
        const uint32_t k = w & 15;
        const uint64_t i = row0 + (k >> 1);
        const uint32_t j = uint32_t(i & (B - 1)), t = uint32_t(i >> J.logb);
        const uint64_t pos = (uint64_t(brev_bits(j, J.logb)) << J.logn) | brev_bits(t, J.logn);
        return ((k & 1) ? J.b : J.a) + pos;
    }
    w -= 32;
    const uint32_t pw = uint32_t(J.trace_target) * 8u;  // > 0 whenever words remain
    const uint64_t leaf8 = (w >= pw ? shifted : idx) / 8;
    if (w >= pw) w -= pw;
    const int l = int(w >> 3);
    const uint64_t sib = (leaf8 >> l) ^ 1;
    return J.trace_tree + level_offset_words(J.trace_leaves, l) + sib * 8 + (w & 7);
}

// One workgroup per query, one thread per output word (all loads
// independent). The output is host memory: fence before finishing.
__global__ __launch_bounds__(WG) void k_open(const OpenJob* tab, uint32_t njobs) {
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (tab[mid].block_begin <= blockIdx.x) lo = mid;
        else hi = mid;
    }
    const OpenJob& J = tab[lo];
    const uint32_t q = blockIdx.x - J.block_begin;
    if (q < J.nq && (J.live == nullptr || *J.live != 0)) {
        const uint64_t leaf = J.queries[q];
        uint32_t* o = J.out + uint64_t(q) * J.words_per_query;
        for (uint32_t w = threadIdx.x; w < J.words_per_query; w += WG) o[w] = *open_src(J, leaf, w);
    }
    __threadfence_system();
}

}  // namespace

void launch_open(const Launch& L, std::vector<OpenJob> jobs) {
    if (jobs.empty()) return;
    unsigned g = assign_blocks(jobs, [](const OpenJob& j) { return uint64_t(j.nq); });
    if (g == 0) return;
    const OpenJob* table = static_cast<const OpenJob*>(L.sink->push(jobs.data(), jobs.size() * sizeof(OpenJob)));
    if (L.dry) return;
    ProfScope ps(L.prof, "open", L.stream);
    hipLaunchKernelGGL(k_open, dim3(g), dim3(WG), 0, L.stream, table, uint32_t(jobs.size()));
}

}  // namespace z0p

namespace z0p {
namespace {

// One workgroup per (step, query) pair, grid-strided; one thread per word.
__global__ __launch_bounds__(WG) void k_fold_open(FoldOpenJob J) {
    const int logm = J.logn + J.logb;
    const uint64_t m = uint64_t(1) << logm, B = uint64_t(1) << J.logb;
    const uint32_t pw = uint32_t(J.target) * 8u;
    const uint64_t pairs = uint64_t(J.nsteps) * J.nq;
    for (uint64_t g = blockIdx.x; g < pairs; g += gridDim.x) {
        const uint32_t k = uint32_t(g / J.nq), q = uint32_t(g % J.nq);
        const FoldStepDesc& S = J.steps[k];
        const uint64_t leaf = J.queries[q];
        const uint64_t idx = brev_bits(uint32_t(leaf), logm);
        const uint64_t sh = (idx + B) & (m - 1);
        uint32_t* o = J.out + g * J.words_per_query;
        for (uint32_t w = threadIdx.x; w < J.words_per_query; w += WG) {
            uint32_t v;
            if (w < 4) {
                const uint64_t pos = w < 2 ? leaf : brev_bits(uint32_t(sh), logm);
                v = gc((w & 1) ? S.b : S.a)[pos];
            } else {
                uint32_t r = w - 4;
                const uint64_t row = r >= pw ? sh : idx;
                if (r >= pw) r -= pw;
                const int l = int(r >> 3);
                v = gc(S.tree)[level_offset_words(m, l) + ((row >> l) ^ 1) * 8 + (r & 7)];
            }
            o[w] = v;
        }
    }
}

}  // namespace

void launch_fold_open(const Launch& L, const FoldOpenJob& job) {
    if (L.dry) return;
    const uint64_t pairs = uint64_t(job.nsteps) * job.nq;
    if (pairs == 0) return;
    ProfScope ps(L.prof, "fold_open", L.stream);
    hipLaunchKernelGGL(k_fold_open, dim3(unsigned(std::min<uint64_t>(pairs, 65536))), dim3(WG), 0, L.stream, job);
}

}  // namespace z0p
