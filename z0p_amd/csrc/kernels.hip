// HIP kernels for gfx950 (MI355X): the proving hot path of the Fibonacci
// AIR / FRI commitment. All arithmetic is exact 32-bit integer work; no
// MFMA. Layout conventions:
//  - LDE columns and FRI layers live in bit-reversed order over the whole
//    domain of size m = B*n: natural index i = j + B*t (coset j, step t)
//    sits at brev_{log B}(j)*n + brev_{log n}(t). Coset j is therefore one
//    contiguous block, FRI quads are adjacent, and the next trace row
//    (i + B) stays inside its block.
//  - Digests are 8 little-endian u32 words; tree levels are stored
//    contiguously leaf level first.
#include <hip/hip_runtime.h>

#include "blake3.hpp"
#include "field.hpp"
#include "kernels.hpp"
#include "profiler.hpp"

namespace z0p {

namespace {

constexpr int WG = 256;

__device__ __forceinline__ uint32_t brev_bits(uint32_t x, int bits) {
    return bits == 0 ? 0u : (__builtin_bitreverse32(x) >> (32 - bits));
}

__device__ __forceinline__ uint32_t pow_lookup(PowTable t, uint32_t e) {
    return mont_mul(t.lo[e & 4095u], t.hi[e >> 12]);
}

// Montgomery-domain inverse: (a^-1)_m from a_m, a != 0 (a = 0 maps to 0).
// Exponent p-2 = 0b111_0_(27 ones): window-3 chain, 30 squarings + 11 mults.
__device__ __forceinline__ uint32_t inv_m(uint32_t a) {
    uint32_t a2 = mont_mul(a, a);
    uint32_t a3 = mont_mul(a2, a);
    uint32_t a6 = mont_mul(a3, a3);
    uint32_t a7 = mont_mul(a6, a);  // t3 = a^(2^3-1)
    uint32_t acc = mont_mul(a7, a7);  // bits 30..27 = 1110
#pragma unroll
    for (int w = 0; w < 9; ++w) {
        acc = mont_mul(acc, acc);
        acc = mont_mul(acc, acc);
        acc = mont_mul(acc, acc);
        acc = mont_mul(acc, a7);
    }
    return acc;
}

// ---------------------------------------------------------------------------
// Fibonacci trace column A: A[i] = F(i-1) a0 + F(i) b0 (F(-1) = 1, F(0) = 0).
// Each thread jumps to its chunk start with fast doubling, then iterates.
constexpr int FIB_CHUNK = 64;

__global__ void k_fib_trace(uint32_t* col, uint64_t n, uint32_t a0, uint32_t b0) {
    uint64_t start = (uint64_t(blockIdx.x) * WG + threadIdx.x) * FIB_CHUNK;
    if (start >= n) return;
    // (f, g) = (F(k), F(k+1)) in Montgomery form by fast doubling over bits of start.
    uint32_t f = 0, g = R1;
    for (int b = 63; b >= 0; --b) {
        // F(2k) = F(k) (2F(k+1) - F(k)); F(2k+1) = F(k)^2 + F(k+1)^2
        uint32_t f2 = mont_mul(f, sub(dbl(g), f));
        uint32_t g2 = add(mont_mul(f, f), mont_mul(g, g));
        if ((start >> b) & 1) {
            f = g2;
            g = add(f2, g2);
        } else {
            f = f2;
            g = g2;
        }
    }
    // F(s-1) = F(s+1) - F(s)
    uint32_t fm1 = sub(g, f);
    uint32_t a = add(mont_mul(a0, fm1), mont_mul(b0, f));       // A[s]
    uint32_t bb = add(mont_mul(a0, f), mont_mul(b0, g));        // B[s]
    uint64_t end = start + FIB_CHUNK < n ? start + FIB_CHUNK : n;
    for (uint64_t i = start; i < end; ++i) {
        col[i] = a;
        uint32_t nb = add(a, bb);
        a = bb;
        bb = nb;
    }
}

// ---------------------------------------------------------------------------
// In-LDS radix-2 decimation-in-frequency over `cnt` transforms of size
// 2^logn at s + k*stride: natural order in, bit-reversed order out.
// tw[e] = w^e (Montgomery), e < n/2, w the n-th root for this direction.
__device__ void lds_dif(uint32_t* s, int logn, int cnt, int stride, const uint32_t* tw) {
    if (logn == 0) return;
    const int half = 1 << (logn - 1);
    const int total = half * cnt;
    for (int lh = logn - 1; lh >= 0; --lh) {
        const int h = 1 << lh;
        const int tws = logn - 1 - lh;
        for (int b = threadIdx.x; b < total; b += blockDim.x) {
            int k = b >> (logn - 1);
            int bi = b & (half - 1);
            int jl = bi & (h - 1);
            int i = ((bi >> lh) << (lh + 1)) | jl;
            uint32_t* x = s + k * stride;
            uint32_t u = x[i], v = x[i + h];
            x[i] = add(u, v);
            x[i + h] = mont_mul(sub(u, v), tw[jl << tws]);
        }
        __syncthreads();
    }
}

// In-LDS radix-2 decimation-in-time: bit-reversed order in, natural out.
__device__ void lds_dit(uint32_t* s, int logn, int cnt, int stride, const uint32_t* tw) {
    if (logn == 0) return;
    const int half = 1 << (logn - 1);
    const int total = half * cnt;
    for (int lh = 0; lh < logn; ++lh) {
        const int h = 1 << lh;
        const int tws = logn - 1 - lh;
        for (int b = threadIdx.x; b < total; b += blockDim.x) {
            int k = b >> (logn - 1);
            int bi = b & (half - 1);
            int jl = bi & (h - 1);
            int i = ((bi >> lh) << (lh + 1)) | jl;
            uint32_t* x = s + k * stride;
            uint32_t u = x[i], v = mont_mul(x[i + h], tw[jl << tws]);
            x[i] = add(u, v);
            x[i + h] = sub(u, v);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ void load_tw(uint32_t* tw, const uint32_t* tw4096, int logn) {
    const int half = logn == 0 ? 0 : 1 << (logn - 1);
    for (int e = threadIdx.x; e < half; e += blockDim.x) tw[e] = tw4096[e << (12 - logn)];
}

// Single-workgroup NTT for n <= 4096.
// INTT: natural evals -> bit-reversed coeffs * 1/n.
__global__ void k_intt_small(const uint32_t* in, uint32_t* out, int logn, const uint32_t* tw4096, uint32_t ninv_m) {
    extern __shared__ uint32_t lds[];
    const int n = 1 << logn;
    uint32_t* s = lds;
    uint32_t* tw = lds + n;
    load_tw(tw, tw4096, logn);
    for (int k = threadIdx.x; k < n; k += blockDim.x) s[k] = in[k];
    __syncthreads();
    lds_dif(s, logn, 1, n, tw);
    for (int q = threadIdx.x; q < n; q += blockDim.x) out[q] = mont_mul(s[q], ninv_m);
}

// LDE coset j = blockIdx.x: bit-reversed coeffs -> coset evals (bit-reversed).
__global__ void k_lde_small(const uint32_t* in, uint32_t* out, int logn, int logb, const uint32_t* tw4096,
                            const uint32_t* scale) {
    extern __shared__ uint32_t lds[];
    const int n = 1 << logn;
    const int j = blockIdx.x;
    uint32_t* s = lds;
    uint32_t* tw = lds + n;
    load_tw(tw, tw4096, logn);
    const uint32_t* sc = scale + size_t(j) * n;
    for (int P = threadIdx.x; P < n; P += blockDim.x) {
        uint32_t k = brev_bits(P, logn);
        s[k] = mont_mul(in[P], sc[k]);
    }
    __syncthreads();
    lds_dif(s, logn, 1, n, tw);
    uint32_t* o = out + (size_t(brev_bits(j, logb)) << logn);
    for (int q = threadIdx.x; q < n; q += blockDim.x) o[q] = s[q];
}

// Two-pass NTT, n = n1 * n2. "Rows" are contiguous length-n2 segments;
// "columns" are strided length-n1 transforms handled COLS at a time.
//
// INTT (natural x[r*n2 + c]):
//   pass A (columns, over r):  Y[u*n2 + c] = w_n^{-uc} * sum_r x[r n2 + c] w_n1^{-ur}
//   pass B (rows, over c):     X[u + n1 v] at brev(u)*n2 + brev(v), times 1/n
// LDE (bit-reversed coeffs, k = r*n1 + c sits at brev(c)*n2 + brev(r)):
//   pass A (rows, over r):     Y[c*n2 + u] = s_j^c w_n^{uc} * sum_r c_k (s_j^{n1})^r w_n2^{ur}
//   pass B (columns, over c):  X[u + n2 v] at block_j*n + brev(u)*n1 + brev(v)
constexpr int COL_LDS_WORDS = 8192;

// INTT pass A. Column transforms use a padded LDS stride (n1 + 1) so the
// transposing loads and stores hit distinct banks.
__global__ void k_intt_cols(const uint32_t* x, uint32_t* y, int log1, int log2, int cols, const uint32_t* tw4096,
                            PowTable winv_m, int logb) {
    extern __shared__ uint32_t lds[];
    const int n1 = 1 << log1, n2 = 1 << log2;
    const int st = n1 + 1;
    uint32_t* s = lds;  // [cols][n1 + 1]
    uint32_t* tw = lds + cols * st;
    load_tw(tw, tw4096, log1);
    const int c0 = blockIdx.x * cols;
    for (int e = threadIdx.x; e < cols * n1; e += blockDim.x) {
        int r = e / cols, c = e % cols;
        s[c * st + r] = x[size_t(r) * n2 + c0 + c];
    }
    __syncthreads();
    lds_dif(s, log1, cols, st, tw);
    for (int e = threadIdx.x; e < cols * n1; e += blockDim.x) {
        int q = e / cols, c = e % cols;
        uint32_t u = brev_bits(q, log1);
        uint32_t cc = uint32_t(c0 + c);
        uint32_t tw_uc = pow_lookup(winv_m, (u * cc) << logb);
        y[size_t(u) * n2 + cc] = mont_mul(s[c * st + q], tw_uc);
    }
}

// INTT pass B: one row u per workgroup.
__global__ void k_intt_rows(const uint32_t* y, uint32_t* out, int log1, int log2, const uint32_t* tw4096,
                            uint32_t ninv_m) {
    extern __shared__ uint32_t lds[];
    const int n2 = 1 << log2;
    uint32_t* s = lds;
    uint32_t* tw = lds + n2;
    load_tw(tw, tw4096, log2);
    const uint32_t u = blockIdx.x;
    const uint32_t* row = y + size_t(u) * n2;
    for (int c = threadIdx.x; c < n2; c += blockDim.x) s[c] = row[c];
    __syncthreads();
    lds_dif(s, log2, 1, n2, tw);
    uint32_t* o = out + size_t(brev_bits(u, log1)) * n2;
    for (int q = threadIdx.x; q < n2; q += blockDim.x) o[q] = mont_mul(s[q], ninv_m);
}

// LDE pass A: one bit-reversed coefficient row R per workgroup, coset
// j = blockIdx.y. The row is already in bit-reversed order, so a DIT
// transform consumes it in place; rowtab is stored bit-reversed to match.
__global__ void k_lde_rows(const uint32_t* coeffs, uint32_t* y, int log1, int log2, int logb, const uint32_t* tw4096,
                           PowTable w_m, const uint32_t* rowtab_br, const uint32_t* coltab) {
    extern __shared__ uint32_t lds[];
    const int n1 = 1 << log1, n2 = 1 << log2;
    const int logn = log1 + log2;
    uint32_t* s = lds;
    uint32_t* tw = lds + n2;
    load_tw(tw, tw4096, log2);
    const uint32_t R = blockIdx.x;
    const int j = blockIdx.y;
    const uint32_t c = brev_bits(R, log1);
    const uint32_t* rt = rowtab_br + size_t(j) * n2;
    const uint32_t* in = coeffs + size_t(R) * n2;
    for (int Cc = threadIdx.x; Cc < n2; Cc += blockDim.x) s[Cc] = mont_mul(in[Cc], rt[Cc]);
    __syncthreads();
    lds_dit(s, log2, 1, n2, tw);
    const uint32_t colscale = coltab[size_t(j) * n1 + c];
    uint32_t* yo = y + (size_t(j) << logn) + size_t(c) * n2;
    for (int u = threadIdx.x; u < n2; u += blockDim.x) {
        uint32_t f = mont_mul(pow_lookup(w_m, (uint32_t(u) * c) << logb), colscale);
        yo[u] = mont_mul(s[u], f);
    }
}

// LDE pass B: `cols` adjacent columns u per workgroup, coset j = blockIdx.y.
__global__ void k_lde_cols(const uint32_t* y, uint32_t* out, int log1, int log2, int logb, int cols,
                           const uint32_t* tw4096) {
    extern __shared__ uint32_t lds[];
    const int n1 = 1 << log1, n2 = 1 << log2;
    const int logn = log1 + log2;
    const int st = n1 + 1;
    uint32_t* s = lds;  // [cols][n1 + 1]
    uint32_t* tw = lds + cols * st;
    load_tw(tw, tw4096, log1);
    const int j = blockIdx.y;
    const int u0 = blockIdx.x * cols;
    const uint32_t* yj = y + (size_t(j) << logn);
    for (int e = threadIdx.x; e < cols * n1; e += blockDim.x) {
        int c = e / cols, uu = e % cols;
        s[uu * st + c] = yj[size_t(c) * n2 + u0 + uu];
    }
    __syncthreads();
    lds_dif(s, log1, cols, st, tw);
    uint32_t* oj = out + (size_t(brev_bits(j, logb)) << logn);
    for (int e = threadIdx.x; e < cols * n1; e += blockDim.x) {
        int uu = e / n1, q = e % n1;
        uint32_t u = uint32_t(u0 + uu);
        oj[size_t(brev_bits(u, log2)) * n1 + q] = s[uu * st + q];
    }
}

__global__ void k_bitrev(const uint32_t* in, uint32_t* out, int logn) {
    uint64_t i = uint64_t(blockIdx.x) * WG + threadIdx.x;
    if (i >> logn) return;
    out[i] = in[brev_bits(uint32_t(i), logn)];
}

// ---------------------------------------------------------------------------
// Per-thread Montgomery batch inversion over K register values (all nonzero
// on the proving domains: the LDE coset 31*<Omega> avoids every 2-power
// root of unity because 31 generates the full multiplicative group).
template <int K>
__device__ __forceinline__ void batch_inv_m(uint32_t (&v)[K]) {
    uint32_t pre[K];
    pre[0] = v[0];
#pragma unroll
    for (int e = 1; e < K; ++e) pre[e] = mont_mul(pre[e - 1], v[e]);
    uint32_t inv = inv_m(pre[K - 1]);
#pragma unroll
    for (int e = K - 1; e > 0; --e) {
        uint32_t ve = v[e];
        v[e] = mont_mul(inv, pre[e - 1]);
        inv = mont_mul(inv, ve);
    }
    v[0] = inv;
}

constexpr int ELT = 16;  // elements per thread in the pointwise kernels

// Position p = block*n + q of the bit-reversed LDE -> (coset j, step t) and
// the position of the next trace row (t + 1 mod n, same coset).
struct PosInfo {
    uint32_t j, t, pnext;
};
__device__ __forceinline__ PosInfo pos_info(uint32_t p, int logn, int logb) {
    uint32_t block = p >> logn;
    uint32_t q = p & ((1u << logn) - 1);
    uint32_t t = brev_bits(q, logn);
    uint32_t t1 = (t + 1) & ((1u << logn) - 1);
    PosInfo r;
    r.j = brev_bits(block, logb);
    r.t = t;
    r.pnext = (block << logn) | brev_bits(t1, logn);
    return r;
}

// Point x = s_j * omega_n^t in Montgomery form; omega_n = Omega_m^B.
__device__ __forceinline__ uint32_t point_m(PowTable w_m, const uint32_t* shift_m, uint32_t j, uint32_t t, int logb) {
    return mont_mul(pow_lookup(w_m, t << logb), shift_m[j]);
}

// B[i] = A[i + B] + coeff*(x^n - 1)/(x - w_last); coeff_van[j] =
// coeff*(s_j^n - 1) (canonical) is constant per coset since x^n = s_j^n there.
__global__ void k_linked_b(const uint32_t* a, uint32_t* b, int logn, int logb, PowTable w_m,
                           const uint32_t* shift_m, const uint32_t* coeff_van, uint32_t w_last_m) {
    const uint32_t m = 1u << (logn + logb);
    const uint32_t base = blockIdx.x * (WG * ELT) + threadIdx.x;
    uint32_t den[ELT];
    PosInfo pi[ELT];
#pragma unroll
    for (int e = 0; e < ELT; ++e) {
        uint32_t p = base + e * WG;
        if (p < m) {
            pi[e] = pos_info(p, logn, logb);
            den[e] = sub(point_m(w_m, shift_m, pi[e].j, pi[e].t, logb), w_last_m);
        } else {
            den[e] = R1;
        }
    }
    batch_inv_m<ELT>(den);
#pragma unroll
    for (int e = 0; e < ELT; ++e) {
        uint32_t p = base + e * WG;
        if (p < m) {
            // den is (1/(x - w_last)) in Montgomery form, coeff_van canonical.
            b[p] = add(a[pi[e].pnext], mont_mul(den[e], coeff_van[pi[e].j]));
        }
    }
}

// Composition value for one point (all scalars canonical, alpha Montgomery):
//   a0*(A'-B)/Z + a1*(B'-A-B)/Z + a2*(A-a0)/(x-1) + a3*(B-b0)/(x-1) + a4*(B-bl)/(x-w)
__global__ void k_compose(const uint32_t* a, const uint32_t* b, uint4* layer0, int logn, int logb, PowTable w_m,
                          const uint32_t* shift_m, const uint32_t* zinv_m, ComposeConsts c) {
    const uint32_t m = 1u << (logn + logb);
    const uint32_t base = blockIdx.x * (WG * ELT) + threadIdx.x;
    uint32_t d1[ELT], dw[ELT], pr[ELT];
#pragma unroll
    for (int e = 0; e < ELT; ++e) {
        uint32_t p = base + e * WG;
        if (p < m) {
            PosInfo pi = pos_info(p, logn, logb);
            uint32_t x = point_m(w_m, shift_m, pi.j, pi.t, logb);
            d1[e] = sub(x, R1);
            dw[e] = sub(x, c.w_last_m);
            pr[e] = mont_mul(d1[e], dw[e]);
        } else {
            d1[e] = dw[e] = pr[e] = R1;
        }
    }
    batch_inv_m<ELT>(pr);
#pragma unroll
    for (int e = 0; e < ELT; ++e) {
        uint32_t p = base + e * WG;
        if (p >= m) continue;
        PosInfo pi = pos_info(p, logn, logb);
        uint32_t i1 = mont_mul(dw[e], pr[e]);                 // 1/(x-1), Montgomery
        uint32_t iw = mont_mul(d1[e], pr[e]);                 // 1/(x-w), Montgomery
        uint32_t iz = mont_mul(dw[e], zinv_m[pi.j]);          // (x-w)/(x^n-1), Montgomery
        uint32_t va = a[p], vb = b[p], na = a[pi.pnext], nb = b[pi.pnext];
        uint32_t s0 = mont_mul(sub(na, vb), iz);
        uint32_t s1 = mont_mul(sub(nb, add(va, vb)), iz);
        uint32_t s2 = mont_mul(sub(va, c.a0), i1);
        uint32_t s3 = mont_mul(sub(vb, c.b0), i1);
        uint32_t s4 = mont_mul(sub(vb, c.b_last), iw);
        uint32_t o[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            uint32_t acc = mont_mul(s0, c.alpha_m[0][k]);
            acc = add(acc, mont_mul(s1, c.alpha_m[1][k]));
            acc = add(acc, mont_mul(s2, c.alpha_m[2][k]));
            acc = add(acc, mont_mul(s3, c.alpha_m[3][k]));
            acc = add(acc, mont_mul(s4, c.alpha_m[4][k]));
            o[k] = acc;
        }
        layer0[p] = make_uint4(o[0], o[1], o[2], o[3]);
    }
}

// ---------------------------------------------------------------------------
// Merkle subtree reduction in LDS. Digests held word-major (SoA) so a
// thread reads both children of word k with one 8-byte access.
constexpr int SUB_IN = 2 * WG;  // digests entering the LDS reduction per workgroup

// cur: [8][SUB_IN] current level in LDS (count valid entries), writes
// parent levels 1..nlevels to global at level_ptr(l) + first_index(l).
__device__ void subtree_levels(uint32_t (*cur)[SUB_IN], uint32_t (*nxt)[SUB_IN], int count, int nlevels,
                               uint32_t* out_level_base, uint64_t level_count, uint64_t group_first) {
    // out_level_base points at the level that `cur` belongs to; subsequent
    // levels follow contiguously (level sizes level_count >> l).
    uint32_t* lvl = out_level_base;
    uint64_t lc = level_count;
    uint64_t first = group_first;
    for (int l = 1; l <= nlevels && count > 1; ++l) {
        uint32_t* next_lvl = lvl + lc * 8;
        int pc = count >> 1;
        for (int t = threadIdx.x; t < pc; t += blockDim.x) {
            uint32_t L[8], Rr[8], o[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                L[k] = cur[k][2 * t];
                Rr[k] = cur[k][2 * t + 1];
            }
            node_words(L, Rr, o);
#pragma unroll
            for (int k = 0; k < 8; ++k) nxt[k][t] = o[k];
            uint4* g = reinterpret_cast<uint4*>(next_lvl + (first / 2 + t) * 8);
            g[0] = make_uint4(o[0], o[1], o[2], o[3]);
            g[1] = make_uint4(o[4], o[5], o[6], o[7]);
        }
        __syncthreads();
        auto tmp = cur;
        cur = nxt;
        nxt = tmp;
        count = pc;
        lvl = next_lvl;
        lc >>= 1;
        first >>= 1;
    }
}

// Generic: digests of one level -> up to nlevels parent levels.
__global__ void k_merkle_levels(uint32_t* level_base, uint64_t count, int nlevels) {
    __shared__ uint32_t bufA[8][SUB_IN];
    __shared__ uint32_t bufB[8][SUB_IN];
    uint64_t first = uint64_t(blockIdx.x) * SUB_IN;
    int here = int(count - first < uint64_t(SUB_IN) ? count - first : uint64_t(SUB_IN));
    for (int t = threadIdx.x; t < here; t += blockDim.x) {
        const uint4* g = reinterpret_cast<const uint4*>(level_base + (first + t) * 8);
        uint4 x = g[0], y = g[1];
        bufA[0][t] = x.x; bufA[1][t] = x.y; bufA[2][t] = x.z; bufA[3][t] = x.w;
        bufA[4][t] = y.x; bufA[5][t] = y.y; bufA[6][t] = y.z; bufA[7][t] = y.w;
    }
    __syncthreads();
    subtree_levels(bufA, bufB, here, nlevels, level_base, count, first);
}

// FRI leaves: 4 consecutive values per leaf (Ext4 or base-field values with
// zero upper coordinates), then fused parent levels.
__global__ void k_fri_leaves(const uint32_t* layer, int base_field, uint32_t* levels, uint64_t leaves, int nlevels) {
    __shared__ uint32_t bufA[8][SUB_IN];
    __shared__ uint32_t bufB[8][SUB_IN];
    uint64_t first = uint64_t(blockIdx.x) * SUB_IN;
    int here = int(leaves - first < uint64_t(SUB_IN) ? leaves - first : uint64_t(SUB_IN));
    for (int t = threadIdx.x; t < here; t += blockDim.x) {
        uint64_t leaf = first + t;
        uint32_t m[16];
        if (base_field) {
            uint4 v = reinterpret_cast<const uint4*>(layer)[leaf];
            m[0] = v.x; m[4] = v.y; m[8] = v.z; m[12] = v.w;
            m[1] = m[2] = m[3] = m[5] = m[6] = m[7] = 0;
            m[9] = m[10] = m[11] = m[13] = m[14] = m[15] = 0;
        } else {
            const uint4* src = reinterpret_cast<const uint4*>(layer) + leaf * 4;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                uint4 v = src[k];
                m[4 * k] = v.x; m[4 * k + 1] = v.y; m[4 * k + 2] = v.z; m[4 * k + 3] = v.w;
            }
        }
        uint32_t d[8];
        leaf64_words(m, d);
        uint4* g = reinterpret_cast<uint4*>(levels + leaf * 8);
        g[0] = make_uint4(d[0], d[1], d[2], d[3]);
        g[1] = make_uint4(d[4], d[5], d[6], d[7]);
#pragma unroll
        for (int k = 0; k < 8; ++k) bufA[k][t] = d[k];
    }
    __syncthreads();
    subtree_levels(bufA, bufB, here, nlevels, levels, leaves, first);
}

// Trace leaves: leaf L packs natural rows 8L..8L+7 as A||B (LE words).
// Thread g handles the rows of trace steps t sharing one position q of the
// bit-reversed blocks: for B >= 8 one step t = brev(q) (B/8 leaves, reduced
// in registers to one node), for B = 4 the step pair (t, t+1) with t even.
__global__ void k_trace_leaves(const uint32_t* a, const uint32_t* b, uint32_t* levels, int logn, int logb,
                               int max_level) {
    const uint32_t n = 1u << logn;
    const int B = 1 << logb;
    uint32_t g = blockIdx.x * WG + threadIdx.x;
    const uint64_t leaves = (uint64_t(n) << logb) >> 3;
    if (logb >= 3) {
        if (g >= n) return;
        const uint32_t q = g;
        const uint32_t t = brev_bits(q, logn);
        const int per = B >> 3;  // leaves per step
        // Leaves L = per*t + h, h < per; rows j = 8h + r.
        uint32_t node[2][8];  // B = 16: both leaves of step t stay in registers
        for (int h = 0; h < per; ++h) {
            uint32_t msg[16];
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                uint32_t j = uint32_t(8 * h + r);
                uint32_t pos = (brev_bits(j, logb) << logn) | q;
                msg[2 * r] = a[pos];
                msg[2 * r + 1] = b[pos];
            }
            uint32_t d[8];
            leaf64_words(msg, d);
            uint64_t L = uint64_t(per) * t + h;
            uint4* gp = reinterpret_cast<uint4*>(levels + L * 8);
            gp[0] = make_uint4(d[0], d[1], d[2], d[3]);
            gp[1] = make_uint4(d[4], d[5], d[6], d[7]);
            if (per == 2) {
#pragma unroll
                for (int k = 0; k < 8; ++k) node[h & 1][k] = d[k];
            }
        }
        if (per == 2 && max_level >= 1) {
            uint32_t o[8];
            node_words(node[0], node[1], o);
            uint4* gp = reinterpret_cast<uint4*>(levels + (leaves + t) * 8);
            gp[0] = make_uint4(o[0], o[1], o[2], o[3]);
            gp[1] = make_uint4(o[4], o[5], o[6], o[7]);
        }
    } else {
        // B = 4: leaf L = t/2 covers steps t (even) and t+1, cosets 0..3 each.
        if (g >= n / 2) return;
        const uint32_t q = g;                 // top bit of q clear -> t even
        const uint32_t t = brev_bits(q, logn);
        const uint32_t q1 = q | (n >> 1);     // position of t + 1
        uint32_t msg[16];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            uint32_t j = uint32_t(r & 3);
            uint32_t pos = (brev_bits(j, 2) << logn) | (r < 4 ? q : q1);
            msg[2 * r] = a[pos];
            msg[2 * r + 1] = b[pos];
        }
        uint32_t d[8];
        leaf64_words(msg, d);
        uint64_t L = t >> 1;
        uint4* gp = reinterpret_cast<uint4*>(levels + L * 8);
        gp[0] = make_uint4(d[0], d[1], d[2], d[3]);
        gp[1] = make_uint4(d[4], d[5], d[6], d[7]);
    }
}

// ---------------------------------------------------------------------------
// Arity-4 fold of a bit-reversed layer. Quad i sits at (x, -x, j0 x, -j0 x)
// with x = point of position 4i; x^{-1} = 31^{-4^d} * Omega_m^{-(brev(i) << 2d)}.
__global__ void k_fold(const uint32_t* layer, int base_field, uint4* next, PowTable winv_m, FoldConsts c) {
    const uint64_t quads = uint64_t(1) << (c.log_len - 2);
    uint64_t i = uint64_t(blockIdx.x) * WG + threadIdx.x;
    if (i >= quads) return;
    uint32_t e = brev_bits(uint32_t(i), c.log_len - 2);
    uint32_t inv_x = mont_mul(pow_lookup(winv_m, e << (2 * c.depth)), c.ginv_m);  // Montgomery
    E4 q[4];
    if (base_field) {
        uint4 v = reinterpret_cast<const uint4*>(layer)[i];
        q[0] = e4_base(v.x); q[1] = e4_base(v.y); q[2] = e4_base(v.z); q[3] = e4_base(v.w);
    } else {
        const uint4* src = reinterpret_cast<const uint4*>(layer) + i * 4;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            uint4 v = src[k];
            q[k] = E4{{v.x, v.y, v.z, v.w}};
        }
    }
    E4 s01 = e4_add(q[0], q[1]), s23 = e4_add(q[2], q[3]);
    E4 odd_diff = e4_sub(q[0], q[1]);
    E4 skew_diff = e4_scale_m(e4_sub(q[2], q[3]), c.unit_m);
    E4 mean = e4_halve(e4_halve(e4_add(s01, s23)));
    E4 odd = e4_halve(e4_halve(e4_sub(odd_diff, skew_diff)));
    E4 even = e4_halve(e4_halve(e4_sub(s01, s23)));
    E4 skew = e4_halve(e4_halve(e4_add(odd_diff, skew_diff)));
    E4 ch0{{c.ch_m[0][0], c.ch_m[0][1], c.ch_m[0][2], c.ch_m[0][3]}};
    E4 ch1{{c.ch_m[1][0], c.ch_m[1][1], c.ch_m[1][2], c.ch_m[1][3]}};
    E4 ch2{{c.ch_m[2][0], c.ch_m[2][1], c.ch_m[2][2], c.ch_m[2][3]}};
    E4 acc = e4_scale_m(e4_mul_m(skew, ch2), inv_x);
    acc = e4_scale_m(e4_add(acc, e4_mul_m(even, ch1)), inv_x);
    acc = e4_add(acc, e4_mul_m(odd, ch0));
    E4 out = e4_add(mean, e4_scale_m(acc, inv_x));
    next[i] = make_uint4(out.v[0], out.v[1], out.v[2], out.v[3]);
}

// ---------------------------------------------------------------------------
// Proof-of-work scan. The 55-byte preimage head||nonce is one BLAKE3 block
// (START|END|ROOT). Workgroups walk chunks in increasing order and stop once
// a passing nonce below their next chunk is known, so the reported minimum
// is the exact smallest passing nonce (every chunk at or below it is scanned).
constexpr int GRIND_PER_THREAD = 4;
constexpr uint32_t POW_HEAD_LEN = 47;  // 32-byte history digest + u64 len + "fri-pow"

__global__ void k_grind(uint4 h0, uint4 h1, uint4 h2, uint4 h3, uint32_t bits, uint64_t limit,
                        unsigned long long* result) {
    const uint32_t hw[16] = {h0.x, h0.y, h0.z, h0.w, h1.x, h1.y, h1.z, h1.w,
                             h2.x, h2.y, h2.z, h2.w, h3.x, h3.y, h3.z, h3.w};
    const uint64_t chunk = uint64_t(WG) * GRIND_PER_THREAD;
    const uint64_t stride = uint64_t(gridDim.x) * chunk;
    for (uint64_t start = uint64_t(blockIdx.x) * chunk; start < limit; start += stride) {
        if (__hip_atomic_load(result, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < start) return;
        for (int k = 0; k < GRIND_PER_THREAD; ++k) {
            uint64_t nonce = start + uint64_t(k) * WG + threadIdx.x;
            if (nonce >= limit) break;
            uint32_t m[16];
#pragma unroll
            for (int w = 0; w < 16; ++w) m[w] = hw[w];
            // nonce bytes occupy preimage bytes 47..54: word 11 byte 3, words 12-13
            uint32_t lo = uint32_t(nonce), hi = uint32_t(nonce >> 32);
            m[11] |= lo << 24;
            m[12] = (lo >> 8) | (hi << 24);
            m[13] = hi >> 8;
            uint32_t iv[8], d[8];
            b3_iv(iv);
            compress8(iv, m, 0, 0, POW_HEAD_LEN + 8, b3::CHUNK_START | b3::CHUNK_END | b3::ROOT, d);
            uint32_t be = __builtin_bswap32(d[0]);
            uint32_t lz = be == 0 ? 32u : uint32_t(__builtin_clz(be));
            if (lz >= bits) atomicMin(result, (unsigned long long)nonce);
        }
    }
}

// ---------------------------------------------------------------------------
__global__ void k_gather(const uint32_t* base, const uint64_t* src, const uint32_t* dst, const uint32_t* len,
                         uint32_t items, uint32_t* out) {
    uint32_t i = blockIdx.x * WG + threadIdx.x;
    if (i >= items) return;
    const uint32_t* s = base + src[i];
    uint32_t* d = out + dst[i];
    for (uint32_t k = 0; k < len[i]; ++k) d[k] = s[k];
}

// Plain streaming copy of `n16` 16-byte words: the calibration reference for
// the FETCH_SIZE / WRITE_SIZE hardware counters.
__global__ void k_copy(const uint4* __restrict__ in, uint4* __restrict__ out, uint64_t n16) {
    uint64_t i = uint64_t(blockIdx.x) * WG + threadIdx.x;
    const uint64_t stride = uint64_t(gridDim.x) * WG;
    for (; i < n16; i += stride) out[i] = in[i];
}

inline unsigned grid_for(uint64_t work, uint64_t per_block) { return unsigned((work + per_block - 1) / per_block); }

}  // namespace

NttShape ntt_shape(int logn) {
    NttShape s;
    s.logn = logn;
    // Rows (contiguous, DIT/DIF in LDS) take up to 4096 points; the strided
    // column pass takes the remainder, COLS_WORDS / n1 columns at a time.
    if (logn <= 12) {
        s.log1 = 0;
        s.log2 = logn;
    } else {
        s.log2 = 12;
        s.log1 = logn - 12;
    }
    return s;
}


void launch_fib_trace(const Launch& L, uint32_t* col, uint64_t n, uint32_t a0, uint32_t b0) {
    unsigned g = grid_for((n + FIB_CHUNK - 1) / FIB_CHUNK, WG);
    ProfScope ps(L.prof, "fib_trace", L.stream);
    hipLaunchKernelGGL(k_fib_trace, dim3(g), dim3(WG), 0, L.stream, col, n, a0, b0);
}

void launch_intt_to_br(const Launch& L, const uint32_t* evals, uint32_t* coeffs_br, uint32_t* scratch, int logn,
                       const uint32_t* tw4096_inv, PowTable winv_m, int log_blowup, uint32_t ninv_m) {
    NttShape s = ntt_shape(logn);
    if (s.log1 == 0) {
        size_t lds = (size_t(1) << logn) * 4 + (size_t(1) << logn) * 2;
        ProfScope ps(L.prof, "intt_small", L.stream);
        hipLaunchKernelGGL(k_intt_small, dim3(1), dim3(WG), lds, L.stream, evals, coeffs_br, logn, tw4096_inv,
                           ninv_m);
        return;
    }
    const int n1 = 1 << s.log1, n2 = 1 << s.log2;
    int cols = COL_LDS_WORDS >> s.log1;
    if (cols > n2) cols = n2;
    {
        size_t lds = size_t(cols) * (n1 + 1) * 4 + size_t(n1) * 2;
        ProfScope ps(L.prof, "intt_cols", L.stream);
        hipLaunchKernelGGL(k_intt_cols, dim3(n2 / cols), dim3(WG), lds, L.stream, evals, scratch, s.log1, s.log2,
                           cols, tw4096_inv, winv_m, log_blowup);
    }
    {
        size_t lds = size_t(n2) * 4 + size_t(n2) * 2;
        ProfScope ps(L.prof, "intt_rows", L.stream);
        hipLaunchKernelGGL(k_intt_rows, dim3(n1), dim3(WG), lds, L.stream, scratch, coeffs_br, s.log1, s.log2,
                           tw4096_inv, ninv_m);
    }
}

void launch_lde_br(const Launch& L, const uint32_t* coeffs_br, uint32_t* lde_br, uint32_t* scratch, int logn,
                   int log_blowup, const uint32_t* tw4096, PowTable w_m, CosetTables ct) {
    NttShape s = ntt_shape(logn);
    const int B = 1 << log_blowup;
    if (s.log1 == 0) {
        size_t lds = (size_t(1) << logn) * 4 + (size_t(1) << logn) * 2;
        ProfScope ps(L.prof, "lde_small", L.stream);
        hipLaunchKernelGGL(k_lde_small, dim3(B), dim3(WG), lds, L.stream, coeffs_br, lde_br, logn, log_blowup,
                           tw4096, ct.small);
        return;
    }
    const int n1 = 1 << s.log1, n2 = 1 << s.log2;
    {
        size_t lds = size_t(n2) * 4 + size_t(n2) * 2;
        ProfScope ps(L.prof, "lde_rows", L.stream);
        hipLaunchKernelGGL(k_lde_rows, dim3(n1, B), dim3(WG), lds, L.stream, coeffs_br, scratch, s.log1, s.log2,
                           log_blowup, tw4096, w_m, ct.row, ct.col);
    }
    int cols = COL_LDS_WORDS >> s.log1;
    if (cols > n2) cols = n2;
    {
        size_t lds = size_t(cols) * (n1 + 1) * 4 + size_t(n1) * 2;
        ProfScope ps(L.prof, "lde_cols", L.stream);
        hipLaunchKernelGGL(k_lde_cols, dim3(n2 / cols, B), dim3(WG), lds, L.stream, scratch, lde_br, s.log1, s.log2,
                           log_blowup, cols, tw4096);
    }
}

void launch_bitrev(const Launch& L, const uint32_t* in, uint32_t* out, int logn) {
    ProfScope ps(L.prof, "bitrev", L.stream);
    hipLaunchKernelGGL(k_bitrev, dim3(grid_for(uint64_t(1) << logn, WG)), dim3(WG), 0, L.stream, in, out, logn);
}

void launch_linked_b(const Launch& L, const uint32_t* a_br, uint32_t* b_br, int logn, int log_blowup, PowTable w_m,
                     const uint32_t* shift_m, const uint32_t* coeff_van, uint32_t w_last_m) {
    uint64_t m = uint64_t(1) << (logn + log_blowup);
    ProfScope ps(L.prof, "linked_b", L.stream);
    hipLaunchKernelGGL(k_linked_b, dim3(grid_for(m, WG * ELT)), dim3(WG), 0, L.stream, a_br, b_br, logn, log_blowup,
                       w_m, shift_m, coeff_van, w_last_m);
}

void launch_compose(const Launch& L, const uint32_t* a_br, const uint32_t* b_br, uint32_t* layer0, int logn,
                    int log_blowup, PowTable w_m, const uint32_t* shift_m, const uint32_t* zinv_m,
                    const ComposeConsts& c) {
    uint64_t m = uint64_t(1) << (logn + log_blowup);
    ProfScope ps(L.prof, "compose", L.stream);
    hipLaunchKernelGGL(k_compose, dim3(grid_for(m, WG * ELT)), dim3(WG), 0, L.stream, a_br, b_br,
                       reinterpret_cast<uint4*>(layer0), logn, log_blowup, w_m, shift_m, zinv_m, c);
}

int launch_trace_leaves(const Launch& L, const uint32_t* a_br, const uint32_t* b_br, uint32_t* levels, int logn,
                        int log_blowup, int max_level) {
    uint64_t n = uint64_t(1) << logn;
    uint64_t threads = log_blowup >= 3 ? n : n / 2;
    {
        ProfScope ps(L.prof, "trace_leaves", L.stream);
        hipLaunchKernelGGL(k_trace_leaves, dim3(grid_for(threads, WG)), dim3(WG), 0, L.stream, a_br, b_br, levels,
                           logn, log_blowup, max_level);
    }
    // levels produced in-thread beyond the leaf level
    return (log_blowup == 4 && max_level >= 1) ? 1 : 0;
}

void launch_merkle_levels(const Launch& L, uint32_t* level_base, uint64_t count, int nlevels) {
    ProfScope ps(L.prof, "merkle_levels", L.stream);
    hipLaunchKernelGGL(k_merkle_levels, dim3(grid_for(count, SUB_IN)), dim3(WG), 0, L.stream, level_base, count,
                       nlevels);
}

void launch_fri_leaves(const Launch& L, const uint32_t* layer, bool base_field, uint32_t* levels, uint64_t leaves,
                       int nlevels) {
    ProfScope ps(L.prof, base_field ? "fri_leaves_base" : "fri_leaves", L.stream);
    hipLaunchKernelGGL(k_fri_leaves, dim3(grid_for(leaves, SUB_IN)), dim3(WG), 0, L.stream, layer,
                       base_field ? 1 : 0, levels, leaves, nlevels);
}

void launch_fold(const Launch& L, const uint32_t* layer, bool base_field, uint32_t* next, PowTable winv_m,
                 const FoldConsts& c) {
    uint64_t quads = uint64_t(1) << (c.log_len - 2);
    ProfScope ps(L.prof, base_field ? "fold_base" : "fold", L.stream);
    hipLaunchKernelGGL(k_fold, dim3(grid_for(quads, WG)), dim3(WG), 0, L.stream, layer, base_field ? 1 : 0,
                       reinterpret_cast<uint4*>(next), winv_m, c);
}

void launch_grind(const Launch& L, const uint32_t hw[16], uint32_t head_len, uint32_t bits, uint64_t limit,
                  unsigned long long* result) {
    ProfScope ps(L.prof, "grind", L.stream);
    uint4 h0 = make_uint4(hw[0], hw[1], hw[2], hw[3]), h1 = make_uint4(hw[4], hw[5], hw[6], hw[7]);
    uint4 h2 = make_uint4(hw[8], hw[9], hw[10], hw[11]), h3 = make_uint4(hw[12], hw[13], hw[14], hw[15]);
    (void)head_len;  // fixed 47-byte head, checked by the caller
    hipLaunchKernelGGL(k_grind, dim3(2048), dim3(WG), 0, L.stream, h0, h1, h2, h3, bits, limit, result);
}

void launch_copy(const Launch& L, const uint32_t* in, uint32_t* out, uint64_t words) {
    ProfScope ps(L.prof, "copy", L.stream);
    hipLaunchKernelGGL(k_copy, dim3(4096), dim3(WG), 0, L.stream, reinterpret_cast<const uint4*>(in),
                       reinterpret_cast<uint4*>(out), words / 4);
}

void launch_gather(const Launch& L, const uint32_t* base, const uint64_t* src, const uint32_t* dst,
                   const uint32_t* len, uint32_t items, uint32_t* out) {
    ProfScope ps(L.prof, "gather", L.stream);
    hipLaunchKernelGGL(k_gather, dim3(grid_for(items, WG)), dim3(WG), 0, L.stream, base, src, dst, len, items, out);
}

}  // namespace z0p
