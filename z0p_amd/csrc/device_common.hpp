// Device helpers shared by the kernels: bit reversal, power-table lookups,
// Montgomery inversion, LDS transforms, batched inversion, layout helpers,
// the LDS Merkle subtree reduction and the job-table lookup.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "blake3.hpp"
#include "field.hpp"
#include "kernels.hpp"

namespace z0p {
namespace dev {

constexpr int WG = 256;

// Compile-time loop: f(integral_constant<int, I>) for I in [B, E). Register
// arrays indexed by the constant stay in registers even where the compiler
// would not fully unroll an ordinary loop with a large body.
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// Descriptor pointers are generic; these casts state that they address
// global memory, so the compiler emits global (not flat) accesses and does
// not have to order them against LDS traffic.
#define Z0P_GLOBAL __attribute__((address_space(1)))
typedef const Z0P_GLOBAL uint32_t* gcptr;
typedef Z0P_GLOBAL uint32_t* gptr;
__device__ __forceinline__ gcptr gc(const uint32_t* p) { return (gcptr)p; }
__device__ __forceinline__ gptr gm(uint32_t* p) { return (gptr)p; }
// 16-byte global accesses (native vector type: HIP's uint4 has no
// address-space-qualified operators).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ u32x4 ld4(const uint32_t* p, uint64_t i) {
    return reinterpret_cast<const Z0P_GLOBAL u32x4*>(gc(p))[i];
}
__device__ __forceinline__ void st4(uint32_t* p, uint64_t i, u32x4 v) {
    reinterpret_cast<Z0P_GLOBAL u32x4*>(gm(p))[i] = v;
}

__device__ __forceinline__ uint32_t brev_bits(uint32_t x, int bits) {
    return bits == 0 ? 0u : (__builtin_bitreverse32(x) >> (32 - bits));
}

__device__ __forceinline__ uint32_t pow_lookup(PowTable t, uint32_t e) {
    return mont_mul(gc(t.lo)[e & 4095u], gc(t.hi)[e >> 12]);
}

// Montgomery-domain inverse: (a^-1)_m from a_m (zero maps to zero).
// Exponent p-2 = 0b111_0_(27 ones): window-3 chain, 30 squarings + 11 mults.
__device__ __forceinline__ uint32_t inv_m(uint32_t a) {
    uint32_t a2 = mont_mul(a, a);
    uint32_t a3 = mont_mul(a2, a);
    uint32_t a6 = mont_mul(a3, a3);
    uint32_t a7 = mont_mul(a6, a);
    uint32_t acc = mont_mul(a7, a7);
#pragma unroll
    for (int w = 0; w < 9; ++w) {
        acc = mont_mul(acc, acc);
        acc = mont_mul(acc, acc);
        acc = mont_mul(acc, acc);
        acc = mont_mul(acc, a7);
    }
    return acc;
}

// Radix-2 decimation-in-frequency in LDS over `cnt` transforms of size
// 2^logn at s + k*stride: natural order in, bit-reversed out.
__device__ __forceinline__ void lds_dif(uint32_t* s, int logn, int cnt, int stride, const uint32_t* tw) {
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

// Radix-2 decimation-in-time in LDS: bit-reversed order in, natural out.
__device__ __forceinline__ void lds_dit(uint32_t* s, int logn, int cnt, int stride, const uint32_t* tw) {
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

// tw[e] = w_n^e for e < n/2 from the 4096-point table.
__device__ __forceinline__ void load_tw(uint32_t* tw, const uint32_t* tw4096, int logn) {
    const int half = logn == 0 ? 0 : 1 << (logn - 1);
    const gcptr src = gc(tw4096);
    for (int e = threadIdx.x; e < half; e += blockDim.x) tw[e] = src[e << (12 - logn)];
}

// Per-thread Montgomery batch inversion of K register values (nonzero on
// every proving domain: the coset 31*<Omega> avoids all 2-power roots of
// unity because 31 generates the full multiplicative group).
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

// Position p = block*n + q of the bit-reversed LDE -> coset j, step t and
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

// x = s_j * omega_n^t (Montgomery form); omega_n = Omega_m^B.
__device__ __forceinline__ uint32_t point_m(PowTable w_m, const uint32_t* shift_m, uint32_t j, uint32_t t, int logb) {
    return mont_mul(pow_lookup(w_m, t << logb), gc(shift_m)[j]);
}

// Word offset of level l in a level-contiguous tree whose level 0 has L
// digests (l <= log2 L): 8 * (L + L/2 + ... + L/2^(l-1)).
__device__ __forceinline__ uint64_t tree_level_words(uint64_t L, int l) { return 16 * (L - (L >> l)); }

__device__ __forceinline__ void store_digest(uint32_t* dst, const uint32_t d[8]) {
    st4(dst, 0, u32x4{d[0], d[1], d[2], d[3]});
    st4(dst, 1, u32x4{d[4], d[5], d[6], d[7]});
}

// This workgroup's job: table sorted by block_begin (njobs > 0), or the
// inline descriptor when njobs == 0. `local` is the block index inside the job.
template <typename D>
__device__ __forceinline__ D job_of(const D& inline_d, const D* table, uint32_t njobs, uint32_t& local) {
    const uint32_t b = blockIdx.x;
    if (njobs == 0) {
        local = b;
        return inline_d;
    }
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (table[mid].block_begin <= b) lo = mid;
        else hi = mid;
    }
    local = b - table[lo].block_begin;
    return table[lo];
}

}  // namespace dev
}  // namespace z0p
