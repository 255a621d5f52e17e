// Device context: one HIP stream, a bump-allocated device arena reused across
// proofs, pinned host staging, and cached per-geometry constant tables.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <vector>

#include "kernels.hpp"
#include "profiler.hpp"
#include "status.hpp"

namespace z0p {

inline void hip_check(hipError_t e) {
    if (e != hipSuccess) {
        if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) throw Status(Z0P_ERR_OUT_OF_MEMORY);
        throw Status(Z0P_ERR_DEVICE);
    }
}

// Device-resident constant tables for one LDE geometry (log n, log B).
struct GeometryTables {
    int logn = 0, logb = 0;
    std::vector<void*> owned;
    PowTable w{}, winv{};           // omega_m^{+e}, omega_m^{-e}
    CosetTables coset{};
    const uint32_t* zinv = nullptr;  // [B] 1/(s_j^n - 1), Montgomery
    ~GeometryTables() {
        for (void* p : owned) (void)hipFree(p);
    }
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    Profiler prof;

    uint8_t* arena = nullptr;
    size_t arena_cap = 0;
    size_t arena_used = 0;

    uint8_t* pinned = nullptr;
    size_t pinned_cap = 0;

    uint32_t* tw4096 = nullptr;      // [2048] omega_4096^e, Montgomery
    uint32_t* tw4096_inv = nullptr;  // [2048] omega_4096^{-e}, Montgomery
    std::map<std::pair<int, int>, std::unique_ptr<GeometryTables>> tables;

    Launch launch() { return Launch{stream, &prof}; }

    // Ensure the arena holds `bytes`; only valid between proofs.
    void reserve_arena(size_t bytes) {
        if (bytes <= arena_cap) return;
        hip_check(hipStreamSynchronize(stream));
        if (arena) hip_check(hipFree(arena));
        arena = nullptr;
        arena_cap = 0;
        size_t cap = bytes + (bytes >> 3);
        hip_check(hipMalloc(&arena, cap));
        arena_cap = cap;
    }
    void reset_arena() { arena_used = 0; }
    template <typename T>
    T* alloc(size_t count) {
        size_t bytes = (count * sizeof(T) + 255) & ~size_t(255);
        require(arena_used + bytes <= arena_cap, Z0P_ERR_OUT_OF_MEMORY);
        T* p = reinterpret_cast<T*>(arena + arena_used);
        arena_used += bytes;
        return p;
    }
    void reserve_pinned(size_t bytes) {
        if (bytes <= pinned_cap) return;
        hip_check(hipStreamSynchronize(stream));
        if (pinned) hip_check(hipHostFree(pinned));
        pinned = nullptr;
        pinned_cap = 0;
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&pinned), bytes, hipHostMallocDefault));
        pinned_cap = bytes;
    }
    void sync() {
        hip_check(hipStreamSynchronize(stream));
        if (prof.enabled) prof.resolve();
    }

    const GeometryTables& geometry(int logn, int logb);

    ~Ctx();
};

}  // namespace z0p
