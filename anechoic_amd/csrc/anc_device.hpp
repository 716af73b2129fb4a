// Device helpers shared by the team kernels (anc_engine.hip) and the large-instance kernel (anc_large.hip): the
// per-lane RNG (fpga/v80_sca/src/sca_ref.hpp:71-91), the status words, the tagged exchange of flip fragments.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace anc {
namespace dev {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

// bit k of v (k < 8) moved to bit 4k
__device__ __forceinline__ uint32_t spread_nibbles(uint32_t v) {
    v = (v | (v << 12)) & 0x000F000Fu;
    v = (v | (v << 6)) & 0x03030303u;
    return (v | (v << 3)) & 0x11111111u;
}

// bit k of v (k < 8) moved to bit 2k
__device__ __forceinline__ uint32_t spread_crumbs(uint32_t v) {
    v = (v | (v << 4)) & 0x0F0Fu;
    v = (v | (v << 2)) & 0x3333u;
    return (v | (v << 1)) & 0x5555u;
}

constexpr int LANES = 256;
constexpr int THREADS = 256;

// ---------------------------------------------------------------------------------------------------------- RNG
__device__ __forceinline__ uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint32_t s0, s1, s2, s3;
    __device__ __forceinline__ void seed(uint64_t seed, uint32_t trial, uint32_t lane) {
        uint64_t x = seed ^ (uint64_t(trial) << 20) ^ (uint64_t(lane) << 48) ^ 0x5CA0F00Dull;
        const uint64_t a = splitmix(x), b = splitmix(x);
        s0 = uint32_t(a); s1 = uint32_t(a >> 32); s2 = uint32_t(b); s3 = uint32_t(b >> 32);
        if (!(s0 | s1 | s2 | s3)) s0 = 1;
    }
    __device__ __forceinline__ uint32_t next() {
        const uint32_t r = __builtin_rotateleft32(s1 * 5u, 7) * 9u, t = s1 << 9;
        s2 ^= s0; s3 ^= s1; s1 ^= s2; s0 ^= s3; s2 ^= t; s3 = __builtin_rotateleft32(s3, 11);
        return r;
    }
};

// ---------------------------------------------------------------------------------------------------------- status
// Status block (int[4], device memory): [0] abort flag of the current launch (cleared before every launch), [1] a
// word nothing ever sets (the host-forced wait of the error-path tests polls it), [2] the first error code since the
// host last read the block (kept across launches, cleared by the host when it reports the error).
__device__ __forceinline__ void raise_error(int* status, int code) {
    atomicCAS(status + 2, 0, code);
    __hip_atomic_store(status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool aborted(int* status) {
    return __hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

// ---------------------------------------------------------------------------------------------------------- barrier
// Workgroup barrier that also tells every wave whether any thread arrived with its flag set: one s_barrier and one
// LDS read (__syncthreads_or takes three barriers, a cross-lane reduction and a load of the workgroup size). `cell`
// is an LDS word zeroed before its first use; a set cell stays set (the callers leave their loops when it reads
// set). Rule: two calls with no other barrier between them use different cells. Then a wave writes a cell only after
// passing a barrier that every wave reached after its last read of that cell, so all waves read the same value.
// (Plain LDS accesses: the barrier's fence keeps the write before it and the read after it.)
__device__ __forceinline__ bool barrier_any(bool flag, int& cell) {
    if (flag) cell = 1;
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(cell) != 0;
}

// Workgroup barrier of a pipeline of direct-to-LDS copies: every wave first waits (s_waitcnt WAITCNT, an encoded
// immediate) for the copies the barrier publishes -- vector memory counters retire in order, so waiting until at most
// the copies issued after them are outstanding suffices -- and then meets the others. Unlike __syncthreads, whose
// workgroup fence drains every outstanding vector memory operation (the copies of later stages included, since
// direct-to-LDS copies are LDS writes counted by vmcnt), the younger copies stay in flight. The empty asm statements
// keep the compiler from moving memory accesses across the barrier.
template <int WAITCNT>
__device__ __forceinline__ void copy_barrier() {
    __builtin_amdgcn_s_waitcnt(WAITCNT);
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    __asm__ volatile("" ::: "memory");
}

// Sum over aligned groups of W in {1, 2, 4, 8} lanes, returned in every lane of the group (DPP lane moves).
template <int W>
__device__ __forceinline__ int32_t group_sum(int32_t v) {
    static_assert(W == 1 || W == 2 || W == 4 || W == 8, "group width");
    if (W >= 2) v += __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);    // quad_perm [1,0,3,2]
    if (W >= 4) v += __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);    // quad_perm [2,3,0,1]
    if (W >= 8) v += __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false);   // row_half_mirror: other quad
    return v;
}

// ---------------------------------------------------------------------------------------------------------- exchange
// Exchange load flavours (cache-policy bits of the buffer loads; stores are device-scope relaxed atomics):
//   1: sc1 -- device scope: correct for any placement of the team's workgroups (misses the XCD's L2);
//   2: nt  -- misses the CU's L1 and is served by the XCD's L2, which also holds the same-XCD producers' stores:
//             fast for a team on one XCD (workgroup b runs on XCD b % 8). A team split over XCDs would only see
//             stale lines, run into the bounded wait and report a time-out (with the context option "fallback" set,
//             the context then switches to 1).
// Correctness never depends on the flavour: every word carries its own validity tag.
template <int XFLAVOR>
__device__ __forceinline__ v4i xload16(__amdgpu_buffer_rsrc_t r, int off) {
    return __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, XFLAVOR >= 2 ? 2 : 16);
}
template <int XFLAVOR>
__device__ __forceinline__ uint32_t xload4(__amdgpu_buffer_rsrc_t r, int off) {
    return __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, XFLAVOR >= 2 ? 2 : 16);
}
//   3: nt loads and plain stores -- the producers' stores stay in the XCD's L2 (write-back) and the consumers read
//      them there; again only for a team on one XCD.
template <int XFLAVOR>
__device__ __forceinline__ void xstore4(uint32_t* p, uint32_t v) {
    if (XFLAVOR == 3) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);   // no cache bits
    else __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int XFLAVOR>
__device__ __forceinline__ void xstore2(uint16_t* p, uint16_t v) {
    if (XFLAVOR == 3) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    else __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t tau_of(int u) { return (uint32_t(u >> 1) & 1u) ^ 1u; }
// number of u' in [0, u] with tau(u') = 1 (pattern 1,1,0,0,...)
__device__ __forceinline__ int tau_count(int u) { return 2 * ((u + 1) >> 2) + min((u + 1) & 3, 2); }

// Byte d in {-2, 0, 2} with the step's tau in bit 0 (d is even, so OR = ADD: the byte is D + tau).
__device__ __forceinline__ uint32_t dbyte(int d, uint32_t tau) { return (uint32_t(d) & 0xFFu) | tau; }   // d even

}  // namespace dev
}  // namespace anc
