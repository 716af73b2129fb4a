// Shared device helpers of the mean-field kernels (anc_meanfield.hip: one launch per step; anc_mf_team.hip: persistent
// teams): block reductions and the per-instance vector pass that quantises a vector into the four base-256 digit
// rows of the exact integer product (see anc_meanfield.hip for the arithmetic). Both kernel families call these, so
// they round and sum in the same order and give identical floats.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "anc_device.hpp"

namespace anc {
namespace mf {

using dev::v4i;

constexpr int THREADS = 256;
constexpr int ROWS = 64;              // rows per sub-block: 4 waves x 16 rows
constexpr int UNIT_COLS = 256;        // columns per unit: 4 MFMAs of 64 per wave and 16 rows
constexpr int DEPTH = 4;              // units whose J loads are in flight per wave
constexpr int GEMV_BLOCKS = 1024;     // workgroups per launch the row groups are sized for (256 CUs x 4)

// Butterfly reductions over the 64 lanes of a wave: at distance o every lane combines its value with lane i ^ o's
// (o = 32, 16, 8, 4, 2, 1), so every lane ends with the same result. The partners come from lane moves that do not go
// through LDS addresses: the two cross-row swaps (v_permlane32_swap / v_permlane16_swap return {own, partner} in
// either order, and the combining operations are commutative), a row rotation by 8, a swizzle and two quad
// permutations. The pairing is that of the xor butterfly, so sums round as its additions do.
template <int O>
__device__ __forceinline__ uint32_t xor_lane(uint32_t v) {
    if constexpr (O == 1) return uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0xB1, 0xF, 0xF, false));    // quad_perm [1,0,3,2]
    else if constexpr (O == 2) return uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x4E, 0xF, 0xF, false)); // quad_perm [2,3,0,1]
    else if constexpr (O == 4) return uint32_t(__builtin_amdgcn_ds_swizzle(int(v), 0x101F));                    // bit mode, xor 4
    else return uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x128, 0xF, 0xF, false));                       // row_ror:8
}

template <class Op>
__device__ __forceinline__ float wave_reduce(float v, Op op) {
    {
        const auto a = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
        v = op(__uint_as_float(a[0]), __uint_as_float(a[1]));
    }
    {
        const auto a = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
        v = op(__uint_as_float(a[0]), __uint_as_float(a[1]));
    }
    v = op(v, __uint_as_float(xor_lane<8>(__float_as_uint(v))));
    v = op(v, __uint_as_float(xor_lane<4>(__float_as_uint(v))));
    v = op(v, __uint_as_float(xor_lane<2>(__float_as_uint(v))));
    v = op(v, __uint_as_float(xor_lane<1>(__float_as_uint(v))));
    return v;
}

__device__ __forceinline__ float wave_sum(float v) {
    return wave_reduce(v, [](float x, float y) { return x + y; });
}

__device__ __forceinline__ float wave_max(float v) {
    return wave_reduce(v, [](float x, float y) { return fmaxf(x, y); });
}

// The FP64 butterfly in two parts: the xor-32, -16 and -8 levels (afterwards lane i holds the sum over the eight lanes
// i mod 8, i mod 8 + 8, ..., i mod 8 + 56), then the xor-4, -2 and -1 levels; wave_sum_d is both.
__device__ __forceinline__ void split_d(double d, uint32_t& lo, uint32_t& hi) {
    const uint64_t u = __double_as_longlong(d);
    lo = uint32_t(u);
    hi = uint32_t(u >> 32);
}
__device__ __forceinline__ double join_d(uint32_t lo, uint32_t hi) {
    return __longlong_as_double((long long)((uint64_t(hi) << 32) | lo));
}
__device__ __forceinline__ double wave_sum_d_hi(double v) {
    uint32_t lo, hi;
    split_d(v, lo, hi);
    {
        const auto a = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
        const auto b = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
        v = join_d(a[0], b[0]) + join_d(a[1], b[1]);
    }
    split_d(v, lo, hi);
    {
        const auto a = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
        const auto b = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
        v = join_d(a[0], b[0]) + join_d(a[1], b[1]);
    }
    split_d(v, lo, hi);
    return v + join_d(xor_lane<8>(lo), xor_lane<8>(hi));
}
__device__ __forceinline__ double wave_sum_d_lo(double v) {
    uint32_t lo, hi;
    split_d(v, lo, hi);
    v += join_d(xor_lane<4>(lo), xor_lane<4>(hi));
    split_d(v, lo, hi);
    v += join_d(xor_lane<2>(lo), xor_lane<2>(hi));
    split_d(v, lo, hi);
    return v + join_d(xor_lane<1>(lo), xor_lane<1>(hi));
}
__device__ __forceinline__ double wave_sum_d(double v) { return wave_sum_d_lo(wave_sum_d_hi(v)); }

// Block-wide reductions of per-thread values (all threads receive them); the instance-level sums that feed the
// schedule gradients are FP64 so that they do not lose the small differences they are made of.
__device__ __forceinline__ float block_sum(float v, float* red) {
    v = wave_sum(v);
    const int w = threadIdx.x / 64;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
}
__device__ __forceinline__ double block_sum_d(double v, double* red) {
    v = wave_sum_d(v);
    const int w = threadIdx.x / 64;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
}
// Three block sums at once (the same additions as three block_sum_d calls, one pair of barriers); red: 12 doubles.
__device__ __forceinline__ void block_sum_d3(double& a, double& b, double& c, double* red) {
    a = wave_sum_d(a);
    b = wave_sum_d(b);
    c = wave_sum_d(c);
    const int w = threadIdx.x / 64;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        red[w] = a;
        red[4 + w] = b;
        red[8 + w] = c;
    }
    __syncthreads();
    a = red[0] + red[1] + red[2] + red[3];
    b = red[4] + red[5] + red[6] + red[7];
    c = red[8] + red[9] + red[10] + red[11];
}
// block_max of mx and block_sum_d3 of a, b, c in one pair of barriers (the same results); red: 16 doubles.
__device__ __forceinline__ void block_max_sum_d3(float& mx, double& a, double& b, double& c, double* red) {
    mx = wave_max(mx);
    a = wave_sum_d(a);
    b = wave_sum_d(b);
    c = wave_sum_d(c);
    const int w = threadIdx.x / 64;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        red[w] = a;
        red[4 + w] = b;
        red[8 + w] = c;
        reinterpret_cast<float*>(red + 12)[w] = mx;
    }
    __syncthreads();
    a = red[0] + red[1] + red[2] + red[3];
    b = red[4] + red[5] + red[6] + red[7];
    c = red[8] + red[9] + red[10] + red[11];
    const float* rf = reinterpret_cast<const float*>(red + 12);
    mx = fmaxf(fmaxf(rf[0], rf[1]), fmaxf(rf[2], rf[3]));
}
__device__ __forceinline__ float block_max(float v, float* red) {
    v = wave_max(v);
    const int w = threadIdx.x / 64;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    return fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

// ---------------------------------------------------------------------------------------------------------- digits
// LDS digit vectors: digit k of column c at xl[k * lstride + c], columns padded with zeros to a multiple of 256;
// the 16-byte row offset spreads the four digit rows over different LDS banks.
__host__ __device__ inline int digit_stride(int n) { return UNIT_COLS * ((n + UNIT_COLS - 1) / UNIT_COLS) + 16; }

// Block exponent: s with max|v| 2^s in [2^29, 2^30) (0 for an all-zero vector).
__device__ __forceinline__ int block_shift(float maxabs) {
    if (!(maxabs > 0.f)) return 0;
    int e;
    (void)frexpf(maxabs, &e);
    return 30 - e;
}

// ---------------------------------------------------------------------------------------------------------- vector
// One pass over the instance vector in 4-column groups (group gi = columns 4 gi .. 4 gi + 3, gi = threadIdx.x +
// THREADS j): ld(i, raw) loads the NV raw values of column i (clamped index), val(raw) makes the value to quantise,
// acc(raw, value) accumulates the caller's instance-level sums (columns < n only). The block maximum of |value| sets
// the shift s, and the digits are written to LDS. For n <= 4 RG THREADS the values stay in registers between the
// maximum and the quantisation (every load of the pass is issued before the first use); larger n take a second pass
// that recomputes the values. Returns s.
constexpr int RG = 2;

// The four digit bytes of columns 4 gi .. 4 gi + 3 (v quantised with shift s; columns >= n give 0).
// The balanced base-256 digits of q (|q| <= 2^30; digits in [-128, 127], a unique representation) are the bytes of
// u = q + 0x80808080 (unsigned digits, no borrow out of the top byte) each minus 128, i.e. the bytes of
// u XOR 0x80808080 as int8. The four values' bytes are transposed into the four digit rows by byte permutes.
__device__ __forceinline__ void write_digits(int8_t* xl, int lstride, int n, int gi, const float (&v)[4], int s) {
    uint32_t u[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int q = 4 * gi + c < n ? int(rintf(ldexpf(v[c], s))) : 0;
        u[c] = uint32_t(q) + 0x80808080u;
    }
    // {u0.b0, u1.b0, u0.b1, u1.b1}, {u0.b2, u1.b2, u0.b3, u1.b3}, and the same of u2, u3
    const uint32_t lo01 = __builtin_amdgcn_perm(u[1], u[0], 0x05010400u), hi01 = __builtin_amdgcn_perm(u[1], u[0], 0x07030602u);
    const uint32_t lo23 = __builtin_amdgcn_perm(u[3], u[2], 0x05010400u), hi23 = __builtin_amdgcn_perm(u[3], u[2], 0x07030602u);
    const uint32_t w[4] = {__builtin_amdgcn_perm(lo23, lo01, 0x05040100u), __builtin_amdgcn_perm(lo23, lo01, 0x07060302u),
                           __builtin_amdgcn_perm(hi23, hi01, 0x05040100u), __builtin_amdgcn_perm(hi23, hi01, 0x07060302u)};
#pragma unroll
    for (int k = 0; k < 4; ++k) *reinterpret_cast<uint32_t*>(xl + k * lstride + 4 * gi) = w[k] ^ 0x80808080u;
}

// The vector pass for n <= 4 RG THREADS from raw values already in registers: raw[j][c] belongs to column
// 4 (threadIdx.x + THREADS j) + c (values of columns >= n are never used).
// Its two halves: the per-thread values, their maximum and the caller's sums (val(j, c) and acc(j, c, value) for
// column 4 (threadIdx.x + THREADS j) + c), then the digits for shift s.
template <class Val, class Acc>
__device__ __forceinline__ float vector_scan(int n, float (&v)[RG][4], Val&& val, Acc&& acc) {
    float mx = 0.f;
#pragma unroll
    for (int j = 0; j < RG; ++j)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int i = 4 * (int(threadIdx.x) + THREADS * j) + c;
            v[j][c] = val(j, c);
            if (i < n) {
                mx = fmaxf(mx, fabsf(v[j][c]));
                acc(j, c, v[j][c]);
            }
        }
    return mx;
}
__device__ __forceinline__ void vector_write(const float (&v)[RG][4], int8_t* xl, int lstride, int n, int s) {
    const int groups = (lstride - 16) / 4;
#pragma unroll
    for (int j = 0; j < RG; ++j) {
        const int gi = int(threadIdx.x) + THREADS * j;
        if (gi < groups) write_digits(xl, lstride, n, gi, v[j], s);
    }
}
template <int NV, class Val, class Acc>
__device__ __forceinline__ int vector_regs(const float (&raw)[RG][4][NV], int8_t* xl, int lstride, int n, float* red, Val&& val,
                                           Acc&& acc) {
    float v[RG][4];
    const float mx = vector_scan(n, v, [&](int j, int c) { return val(raw[j][c]); },
                                 [&](int j, int c, float x) { acc(raw[j][c], x); });
    const int s = block_shift(block_max(mx, red));
    vector_write(v, xl, lstride, n, s);
    return s;
}

template <int NV, class Ld, class Val, class Acc>
__device__ __forceinline__ int stage_vector(int8_t* xl, int lstride, int n, float* red, Ld&& ld, Val&& val, Acc&& acc) {
    const int groups = (lstride - 16) / 4;
    if (n <= 4 * RG * THREADS) {
        float raw[RG][4][NV];
#pragma unroll
        for (int j = 0; j < RG; ++j)
#pragma unroll
            for (int c = 0; c < 4; ++c) ld(min(4 * (int(threadIdx.x) + THREADS * j) + c, n - 1), raw[j][c]);
        return vector_regs<NV>(raw, xl, lstride, n, red, val, acc);
    }
    float mx = 0.f;
    for (int gi = threadIdx.x; gi < groups; gi += THREADS) {
        float raw[4][NV];
#pragma unroll
        for (int c = 0; c < 4; ++c) ld(min(4 * gi + c, n - 1), raw[c]);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            if (4 * gi + c < n) {
                const float x = val(raw[c]);
                mx = fmaxf(mx, fabsf(x));
                acc(raw[c], x);
            }
        }
    }
    const int s = block_shift(block_max(mx, red));
    for (int gi = threadIdx.x; gi < groups; gi += THREADS) {
        float raw[4][NV], v[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) ld(min(4 * gi + c, n - 1), raw[c]);
#pragma unroll
        for (int c = 0; c < 4; ++c) v[c] = val(raw[c]);
        write_digits(xl, lstride, n, gi, v, s);
    }
    return s;
}

}  // namespace mf
}  // namespace anc
