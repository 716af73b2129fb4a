// Differentiable mean-field relaxation of the synchronous SCA dynamics with the Onsager reaction term (build-defined;
// no reference implementation). Per instance b (dense int8 J_b, FP32 magnetisations m in [-1, 1]^N):
//
//   u_t     = J m_t + q m_t - beta_t m_{t-1}
//   m_{t+1} = tanh(u_t / (2 T_t))
//   beta_t  = coef_t                          (temperature-scaled mode: coef_t = kappa gamma T_{t-1} rho_t)
//   beta_t  = coef_t * sum_i (1 - m_{t,i}^2)  (counted mode: coef_t = lambda_t gamma / (2 T_{t-1}))
//
// This is the expectation of the discrete decision (fpga/v80_sca/src/sca_ref_bias.hpp:53-56): with the threshold
// uniform on [-2T, 2T), E[s_new] = clip((h + q s - corr s') / 2T); tanh replaces the clip so that gradients exist
// everywhere, and sum_i (1 - m_i^2) replaces the band count n_lin (Sigma f' = n_lin / 2T, the reference's F2).
//
// Kernels (one launch per step, 256 threads; every block stages the instance's whole input vector in LDS, zero-padded
// to a multiple of 16, and recomputes the instance-level reductions redundantly; each wave computes 16 rows at a time):
//   mf_fwd_step   m_{t+1} (and the tape u_t, beta_t for the backward recomputation)
//   mf_bwd_step   adjoint: delta_t = g_{t+1} (1 - m_{t+1}^2) / (2 T_t);
//                 g_t = J delta_t + q delta_t + 2 coef_t S2_t m_t [counted] - beta_{t+1} delta_{t+1}
//                 dL/dT_t (direct) = -(1/T_t) sum_i delta_i u_i ; dL/dcoef_t = -S2_t (n_t in counted mode); these
//                 instance-level sums are accumulated in FP64
//                 with S2_t = sum_i delta_{t,i} m_{t-1,i}, n_t = sum_i (1 - m_{t,i}^2)
//   mf_gemv       y = J x (energies and the loss gradient)
// for int8 J (grid = (64-row blocks, instances)), their *_stream counterparts for packed J (grid = (row groups,
// instances), each block walking several 64-row sub-blocks), and
//   mf_pack4      packed 4-bit copy of int8 J (entries in [-8, 7])
//
// Packed storage: two entries per byte, row stride ldj bytes: in each 4-byte group g of a row, byte e holds column
// 8g + e in its low nibble and column 8g + 4 + e in its high nibble (two's complement). The GEMV reads the same
// 16-column chunks per lane as for int8 J (8 instead of 16 bytes per row) and unpacks a 4-bit word w as
// (w << 4) & 0xF0F0F0F0 (columns 8g..8g+3) and w & 0xF0F0F0F0 (8g+4..8g+7): bytes that hold 16 J exactly, converted
// by the same byte-select conversion as int8 J. Every product and partial sum is then exactly 16 times the int8
// storage's (a power-of-two scale commutes with FP32 rounding) and the row sum is scaled by 2^-4 (ldexp): both
// storages give identical floats, and the packed one streams half the bytes.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "anc_kernels.hpp"

namespace anc {
namespace mf {

constexpr int THREADS = 256;
constexpr int ROWS = 64;   // rows per block: 4 waves x 16 rows

__device__ __forceinline__ float wave_sum(float v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ double wave_sum_d(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Block-wide sums of per-thread values (all threads receive them); FP64 so that the instance-level reductions that
// feed the schedule gradients do not lose the small differences they are made of.
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

// Per-thread walk over i = threadIdx.x + THREADS j < count (increasing i) in batches of STAGE_BATCH: ld(i, v) loads
// the NV values of entry i for every i of a batch (unconditional loads at clamped indices, so that the whole batch
// is in flight before the first use), then use(i, v) consumes them in order. A plain loop waits for each load before
// issuing the next, which made the staging of a 2000-entry vector cost several memory latencies per workgroup.
constexpr int STAGE_BATCH = 8;
template <int NV, class Ld, class Use>
__device__ __forceinline__ void batched_walk(int count, Ld&& ld, Use&& use) {
    for (int i0 = threadIdx.x; i0 < count; i0 += STAGE_BATCH * THREADS) {
        float v[STAGE_BATCH][NV];
#pragma unroll
        for (int k = 0; k < STAGE_BATCH; ++k) ld(min(i0 + k * THREADS, count - 1), v[k]);
#pragma unroll
        for (int k = 0; k < STAGE_BATCH; ++k)
            if (i0 + k * THREADS < count) use(i0 + k * THREADS, v[k]);
    }
}

// Stages x (n entries) in LDS, zero-padded to 16 ceil(n / 16) entries.
__device__ __forceinline__ void stage_padded(float* xs, int n, const float* src) {
    const int np16 = 16 * ((n + 15) / 16);
    batched_walk<1>(np16, [&](int i, float* v) { v[0] = src[min(i, n - 1)]; },
                    [&](int i, const float* v) { xs[i] = i < n ? v[0] : 0.f; });
}

// Dot products of the 16 rows r0..r0+15 of int8 J (row stride ldj) with the zero-padded LDS vector x. Lane L reads
// the 16-column chunks L, L + 64, ... of all 16 rows (16 independent 16-byte loads per chunk), then the 16 partial
// sums are reduce-scattered over the wave (halving exchanges at lane distances 32, 16, 8, 4, then a sum over 2 and 1):
// lane L ends with the full sum of row r0 + rowsel(L), rowsel(L) = 8 b5 + 4 b4 + 2 b3 + b2 (b_k = bit k of L), the
// same in the four lanes of each aligned quad.
__device__ __forceinline__ float rows16(const int8_t* __restrict__ Jb, int64_t ldj, int n, int r0, const float* x, int& rowsel) {
    const int L = threadIdx.x & 63;
    float acc[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.f;
    const int chunks = (n + 15) / 16;
    for (int ch = L; ch < chunks; ch += 64) {
        const int x0 = ch * 16;
        int4 w[16];
#pragma unroll
        for (int i = 0; i < 16; ++i)
            w[i] = r0 + i < n ? *reinterpret_cast<const int4*>(Jb + size_t(r0 + i) * size_t(ldj) + x0) : int4{0, 0, 0, 0};
        const float4* xq = reinterpret_cast<const float4*>(x + x0);
        const float4 xa = xq[0], xb = xq[1], xc = xq[2], xd = xq[3];
        const float xv[16] = {xa.x, xa.y, xa.z, xa.w, xb.x, xb.y, xb.z, xb.w, xc.x, xc.y, xc.z, xc.w, xd.x, xd.y, xd.z, xd.w};
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int wd[4] = {w[i].x, w[i].y, w[i].z, w[i].w};
            float a0 = acc[i];
#pragma unroll
            for (int d = 0; d < 4; ++d)
#pragma unroll
                for (int e = 0; e < 4; ++e) a0 += float(int8_t(wd[d] >> (8 * e))) * xv[4 * d + e];
            acc[i] = a0;
        }
    }
    // reduce-scatter: at distance o, a lane keeps the lower half of its values if bit o of L is clear, else the upper
    // half, and adds its partner's copy of the half it keeps
    float v8[8], v4[4], v2[2];
    const bool h32 = L & 32, h16 = L & 16, h8 = L & 8, h4 = L & 4;
#pragma unroll
    for (int i = 0; i < 8; ++i) v8[i] = (h32 ? acc[i + 8] : acc[i]) + __shfl_xor(h32 ? acc[i] : acc[i + 8], 32);
#pragma unroll
    for (int i = 0; i < 4; ++i) v4[i] = (h16 ? v8[i + 4] : v8[i]) + __shfl_xor(h16 ? v8[i] : v8[i + 4], 16);
#pragma unroll
    for (int i = 0; i < 2; ++i) v2[i] = (h8 ? v4[i + 2] : v4[i]) + __shfl_xor(h8 ? v4[i] : v4[i + 2], 8);
    float v = (h4 ? v2[1] : v2[0]) + __shfl_xor(h4 ? v2[0] : v2[1], 4);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 1);
    rowsel = (h32 ? 8 : 0) + (h16 ? 4 : 0) + (h8 ? 2 : 0) + (h4 ? 1 : 0);
    return v;
}

template <bool COUNTED>
__global__ void __launch_bounds__(THREADS) mf_fwd_step(MFArgs a, int t, const float* __restrict__ m_prev,
                                                       const float* __restrict__ m_cur, float* __restrict__ m_next,
                                                       float* __restrict__ u_tape, float* __restrict__ beta_tape) {
    extern __shared__ float xs[];
    __shared__ float red[4];
    const int b = blockIdx.y, n = a.n;
    const float* mc = m_cur + size_t(b) * a.n;
    float nsoft = 0.f;
    const int np16 = 16 * ((n + 15) / 16);
    batched_walk<1>(np16, [&](int i, float* v) { v[0] = mc[min(i, n - 1)]; }, [&](int i, const float* vv) {
        const float v = i < n ? vv[0] : 0.f;
        xs[i] = v;
        nsoft += i < n ? 1.f - v * v : 0.f;
    });
    __syncthreads();
    float beta = a.coef[t];
    if (COUNTED) beta *= block_sum(nsoft, red);
    if (t == 0) beta = 0.f;
    const int w = threadIdx.x / 64, L = threadIdx.x & 63;
    const int r0 = blockIdx.x * ROWS + 16 * w;
    int sel;
    const float hv = rows16(a.J + size_t(b) * size_t(a.bstride), a.ldj, n, r0, xs, sel);
    const float inv2T = 0.5f / a.T[t];
    if ((L & 3) == 0) {
        const int r = r0 + sel;
        if (r < n) {
            const size_t idx = size_t(b) * n + r;
            const float u = hv + a.q * xs[r] - beta * m_prev[idx];
            m_next[idx] = tanhf(u * inv2T);
            if (u_tape) u_tape[idx] = u;
        }
    }
    if (beta_tape && blockIdx.x == 0 && threadIdx.x == 0) beta_tape[b] = beta;
}

template <bool COUNTED>
__global__ void __launch_bounds__(THREADS) mf_bwd_step(MFArgs a, int t, const float* __restrict__ g_next,
                                                       const float* __restrict__ m_next, const float* __restrict__ m_cur,
                                                       const float* __restrict__ m_prev, const float* __restrict__ u_t,
                                                       const float* __restrict__ delta_next, const float* __restrict__ beta_next,
                                                       float* __restrict__ delta_out, float* __restrict__ g_out,
                                                       double* __restrict__ dT, double* __restrict__ dcoef) {
    extern __shared__ float xs[];
    __shared__ double red[4];
    const int b = blockIdx.y, n = a.n;
    const size_t off = size_t(b) * n;
    const float inv2T = 0.5f / a.T[t];
    double s1 = 0.0, s2 = 0.0, nsoft = 0.0;
    const int np16 = 16 * ((n + 15) / 16);
    for (int i = n + threadIdx.x; i < np16; i += THREADS) xs[i] = 0.f;
    batched_walk<5>(n, [&](int i, float* v) {
        v[0] = m_next[off + i]; v[1] = g_next[off + i]; v[2] = u_t[off + i]; v[3] = m_prev[off + i];
        v[4] = COUNTED ? m_cur[off + i] : 0.f;
    }, [&](int i, const float* v) {
        const float mn = v[0];
        const float d = v[1] * (1.f - mn * mn) * inv2T;
        xs[i] = d;
        s1 += double(d) * double(v[2]);
        s2 += double(d) * double(v[3]);
        if (COUNTED) { const double mc = v[4]; nsoft += 1.0 - mc * mc; }
    });
    __syncthreads();
    s1 = block_sum_d(s1, red);
    s2 = block_sum_d(s2, red);
    if (COUNTED) nsoft = block_sum_d(nsoft, red);
    if (t == 0) s2 = 0.0;   // beta_0 = 0: no dependence on m_{-1} or coef_0
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        dT[size_t(t) * a.B + b] = -s1 / double(a.T[t]);
        dcoef[size_t(t) * a.B + b] = COUNTED ? -s2 * nsoft : -s2;
    }
    const int w = threadIdx.x / 64, L = threadIdx.x & 63;
    const int r0 = blockIdx.x * ROWS + 16 * w;
    int sel;
    const float v = rows16(a.J + size_t(b) * size_t(a.bstride), a.ldj, n, r0, xs, sel);
    if ((L & 3) == 0) {
        const int r = r0 + sel;
        if (r < n) {
            float g = v + a.q * xs[r];
            if (COUNTED) g += float(2.0 * double(a.coef[t]) * s2) * m_cur[off + r];
            if (delta_next) g -= beta_next[b] * delta_next[off + r];
            g_out[off + r] = g;
            delta_out[off + r] = xs[r];
        }
    }
}

__global__ void __launch_bounds__(THREADS) mf_gemv(MFArgs a, const float* __restrict__ x, float* __restrict__ y) {
    extern __shared__ float xs[];
    const int b = blockIdx.y, n = a.n;
    stage_padded(xs, n, x + size_t(b) * n);
    __syncthreads();
    const int w = threadIdx.x / 64, L = threadIdx.x & 63;
    const int r0 = blockIdx.x * ROWS + 16 * w;
    int sel;
    const float h = rows16(a.J + size_t(b) * size_t(a.bstride), a.ldj, n, r0, xs, sel);
    if ((L & 3) == 0 && r0 + sel < n) y[size_t(b) * n + r0 + sel] = h;
}

// Packed 4-bit copy of int8 J: one thread per 4-byte group (8 columns) of a row; block (row, instance). Columns at
// or beyond n become 0; entries outside [-8, 7] are counted in *bad (the copy is then not usable).
__global__ void __launch_bounds__(THREADS) mf_pack4(const int8_t* __restrict__ J, int64_t ldj, int64_t bstride, int n,
                                                     uint8_t* __restrict__ J4, int64_t ldj4, int64_t bstride4, int* bad) {
    const int r = blockIdx.x, b = blockIdx.y;
    const int8_t* src = J + size_t(b) * size_t(bstride) + size_t(r) * size_t(ldj);
    uint32_t* dst = reinterpret_cast<uint32_t*>(J4 + size_t(b) * size_t(bstride4) + size_t(r) * size_t(ldj4));
    int nbad = 0;
    for (int g = threadIdx.x; g < int(ldj4 / 4); g += THREADS) {
        uint32_t word = 0;
        for (int e = 0; e < 8; ++e) {
            const int col = 8 * g + e;
            const int v = col < n ? int(src[col]) : 0;
            nbad += (v < -8 || v > 7) ? 1 : 0;
            const int byte = e & 3, hi = e >> 2;
            word |= (uint32_t(v) & 0xFu) << (8 * byte + 4 * hi);
        }
        dst[g] = word;
    }
    if (nbad) atomicAdd(bad, nbad);
}

// ---------------------------------------------------------------------------------------------------------- packed
// Row-stream kernels for the packed storage. A block stages its instance's vector once and walks `nsb` consecutive
// 64-row sub-blocks (each wave 16 rows of each), so that fewer, longer-lived blocks keep the chip full. A wave's work
// is a sequence of units (sub-block, chunk iteration): 16 rows x the lane's 16-column chunk, 16 loads of 8 bytes.
// Two unit buffers alternate: the loads of unit u + 2 are issued right after unit u's arithmetic, so one unit's loads
// are always in flight while the other is being multiplied. Loads go through a buffer resource bounded by the
// instance's n rows: rows >= n read as 0 without a branch. Per row and lane the chunks are accumulated in the same
// order as rows16 (int8 storage) and the reduce-scatter is the same; with the exact 16 J bytes and the 2^-4 scale
// (above) the results equal the int8 kernels' bit for bit, and the epilogues below repeat their expressions.
constexpr int STREAM_BLOCKS = 512;   // blocks kept resident: 256 CUs x 2 (8 waves per CU)

// Loads one unit (rows r0..r0+15, chunk ch); base offsets at or beyond n ldj read as 0 without touching memory.
__device__ __forceinline__ void load_unit4(int (&w)[16][2], __amdgpu_buffer_rsrc_t rs, int ldj, int base) {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(rs, base + i * ldj, 0, 0);
        w[i][0] = int(v[0]);
        w[i][1] = int(v[1]);
    }
}

__device__ __forceinline__ void fma_unit4(float (&acc)[16], const int (&w)[16][2], const float* x, int x0) {
    const float4* xq = reinterpret_cast<const float4*>(x + x0);
    const float4 xa = xq[0], xb = xq[1], xc = xq[2], xd = xq[3];
    const float xv[16] = {xa.x, xa.y, xa.z, xa.w, xb.x, xb.y, xb.z, xb.w, xc.x, xc.y, xc.z, xc.w, xd.x, xd.y, xd.z, xd.w};
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const uint32_t a = uint32_t(w[i][0]), c = uint32_t(w[i][1]);
        const int wd[4] = {int((a << 4) & 0xF0F0F0F0u), int(a & 0xF0F0F0F0u), int((c << 4) & 0xF0F0F0F0u), int(c & 0xF0F0F0F0u)};
        float a0 = acc[i];
#pragma unroll
        for (int d = 0; d < 4; ++d)
#pragma unroll
            for (int e = 0; e < 4; ++e) a0 += float(int8_t(wd[d] >> (8 * e))) * xv[4 * d + e];
        acc[i] = a0;
    }
}

// The reduce-scatter of rows16 (the sum of row rowsel(L) in lane L) followed by the 2^-4 scale, acc cleared.
__device__ __forceinline__ float reduce_rows4(float (&acc)[16], int& rowsel) {
    const int L = threadIdx.x & 63;
    float v8[8], v4[4], v2[2];
    const bool h32 = L & 32, h16 = L & 16, h8 = L & 8, h4 = L & 4;
#pragma unroll
    for (int i = 0; i < 8; ++i) v8[i] = (h32 ? acc[i + 8] : acc[i]) + __shfl_xor(h32 ? acc[i] : acc[i + 8], 32);
#pragma unroll
    for (int i = 0; i < 4; ++i) v4[i] = (h16 ? v8[i + 4] : v8[i]) + __shfl_xor(h16 ? v8[i] : v8[i + 4], 16);
#pragma unroll
    for (int i = 0; i < 2; ++i) v2[i] = (h8 ? v4[i + 2] : v4[i]) + __shfl_xor(h8 ? v4[i] : v4[i + 2], 8);
    float v = (h4 ? v2[1] : v2[0]) + __shfl_xor(h4 ? v2[0] : v2[1], 4);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 1);
    rowsel = (h32 ? 8 : 0) + (h16 ? 4 : 0) + (h8 ? 2 : 0) + (h4 ? 1 : 0);
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.f;
    return __builtin_ldexpf(v, -4);
}

// Rows row_base .. row_base + 64 nsb - 1 of instance J (packed, n rows of ldj bytes) times the LDS vector x; calls
// epi(r, value) once per row r < n (in lane 4k of the wave that owns it). The loads are unconditional (the ones past
// the last unit point beyond the instance and read nothing) and the odd last unit runs after the loop: with no
// branch around a load, the compiler's memory-counter waits stay at "one unit outstanding" in the loop.
template <class Epi>
__device__ __forceinline__ void stream_rows4(const int8_t* J, int64_t ldj64, int n, int row_base, int nsb, const float* x, Epi&& epi) {
    const int w = threadIdx.x / 64, L = threadIdx.x & 63;
    const int ldj = int(ldj64), end = n * ldj;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<int8_t*>(J), 0, end, 0x00020000);
    const int chunks = (n + 15) / 16, C = (chunks + 63) / 64, U = nsb * C;
    int A[16][2], B[16][2];
    float acc[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.f;
    auto row0 = [&](int u) { return row_base + 64 * (u / C) + 16 * w; };
    auto chunk = [&](int u) { return L + 64 * (u % C); };
    auto base = [&](int u) { return u < U ? row0(u) * ldj + 8 * chunk(u) : end; };
    auto step = [&](int (&buf)[16][2], int u) {
        const int ch = chunk(u);
        if (ch < chunks) fma_unit4(acc, buf, x, 16 * ch);
        load_unit4(buf, rs, ldj, base(u + 2));
        if (u % C == C - 1) {
            int sel;
            const float v = reduce_rows4(acc, sel);
            const int r = row0(u) + sel;
            if ((L & 3) == 0 && r < n) epi(r, v);
        }
    };
    load_unit4(A, rs, ldj, base(0));
    load_unit4(B, rs, ldj, base(1));
    int u = 0;
    for (; u + 1 < U; u += 2) {
        step(A, u);
        step(B, u + 1);
    }
    if (u < U) step(A, u);
}

// Blocks per instance and sub-blocks per block for the row-stream kernels.
__host__ __device__ inline void stream_shape(int n, int B, int& blocks, int& nsb) {
    const int nsub = (n + ROWS - 1) / ROWS;
    const int want = (STREAM_BLOCKS + B - 1) / B;
    const int g = want < 1 ? 1 : (want > nsub ? nsub : want);
    nsb = (nsub + g - 1) / g;
    blocks = (nsub + nsb - 1) / nsb;
}

template <bool COUNTED>
__global__ void __launch_bounds__(THREADS) mf_fwd_stream(MFArgs a, int nsb, int t, const float* __restrict__ m_prev,
                                                         const float* __restrict__ m_cur, float* __restrict__ m_next,
                                                         float* __restrict__ u_tape, float* __restrict__ beta_tape) {
    extern __shared__ float xs[];
    __shared__ float red[4];
    const int b = blockIdx.y, n = a.n;
    const int np16 = 16 * ((n + 15) / 16);
    const int row_base = blockIdx.x * ROWS * nsb;
    float* mp = xs + np16;                         // m_prev of the block's rows
    const float* mc = m_cur + size_t(b) * n;
    const size_t off = size_t(b) * n;
    float nsoft = 0.f;
    batched_walk<1>(np16, [&](int i, float* v) { v[0] = mc[min(i, n - 1)]; }, [&](int i, const float* vv) {
        const float v = i < n ? vv[0] : 0.f;
        xs[i] = v;
        nsoft += i < n ? 1.f - v * v : 0.f;
    });
    for (int i = threadIdx.x; i < ROWS * nsb; i += THREADS) mp[i] = row_base + i < n ? m_prev[off + row_base + i] : 0.f;
    __syncthreads();
    float beta = a.coef[t];
    if (COUNTED) beta *= block_sum(nsoft, red);
    if (t == 0) beta = 0.f;
    const float inv2T = 0.5f / a.T[t];
    const float q = a.q;
    stream_rows4(a.J + size_t(b) * size_t(a.bstride), a.ldj, n, row_base, nsb, xs, [&](int r, float hv) {
        const float u = hv + q * xs[r] - beta * mp[r - row_base];
        m_next[off + r] = tanhf(u * inv2T);
        if (u_tape) u_tape[off + r] = u;
    });
    if (beta_tape && blockIdx.x == 0 && threadIdx.x == 0) beta_tape[b] = beta;
}

template <bool COUNTED>
__global__ void __launch_bounds__(THREADS) mf_bwd_stream(MFArgs a, int nsb, int t, const float* __restrict__ g_next,
                                                         const float* __restrict__ m_next, const float* __restrict__ m_cur,
                                                         const float* __restrict__ m_prev, const float* __restrict__ u_t,
                                                         const float* __restrict__ delta_next, const float* __restrict__ beta_next,
                                                         float* __restrict__ delta_out, float* __restrict__ g_out,
                                                         double* __restrict__ dT, double* __restrict__ dcoef) {
    extern __shared__ float xs[];
    __shared__ double red[4];
    const int b = blockIdx.y, n = a.n;
    const size_t off = size_t(b) * n;
    const int np16 = 16 * ((n + 15) / 16);
    const int row_base = blockIdx.x * ROWS * nsb;
    float* mcr = xs + np16;                        // m_cur of the block's rows
    float* dnr = mcr + ROWS * nsb;                 // delta_next of the block's rows
    const float inv2T = 0.5f / a.T[t];
    double s1 = 0.0, s2 = 0.0, nsoft = 0.0;
    for (int i = n + threadIdx.x; i < np16; i += THREADS) xs[i] = 0.f;
    batched_walk<5>(n, [&](int i, float* v) {
        v[0] = m_next[off + i]; v[1] = g_next[off + i]; v[2] = u_t[off + i]; v[3] = m_prev[off + i];
        v[4] = COUNTED ? m_cur[off + i] : 0.f;
    }, [&](int i, const float* v) {
        const float mn = v[0];
        const float d = v[1] * (1.f - mn * mn) * inv2T;
        xs[i] = d;
        s1 += double(d) * double(v[2]);
        s2 += double(d) * double(v[3]);
        if (COUNTED) { const double mc = v[4]; nsoft += 1.0 - mc * mc; }
    });
    for (int i = threadIdx.x; i < ROWS * nsb; i += THREADS) {
        const bool in = row_base + i < n;
        mcr[i] = COUNTED && in ? m_cur[off + row_base + i] : 0.f;
        dnr[i] = delta_next && in ? delta_next[off + row_base + i] : 0.f;
    }
    __syncthreads();
    s1 = block_sum_d(s1, red);
    s2 = block_sum_d(s2, red);
    if (COUNTED) nsoft = block_sum_d(nsoft, red);
    if (t == 0) s2 = 0.0;   // beta_0 = 0: no dependence on m_{-1} or coef_0
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        dT[size_t(t) * a.B + b] = -s1 / double(a.T[t]);
        dcoef[size_t(t) * a.B + b] = COUNTED ? -s2 * nsoft : -s2;
    }
    const float q = a.q;
    const float cm = COUNTED ? float(2.0 * double(a.coef[t]) * s2) : 0.f;
    const float bn = delta_next ? beta_next[b] : 0.f;
    const bool has_next = delta_next != nullptr;
    stream_rows4(a.J + size_t(b) * size_t(a.bstride), a.ldj, n, row_base, nsb, xs, [&](int r, float v) {
        float g = v + q * xs[r];
        if (COUNTED) g += cm * mcr[r - row_base];
        if (has_next) g -= bn * dnr[r - row_base];
        g_out[off + r] = g;
        delta_out[off + r] = xs[r];
    });
}

__global__ void __launch_bounds__(THREADS) mf_gemv_stream(MFArgs a, int nsb, const float* __restrict__ x, float* __restrict__ y) {
    extern __shared__ float xs[];
    const int b = blockIdx.y, n = a.n;
    stage_padded(xs, n, x + size_t(b) * n);
    __syncthreads();
    stream_rows4(a.J + size_t(b) * size_t(a.bstride), a.ldj, n, blockIdx.x * ROWS * nsb, nsb, xs,
                 [&](int r, float h) { y[size_t(b) * n + r] = h; });
}

}  // namespace mf

static dim3 mf_grid(const MFArgs& a) { return dim3((a.n + mf::ROWS - 1) / mf::ROWS, a.B); }
static size_t mf_lds(const MFArgs& a) { return sizeof(float) * 16 * size_t((a.n + 15) / 16); }

// Launch shape of the row-stream kernels (packed storage).
struct StreamLaunch {
    dim3 grid;
    int nsb;
    size_t lds;   // x (padded) + two row vectors of the block's rows
    explicit StreamLaunch(const MFArgs& a) {
        int blocks;
        mf::stream_shape(a.n, a.B, blocks, nsb);
        grid = dim3(blocks, a.B);
        lds = sizeof(float) * (16 * size_t((a.n + 15) / 16) + 2 * size_t(mf::ROWS) * nsb);
    }
};

int launch_mf_forward(const MFArgs& a, const float* m_init_prev, const float* m_init, int t0, int t1, float* m_buf,
                      float* u_tape, float* beta_tape, hipStream_t st) {
    // m_buf holds (t1 - t0 + 1) vectors of B x n: slot 0 = m_{t0}, slot k = m_{t0+k}; m_init_prev is m_{t0-1}.
    const size_t vec = size_t(a.B) * a.n;
    if (m_buf != m_init) {
        if (hipMemcpyAsync(m_buf, m_init, vec * sizeof(float), hipMemcpyDeviceToDevice, st) != hipSuccess) return -1;
    }
    const StreamLaunch sl(a);
    for (int t = t0; t < t1; ++t) {
        const int k = t - t0;
        const float* mp = k == 0 ? m_init_prev : m_buf + (k - 1) * vec;
        const float* mc = m_buf + k * vec;
        float* mn = m_buf + (k + 1) * vec;
        float* ut = u_tape ? u_tape + k * vec : nullptr;
        float* bt = beta_tape ? beta_tape + size_t(k) * a.B : nullptr;
        if (a.jbits == 4) {
            if (a.counted) hipLaunchKernelGGL(mf::mf_fwd_stream<true>, sl.grid, dim3(mf::THREADS), sl.lds, st, a, sl.nsb, t, mp, mc, mn, ut, bt);
            else hipLaunchKernelGGL(mf::mf_fwd_stream<false>, sl.grid, dim3(mf::THREADS), sl.lds, st, a, sl.nsb, t, mp, mc, mn, ut, bt);
        } else {
            if (a.counted) hipLaunchKernelGGL(mf::mf_fwd_step<true>, mf_grid(a), dim3(mf::THREADS), mf_lds(a), st, a, t, mp, mc, mn, ut, bt);
            else hipLaunchKernelGGL(mf::mf_fwd_step<false>, mf_grid(a), dim3(mf::THREADS), mf_lds(a), st, a, t, mp, mc, mn, ut, bt);
        }
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_mf_bwd_step(const MFArgs& a, int t, const float* g_next, const float* m_next, const float* m_cur,
                       const float* m_prev, const float* u_t, const float* delta_next, const float* beta_next, float* delta_out,
                       float* g_out, double* dT, double* dcoef, hipStream_t st) {
    if (a.jbits == 4) {
        const StreamLaunch sl(a);
        if (a.counted)
            hipLaunchKernelGGL(mf::mf_bwd_stream<true>, sl.grid, dim3(mf::THREADS), sl.lds, st, a, sl.nsb, t, g_next, m_next, m_cur,
                               m_prev, u_t, delta_next, beta_next, delta_out, g_out, dT, dcoef);
        else
            hipLaunchKernelGGL(mf::mf_bwd_stream<false>, sl.grid, dim3(mf::THREADS), sl.lds, st, a, sl.nsb, t, g_next, m_next, m_cur,
                               m_prev, u_t, delta_next, beta_next, delta_out, g_out, dT, dcoef);
    } else if (a.counted) {
        hipLaunchKernelGGL(mf::mf_bwd_step<true>, mf_grid(a), dim3(mf::THREADS), mf_lds(a), st, a, t, g_next, m_next, m_cur, m_prev, u_t,
                           delta_next, beta_next, delta_out, g_out, dT, dcoef);
    } else {
        hipLaunchKernelGGL(mf::mf_bwd_step<false>, mf_grid(a), dim3(mf::THREADS), mf_lds(a), st, a, t, g_next, m_next, m_cur, m_prev, u_t,
                           delta_next, beta_next, delta_out, g_out, dT, dcoef);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_mf_gemv(const MFArgs& a, const float* x, float* y, hipStream_t st) {
    if (a.jbits == 4) {
        const StreamLaunch sl(a);
        hipLaunchKernelGGL(mf::mf_gemv_stream, sl.grid, dim3(mf::THREADS), sl.lds, st, a, sl.nsb, x, y);
    } else {
        hipLaunchKernelGGL(mf::mf_gemv, mf_grid(a), dim3(mf::THREADS), mf_lds(a), st, a, x, y);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_mf_pack4(int B, int n, const int8_t* J, int64_t ldj, int64_t bstride, uint8_t* J4, int64_t ldj4, int64_t bstride4,
                    int* bad, hipStream_t st) {
    hipLaunchKernelGGL(mf::mf_pack4, dim3(n, B), dim3(mf::THREADS), 0, st, J, ldj, bstride, n, J4, ldj4, bstride4, bad);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace anc
