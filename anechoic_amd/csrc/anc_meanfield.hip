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
// Kernels (one launch per step; grid = (row blocks, instances), 256 threads; every block stages the instance's
// whole input vector in LDS (zero-padded to a multiple of 16), recomputes the instance-level reductions redundantly,
// and produces 64 rows, 16 per wave):
//   mf_fwd_step   m_{t+1} (and the tape u_t, beta_t for the backward recomputation)
//   mf_bwd_step   adjoint: delta_t = g_{t+1} (1 - m_{t+1}^2) / (2 T_t);
//                 g_t = J delta_t + q delta_t + 2 coef_t S2_t m_t [counted] - beta_{t+1} delta_{t+1}
//                 dL/dT_t (direct) = -(1/T_t) sum_i delta_i u_i ; dL/dcoef_t = -S2_t (n_t in counted mode); these
//                 instance-level sums are accumulated in FP64
//                 with S2_t = sum_i delta_{t,i} m_{t-1,i}, n_t = sum_i (1 - m_{t,i}^2)
//   mf_gemv       y = J x (energies and the loss gradient)
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

// Stages x (n entries) in LDS, zero-padded to 16 ceil(n / 16) entries.
__device__ __forceinline__ void stage_padded(float* xs, int n, const float* src) {
    const int np16 = 16 * ((n + 15) / 16);
    for (int i = threadIdx.x; i < np16; i += THREADS) xs[i] = i < n ? src[i] : 0.f;
}

// Dot products of the 16 rows r0..r0+15 of J (int8, row stride ldj) with the zero-padded LDS vector x. Lane L reads
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
    for (int i = threadIdx.x; i < np16; i += THREADS) {
        const float v = i < n ? mc[i] : 0.f;
        xs[i] = v;
        nsoft += i < n ? 1.f - v * v : 0.f;
    }
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
    for (int i = threadIdx.x; i < n; i += THREADS) {
        const float mn = m_next[off + i];
        const float d = g_next[off + i] * (1.f - mn * mn) * inv2T;
        xs[i] = d;
        s1 += double(d) * double(u_t[off + i]);
        s2 += double(d) * double(m_prev[off + i]);
        if (COUNTED) { const double mc = m_cur[off + i]; nsoft += 1.0 - mc * mc; }
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

}  // namespace mf

static dim3 mf_grid(const MFArgs& a) { return dim3((a.n + mf::ROWS - 1) / mf::ROWS, a.B); }

int launch_mf_forward(const MFArgs& a, const float* m_init_prev, const float* m_init, int t0, int t1, float* m_buf,
                      float* u_tape, float* beta_tape, hipStream_t st) {
    // m_buf holds (t1 - t0 + 1) vectors of B x n: slot 0 = m_{t0}, slot k = m_{t0+k}; m_init_prev is m_{t0-1}.
    const size_t vec = size_t(a.B) * a.n;
    const size_t lds = sizeof(float) * 16 * size_t((a.n + 15) / 16);
    if (m_buf != m_init) {
        if (hipMemcpyAsync(m_buf, m_init, vec * sizeof(float), hipMemcpyDeviceToDevice, st) != hipSuccess) return -1;
    }
    for (int t = t0; t < t1; ++t) {
        const int k = t - t0;
        const float* mp = k == 0 ? m_init_prev : m_buf + (k - 1) * vec;
        const float* mc = m_buf + k * vec;
        float* mn = m_buf + (k + 1) * vec;
        float* ut = u_tape ? u_tape + k * vec : nullptr;
        float* bt = beta_tape ? beta_tape + size_t(k) * a.B : nullptr;
        if (a.counted) hipLaunchKernelGGL(mf::mf_fwd_step<true>, mf_grid(a), dim3(mf::THREADS), lds, st, a, t, mp, mc, mn, ut, bt);
        else hipLaunchKernelGGL(mf::mf_fwd_step<false>, mf_grid(a), dim3(mf::THREADS), lds, st, a, t, mp, mc, mn, ut, bt);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_mf_bwd_step(const MFArgs& a, int t, const float* g_next, const float* m_next, const float* m_cur,
                       const float* m_prev, const float* u_t, const float* delta_next, const float* beta_next, float* delta_out,
                       float* g_out, double* dT, double* dcoef, hipStream_t st) {
    const size_t lds = sizeof(float) * 16 * size_t((a.n + 15) / 16);
    if (a.counted)
        hipLaunchKernelGGL(mf::mf_bwd_step<true>, mf_grid(a), dim3(mf::THREADS), lds, st, a, t, g_next, m_next, m_cur, m_prev, u_t,
                           delta_next, beta_next, delta_out, g_out, dT, dcoef);
    else
        hipLaunchKernelGGL(mf::mf_bwd_step<false>, mf_grid(a), dim3(mf::THREADS), lds, st, a, t, g_next, m_next, m_cur, m_prev, u_t,
                           delta_next, beta_next, delta_out, g_out, dT, dcoef);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_mf_gemv(const MFArgs& a, const float* x, float* y, hipStream_t st) {
    hipLaunchKernelGGL(mf::mf_gemv, mf_grid(a), dim3(mf::THREADS), sizeof(float) * 16 * size_t((a.n + 15) / 16), st, a, x, y);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace anc
