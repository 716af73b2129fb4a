// Differentiable mean-field relaxation of the synchronous SCA dynamics with the Onsager reaction term (build-defined;
// no reference implementation). Per instance b (dense integer J_b, FP32 magnetisations m in [-1, 1]^N):
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
// Kernels (one launch per step; grid = (row groups, instances), 256 threads; each workgroup walks `nsb` consecutive
// 64-row sub-blocks of its instance, each wave 16 rows of each):
//   mf_fwd        m_{t+1} (and the tape u_t, beta_t for the backward recomputation)
//   mf_bwd        adjoint: delta_t = g_{t+1} (1 - m_{t+1}^2) / (2 T_t);
//                 g_t = J delta_t + q delta_t + 2 coef_t S2_t m_t [counted] - beta_{t+1} delta_{t+1}
//                 dL/dT_t (direct) = -(1/T_t) sum_i delta_i u_i ; dL/dcoef_t = -S2_t (n_t in counted mode); these
//                 instance-level sums are accumulated in FP64
//                 with S2_t = sum_i delta_{t,i} m_{t-1,i}, n_t = sum_i (1 - m_{t,i}^2)
//   mf_gemv       y = J x (energies and the loss gradient)
//   mf_pack       packed 4-bit (entries in [-8, 7]) or 2-bit ([-2, 1]) copy of int8 J
//
// The product J v (v = m_t, delta_t or x) is an exact integer product on the int8 matrix cores. v is quantised per
// instance and step as a block floating-point vector: s = 30 - e with max_i |v_i| in [2^(e-1), 2^e), q_i =
// rint(v_i 2^s) (|q_i| <= 2^30; entries within 2^-6 of the maximum are exact), split into four balanced base-256
// digits q = l0 + 2^8 l1 + 2^16 l2 + 2^24 l3 (l_k in [-128, 127]). One v_mfma_i32_16x16x64_i8 multiplies 16 rows x
// 64 columns of J (the A operand) with the four digit vectors, which occupy four of the 16 B columns (the others
// repeat them and are ignored); the int32 sums are exact (|sum| <= 128 * 128 * n < 2^31), and sum_k 2^(8k) C_k
// (< 2^53, int64) is converted to double exactly, scaled by 2^-s exactly and rounded once to FP32. Each row sum is
// therefore the correctly rounded value of J q 2^-s, independent of summation order, workgroup shape and J storage.
//
// J storage: 8 = int8 rows (row stride ldj bytes); 4 = packed: two entries per byte, in each 4-byte group g of a row
// byte e holds column 8g + e in its low nibble and column 8g + 4 + e in its high nibble (two's complement). A packed
// dword w unpacks to the A operand bytes (w << 4) & 0xF0F0F0F0 (columns 8g..8g+3) and w & 0xF0F0F0F0 (8g+4..8g+7),
// i.e. 16 J; the extra 2^4 is taken out of the final scale. 2 = two-bit: in each 4-byte group g of a row (columns
// 16g .. 16g + 15), bits 2k, 2k + 1 of byte e hold column 16g + 4k + e (two's complement, [-2, 1]); a dword d
// unpacks to (d << 6), (d << 4), (d << 2), d, each & 0xC0C0C0C0: 64 J in K order, scale 2^6. All storages give
// identical results.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>

#include "anc_device.hpp"
#include "anc_kernels.hpp"
#include "anc_mf_common.hpp"

namespace anc {
namespace mf {

// ---------------------------------------------------------------------------------------------------------- rows
// Rows row_base .. row_base + 64 nsb - 1 of instance J (storage JB, n rows of ldj bytes) times the quantised vector
// (digits in LDS, shift s): start() issues the first DEPTH units' loads (before the caller stages the vector, so that
// their latency overlaps it), run() calls epi(r, value) once per row r < n.
//
// A wave's work is a sequence of units (sub-block, 256-column block): lane (g, ml) = (L / 16, L % 16) loads the 64
// columns 256 uc + 64 g .. + 63 of row r0 + ml (16, 32 or 64 bytes) -- 16 rows x 64, 128 or 256 contiguous bytes
// per unit -- and issues 4 MFMAs; MFMA j takes columns + 16 j .. + 15 as its A bytes (K = 16 g .. 16 g + 15) and the
// same columns of digit ml & 3 from LDS as its B bytes. DEPTH unit buffers form a ring: unit u + DEPTH is loaded
// right after unit u's MFMAs, with no branch around a load (units past the end point beyond the instance, which the
// buffer resource turns into zeros without a memory access), so the compiler's memory-counter waits stay at
// DEPTH - 1 units outstanding. After a sub-block's last unit, lane (g, 0) of each quad collects the four digit sums
// of rows 4 g .. 4 g + 3 (output element e of lane (g, ml) is C[4 g + e][ml]) by quad broadcasts and combines them.
template <int JB>
struct GemvRows {
    static_assert(JB == 8 || JB == 4 || JB == 2, "J storage");
    static_assert(DEPTH == 4, "unit ring");
    static constexpr int NW = 2 * JB;                    // dwords per lane and unit
    static constexpr int UB = UNIT_COLS * JB / 8;        // bytes per row and unit
    __amdgpu_buffer_rsrc_t rs;
    int ldj, n, row_base, UC, U, end;
    uint32_t b0[NW], b1[NW], b2[NW], b3[NW];

    __device__ __forceinline__ int base(int u) const {
        const int w = threadIdx.x / 64, L = threadIdx.x & 63;
        return u < U ? (row_base + ROWS * (u / UC) + 16 * w + (L & 15)) * ldj + UB * (u % UC) + (UB / 4) * (L >> 4) : end;
    }
    __device__ __forceinline__ void load(uint32_t (&b)[NW], int u) {
        const int o = base(u);
#pragma unroll
        for (int k = 0; k < NW / 4; ++k) {
            const v4i v = __builtin_amdgcn_raw_buffer_load_b128(rs, o + 16 * k, 0, 0);
            b[4 * k] = v[0]; b[4 * k + 1] = v[1]; b[4 * k + 2] = v[2]; b[4 * k + 3] = v[3];
        }
    }
    __device__ __forceinline__ void start(const int8_t* J, int ldj_, int n_, int row_base_, int nsb) {
        ldj = ldj_; n = n_; row_base = row_base_;
        end = n * ldj;
        rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<int8_t*>(J), 0, end, 0x00020000);
        UC = (n + UNIT_COLS - 1) / UNIT_COLS;
        U = nsb * UC;
        load(b0, 0);
        load(b1, 1);
        load(b2, 2);
        load(b3, 3);
    }
    template <class Epi>
    __device__ __forceinline__ void run(const int8_t* xl, int lstride, int s, Epi&& epi) {
        const int w = threadIdx.x / 64, L = threadIdx.x & 63, g = L >> 4, ml = L & 15;
        const int8_t* xb = xl + (ml & 3) * lstride + 64 * g;
        const int shift = s + (JB == 4 ? 4 : JB == 2 ? 6 : 0);   // A holds 16 J (packed) or 64 J (2-bit)
        v4i acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
        auto step = [&](uint32_t (&b)[NW], int u) {
            const int uc = u % UC;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                v4i a;
                if constexpr (JB == 8) {
                    a = v4i{int(b[4 * j]), int(b[4 * j + 1]), int(b[4 * j + 2]), int(b[4 * j + 3])};
                } else if constexpr (JB == 2) {
                    const uint32_t d = b[j];
                    a = v4i{int((d << 6) & 0xC0C0C0C0u), int((d << 4) & 0xC0C0C0C0u), int((d << 2) & 0xC0C0C0C0u), int(d & 0xC0C0C0C0u)};
                } else {
                    const uint32_t d0 = b[2 * j], d1 = b[2 * j + 1];
                    a = v4i{int((d0 << 4) & 0xF0F0F0F0u), int(d0 & 0xF0F0F0F0u), int((d1 << 4) & 0xF0F0F0F0u), int(d1 & 0xF0F0F0F0u)};
                }
                const v4i bx = *reinterpret_cast<const v4i*>(xb + UNIT_COLS * uc + 16 * j);
                if (j & 1) acc1 = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, bx, acc1, 0, 0, 0);
                else acc0 = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, bx, acc0, 0, 0, 0);
            }
            load(b, u + DEPTH);
            if (uc == UC - 1) {
                const int r0 = row_base + ROWS * (u / UC) + 16 * w + 4 * g;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int c = acc0[e] + acc1[e];
                    const int c1 = __builtin_amdgcn_update_dpp(0, c, 0x55, 0xF, 0xF, false);   // quad_perm [1,1,1,1]
                    const int c2 = __builtin_amdgcn_update_dpp(0, c, 0xAA, 0xF, 0xF, false);   // [2,2,2,2]
                    const int c3 = __builtin_amdgcn_update_dpp(0, c, 0xFF, 0xF, 0xF, false);   // [3,3,3,3]
                    if (ml == 0 && r0 + e < n) {
                        const int64_t tot = int64_t(c) + (int64_t(c1) << 8) + (int64_t(c2) << 16) + (int64_t(c3) << 24);
                        epi(r0 + e, float(ldexp(double(tot), -shift)));
                    }
                }
                acc0 = v4i{0, 0, 0, 0};
                acc1 = v4i{0, 0, 0, 0};
            }
        };
        int u = 0;
        for (; u + 3 < U; u += 4) {
            step(b0, u);
            step(b1, u + 1);
            step(b2, u + 2);
            step(b3, u + 3);
        }
        if (u < U) step(b0, u);
        if (u + 1 < U) step(b1, u + 1);
        if (u + 2 < U) step(b2, u + 2);
    }
};

// Row groups per instance and sub-blocks per workgroup.
__host__ __device__ inline void gemv_shape(int n, int B, int& blocks, int& nsb) {
    const int nsub = (n + ROWS - 1) / ROWS;
    const int want = (GEMV_BLOCKS + B - 1) / B;
    const int g = want < 1 ? 1 : (want > nsub ? nsub : want);
    nsb = (nsub + g - 1) / g;
    blocks = (nsub + nsb - 1) / nsb;
}

// LDS of a workgroup: the digit vectors, then two row vectors of ROWS * nsb floats.
__host__ __device__ inline size_t gemv_lds(int n, int nsb) { return 4 * size_t(digit_stride(n)) + 2 * sizeof(float) * ROWS * nsb; }

// ---------------------------------------------------------------------------------------------------------- kernels
// Per-row inputs of the epilogues, loaded for the workgroup's rows before the vector pass and kept in LDS.
template <class Ld>
__device__ __forceinline__ void stage_rows(float* dst0, float* dst1, int nr, int row_base, int n, Ld&& ld) {
    for (int i0 = 0; i0 < nr; i0 += THREADS) {
        const int i = i0 + threadIdx.x;
        if (i < nr) {
            float v0 = 0.f, v1 = 0.f;
            if (row_base + i < n) ld(row_base + i, v0, v1);
            dst0[i] = v0;
            dst1[i] = v1;
        }
    }
}

template <bool COUNTED, int JB>
__global__ void __launch_bounds__(THREADS) mf_fwd(MFArgs a, int nsb, int t, const float* __restrict__ m_prev,
                                                  const float* __restrict__ m_cur, float* __restrict__ m_next,
                                                  float* __restrict__ u_tape, float* __restrict__ beta_tape) {
    extern __shared__ __align__(16) int8_t lds[];
    __shared__ float red[4];
    const int b = blockIdx.y, n = a.n;
    const size_t off = size_t(b) * n;
    const int row_base = blockIdx.x * ROWS * nsb, nr = ROWS * nsb;
    const int lstride = digit_stride(n);
    GemvRows<JB> rows;
    rows.start(a.J + size_t(b) * size_t(a.bstride), int(a.ldj), n, row_base, nsb);
    float* mp = reinterpret_cast<float*>(lds + 4 * lstride);   // m_prev and m_cur of the workgroup's rows
    float* mcr = mp + nr;
    const float* mc = m_cur + off;
    stage_rows(mp, mcr, nr, row_base, n, [&](int r, float& v0, float& v1) { v0 = m_prev[off + r]; v1 = mc[r]; });
    float nsoft = 0.f;
    const int s = stage_vector<1>(lds, lstride, n, red, [&](int i, float* raw) { raw[0] = mc[i]; },
                                  [&](const float* raw) { return raw[0]; },
                                  [&](const float*, float v) { nsoft += 1.f - v * v; });
    float beta = a.coef[t];
    if (COUNTED) beta *= block_sum(nsoft, red);
    if (t == 0) beta = 0.f;
    __syncthreads();
    const float inv2T = 0.5f / a.T[t], q = a.q;
    rows.run(lds, lstride, s, [&](int r, float hv) {
        const float u = hv + q * mcr[r - row_base] - beta * mp[r - row_base];
        m_next[off + r] = tanhf(u * inv2T);
        if (u_tape) u_tape[off + r] = u;
    });
    if (beta_tape && blockIdx.x == 0 && threadIdx.x == 0) beta_tape[b] = beta;
}

template <bool COUNTED, int JB>
__global__ void __launch_bounds__(THREADS) mf_bwd(MFArgs a, int nsb, int t, const float* __restrict__ g_next,
                                                  const float* __restrict__ m_next, const float* __restrict__ m_cur,
                                                  const float* __restrict__ m_prev, const float* __restrict__ u_t,
                                                  const float* __restrict__ delta_next, const float* __restrict__ beta_next,
                                                  float* __restrict__ delta_out, float* __restrict__ g_out,
                                                  double* __restrict__ dT, double* __restrict__ dcoef) {
    extern __shared__ __align__(16) int8_t lds[];
    __shared__ double red[12];
    __shared__ float redf[4];
    const int b = blockIdx.y, n = a.n;
    const size_t off = size_t(b) * n;
    const int row_base = blockIdx.x * ROWS * nsb, nr = ROWS * nsb;
    const int lstride = digit_stride(n);
    GemvRows<JB> rows;
    rows.start(a.J + size_t(b) * size_t(a.bstride), int(a.ldj), n, row_base, nsb);
    float* mcr = reinterpret_cast<float*>(lds + 4 * lstride);   // m_cur and delta_next of the workgroup's rows
    float* dnr = mcr + nr;
    stage_rows(mcr, dnr, nr, row_base, n, [&](int r, float& v0, float& v1) {
        v0 = COUNTED ? m_cur[off + r] : 0.f;
        v1 = delta_next ? delta_next[off + r] : 0.f;
    });
    const float inv2T = 0.5f / a.T[t];
    auto delta = [&](float gn, float mn) { return gn * (1.f - mn * mn) * inv2T; };
    double s1 = 0.0, s2 = 0.0, nsoft = 0.0;
    const int s = stage_vector<5>(lds, lstride, n, redf, [&](int i, float* raw) {
        raw[0] = m_next[off + i]; raw[1] = g_next[off + i]; raw[2] = u_t[off + i]; raw[3] = m_prev[off + i];
        raw[4] = COUNTED ? m_cur[off + i] : 0.f;
    }, [&](const float* raw) { return delta(raw[1], raw[0]); }, [&](const float* raw, float d) {
        s1 += double(d) * double(raw[2]);
        s2 += double(d) * double(raw[3]);
        if (COUNTED) { const double mc = raw[4]; nsoft += 1.0 - mc * mc; }
    });
    block_sum_d3(s1, s2, nsoft, red);
    if (t == 0) s2 = 0.0;   // beta_0 = 0: no dependence on m_{-1} or coef_0
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        dT[size_t(t) * a.B + b] = -s1 / double(a.T[t]);
        dcoef[size_t(t) * a.B + b] = COUNTED ? -s2 * nsoft : -s2;
    }
    __syncthreads();
    const float q = a.q;
    const float cm = COUNTED ? float(2.0 * double(a.coef[t]) * s2) : 0.f;
    const float bn = delta_next ? beta_next[b] : 0.f;
    const bool has_next = delta_next != nullptr;
    rows.run(lds, lstride, s, [&](int r, float v) {
        const float d = delta(g_next[off + r], m_next[off + r]);
        float g = v + q * d;
        if (COUNTED) g += cm * mcr[r - row_base];
        if (has_next) g -= bn * dnr[r - row_base];
        g_out[off + r] = g;
        delta_out[off + r] = d;
    });
}

template <int JB>
__global__ void __launch_bounds__(THREADS) mf_gemv(MFArgs a, int nsb, const float* __restrict__ x, float* __restrict__ y) {
    extern __shared__ __align__(16) int8_t lds[];
    __shared__ float red[4];
    const int b = blockIdx.y, n = a.n;
    const float* xb = x + size_t(b) * n;
    const int lstride = digit_stride(n);
    GemvRows<JB> rows;
    rows.start(a.J + size_t(b) * size_t(a.bstride), int(a.ldj), n, blockIdx.x * ROWS * nsb, nsb);
    const int s = stage_vector<1>(lds, lstride, n, red, [&](int i, float* raw) { raw[0] = xb[i]; },
                                  [&](const float* raw) { return raw[0]; }, [&](const float*, float) {});
    __syncthreads();
    rows.run(lds, lstride, s, [&](int r, float h) { y[size_t(b) * n + r] = h; });
}

// Packed copy of int8 J with BITS = 4 or 2 bits per entry: one thread per 4-byte group (8 or 16 columns) of a row;
// block (row, instance). Columns at or beyond n become 0; entries outside [-8, 7] or [-2, 1] are counted in *bad
// (the copy is then not usable).
template <int BITS>
__global__ void __launch_bounds__(THREADS) mf_pack(const int8_t* __restrict__ J, int64_t ldj, int64_t bstride, int n,
                                                    uint8_t* __restrict__ Jp, int64_t ldp, int64_t bstridep, int* bad) {
    constexpr int PER = 32 / BITS, LO = -(1 << (BITS - 1)), HI = (1 << (BITS - 1)) - 1;
    const int r = blockIdx.x, b = blockIdx.y;
    const int8_t* src = J + size_t(b) * size_t(bstride) + size_t(r) * size_t(ldj);
    uint32_t* dst = reinterpret_cast<uint32_t*>(Jp + size_t(b) * size_t(bstridep) + size_t(r) * size_t(ldp));
    int nbad = 0;
    for (int g = threadIdx.x; g < int(ldp / 4); g += THREADS) {
        uint32_t word = 0;
        for (int c = 0; c < PER; ++c) {          // column PER g + 4 k + e -> byte e, field k
            const int col = PER * g + c;
            const int v = col < n ? int(src[col]) : 0;
            nbad += (v < LO || v > HI) ? 1 : 0;
            const int e = c & 3, k = c >> 2;
            word |= (uint32_t(v) & ((1u << BITS) - 1u)) << (8 * e + BITS * k);
        }
        dst[g] = word;
    }
    if (nbad) atomicAdd(bad, nbad);
}

}  // namespace mf

// Launch shape of the kernels: row groups per instance, sub-blocks per workgroup, LDS bytes.
struct MFLaunch {
    dim3 grid;
    int nsb;
    size_t lds;
    explicit MFLaunch(const MFArgs& a) {
        int blocks;
        mf::gemv_shape(a.n, a.B, blocks, nsb);
        grid = dim3(blocks, a.B);
        lds = mf::gemv_lds(a.n, nsb);
    }
};

// Kernels whose workgroups need more than 64 KiB of dynamic LDS (n above about 16000) must be allowed to.
template <class K>
static void allow_lds(K kernel, size_t lds) {
    static std::atomic<size_t> allowed{65536};
    if (lds > allowed.load()) {
        if (hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)) == hipSuccess)
            allowed.store(lds);
    }
}

template <bool COUNTED, int JB>
static void fwd_launch(const MFArgs& a, const MFLaunch& l, int t, const float* mp, const float* mc, float* mn, float* ut, float* bt,
                       hipStream_t st) {
    allow_lds(mf::mf_fwd<COUNTED, JB>, l.lds);
    hipLaunchKernelGGL((mf::mf_fwd<COUNTED, JB>), l.grid, dim3(mf::THREADS), l.lds, st, a, l.nsb, t, mp, mc, mn, ut, bt);
}

int launch_mf_forward(const MFArgs& a, const float* m_init_prev, const float* m_init, int t0, int t1, float* m_buf,
                      float* u_tape, float* beta_tape, hipStream_t st) {
    // m_buf holds (t1 - t0 + 1) vectors of B x n: slot 0 = m_{t0}, slot k = m_{t0+k}; m_init_prev is m_{t0-1}.
    const size_t vec = size_t(a.B) * a.n;
    if (m_buf != m_init) {
        if (hipMemcpyAsync(m_buf, m_init, vec * sizeof(float), hipMemcpyDeviceToDevice, st) != hipSuccess) return -1;
    }
    const MFLaunch l(a);
    for (int t = t0; t < t1; ++t) {
        const int k = t - t0;
        const float* mp = k == 0 ? m_init_prev : m_buf + (k - 1) * vec;
        const float* mc = m_buf + k * vec;
        float* mn = m_buf + (k + 1) * vec;
        float* ut = u_tape ? u_tape + k * vec : nullptr;
        float* bt = beta_tape ? beta_tape + size_t(k) * a.B : nullptr;
        if (a.jbits == 4) {
            if (a.counted) fwd_launch<true, 4>(a, l, t, mp, mc, mn, ut, bt, st);
            else fwd_launch<false, 4>(a, l, t, mp, mc, mn, ut, bt, st);
        } else if (a.jbits == 2) {
            if (a.counted) fwd_launch<true, 2>(a, l, t, mp, mc, mn, ut, bt, st);
            else fwd_launch<false, 2>(a, l, t, mp, mc, mn, ut, bt, st);
        } else {
            if (a.counted) fwd_launch<true, 8>(a, l, t, mp, mc, mn, ut, bt, st);
            else fwd_launch<false, 8>(a, l, t, mp, mc, mn, ut, bt, st);
        }
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <bool COUNTED, int JB>
static void bwd_launch(const MFArgs& a, int t, const float* g_next, const float* m_next, const float* m_cur, const float* m_prev,
                       const float* u_t, const float* delta_next, const float* beta_next, float* delta_out, float* g_out, double* dT,
                       double* dcoef, hipStream_t st) {
    const MFLaunch l(a);
    allow_lds(mf::mf_bwd<COUNTED, JB>, l.lds);
    hipLaunchKernelGGL((mf::mf_bwd<COUNTED, JB>), l.grid, dim3(mf::THREADS), l.lds, st, a, l.nsb, t, g_next, m_next, m_cur, m_prev,
                       u_t, delta_next, beta_next, delta_out, g_out, dT, dcoef);
}

int launch_mf_bwd_step(const MFArgs& a, int t, const float* g_next, const float* m_next, const float* m_cur,
                       const float* m_prev, const float* u_t, const float* delta_next, const float* beta_next, float* delta_out,
                       float* g_out, double* dT, double* dcoef, hipStream_t st) {
    if (a.jbits == 4) {
        if (a.counted) bwd_launch<true, 4>(a, t, g_next, m_next, m_cur, m_prev, u_t, delta_next, beta_next, delta_out, g_out, dT, dcoef, st);
        else bwd_launch<false, 4>(a, t, g_next, m_next, m_cur, m_prev, u_t, delta_next, beta_next, delta_out, g_out, dT, dcoef, st);
    } else if (a.jbits == 2) {
        if (a.counted) bwd_launch<true, 2>(a, t, g_next, m_next, m_cur, m_prev, u_t, delta_next, beta_next, delta_out, g_out, dT, dcoef, st);
        else bwd_launch<false, 2>(a, t, g_next, m_next, m_cur, m_prev, u_t, delta_next, beta_next, delta_out, g_out, dT, dcoef, st);
    } else {
        if (a.counted) bwd_launch<true, 8>(a, t, g_next, m_next, m_cur, m_prev, u_t, delta_next, beta_next, delta_out, g_out, dT, dcoef, st);
        else bwd_launch<false, 8>(a, t, g_next, m_next, m_cur, m_prev, u_t, delta_next, beta_next, delta_out, g_out, dT, dcoef, st);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_mf_gemv(const MFArgs& a, const float* x, float* y, hipStream_t st) {
    const MFLaunch l(a);
    if (a.jbits == 4) {
        allow_lds(mf::mf_gemv<4>, l.lds);
        hipLaunchKernelGGL(mf::mf_gemv<4>, l.grid, dim3(mf::THREADS), l.lds, st, a, l.nsb, x, y);
    } else if (a.jbits == 2) {
        allow_lds(mf::mf_gemv<2>, l.lds);
        hipLaunchKernelGGL(mf::mf_gemv<2>, l.grid, dim3(mf::THREADS), l.lds, st, a, l.nsb, x, y);
    } else {
        allow_lds(mf::mf_gemv<8>, l.lds);
        hipLaunchKernelGGL(mf::mf_gemv<8>, l.grid, dim3(mf::THREADS), l.lds, st, a, l.nsb, x, y);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_mf_pack(int bits, int B, int n, const int8_t* J, int64_t ldj, int64_t bstride, uint8_t* Jp, int64_t ldp, int64_t bstridep,
                   int* bad, hipStream_t st) {
    if (bits == 4) hipLaunchKernelGGL(mf::mf_pack<4>, dim3(n, B), dim3(mf::THREADS), 0, st, J, ldj, bstride, n, Jp, ldp, bstridep, bad);
    else if (bits == 2) hipLaunchKernelGGL(mf::mf_pack<2>, dim3(n, B), dim3(mf::THREADS), 0, st, J, ldj, bstride, n, Jp, ldp, bstridep, bad);
    else return -1;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace anc
