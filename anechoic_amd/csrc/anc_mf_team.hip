// Persistent team kernels of the mean-field relaxation for n <= 2048 (build-defined; no reference counterpart): the
// same arithmetic as the one-launch-per-step kernels of anc_meanfield.hip (same quantisation, exact integer product,
// FP32 epilogue in the same order, FP64 instance sums in the same order), so both give identical floats.
//
// Shape. 16 teams of 32 workgroups, two workgroups per CU from two different teams; the 32 members of a team run on
// one XCD (workgroup b runs on XCD b % 8), two teams per XCD. A team takes instances team, team + 16, ... one after
// the other and runs a whole checkpoint segment of each in one launch. Member c owns rows 64 c .. 64 c + 63 of the
// instance (padded to 2048); wave w holds rows 16 w .. 16 w + 15 of them over all 2048 columns as int8 MFMA A
// fragments in registers (128 per lane), loaded once per instance and segment instead of once per
// step. Per step every member needs the whole vector (m_t forward, g_{t+1} backward): the members publish their rows
// into a per-step exchange slot and read the other members' rows from it. The two workgroups on a CU belong to
// different teams, so one waits for its team's exchange while the other runs its product.
//
// Exchange. Slot k of instance b holds 2048 words; row r's value v is stored as bits(v) XOR 0x7FC00001 with a plain
// 4-byte store (a quiet NaN pattern no arithmetic here produces, so a stored word is never 0). The slots are zeroed
// before each launch, so a reader polls a word until it is nonzero (non-temporal loads, served by the XCD's L2 that
// holds the producers' stores); a 4-byte store is single-copy atomic, so a nonzero word is the value. No fences and
// no ordering between different words are needed. Every wait is bounded: after `timeout_ticks` the workgroup stores
// an error code into the status block and sets its abort flag, which ends every other wait.
//
// Product. Lane (g, ml) of a wave: A fragment (K-block kb) = row ml of the wave's 16 rows, columns 64 kb + 16 g ..
// + 15; B = digit ml & 3 of the vector at the same columns (LDS, read four K-blocks ahead). After the 32 MFMAs (two
// accumulator chains) lanes ml = 0..3 of each 16-lane row hold the four digit sums of rows 4 g + e; quad broadcasts
// (DPP) give lane (g, ml < 4) the sums of row 4 g + ml, which then runs the per-row epilogue: 16 of a wave's 64
// lanes, one row each.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "anc_device.hpp"
#include "anc_kernels.hpp"
#include "anc_mf_common.hpp"

namespace anc {
namespace mf {

constexpr int TM = 32;            // members per team
constexpr int TRW = 2048 / TM;    // rows per member
constexpr int TNP = 2048;         // padded instance size of the team kernels
constexpr int TKB = TNP / 64;     // K-blocks of 64 columns
// digit rows 2112 bytes apart: the 16 addresses of one B read (4 digits x 4 lane groups 16 bytes apart) start 64 d +
// 16 g bytes into a 256-byte window, so they fall on different LDS banks
constexpr int TLS = TNP + 64;
constexpr int TLDS = 4 * TLS;
constexpr uint32_t XOR_PAT = 0x7FC00001u;

__device__ __forceinline__ uint32_t enc(float v) { return __float_as_uint(v) ^ XOR_PAT; }
__device__ __forceinline__ float dec(uint32_t w) { return __uint_as_float(w ^ XOR_PAT); }

// A operand bytes of row `row`, columns col .. col + 15 (col a multiple of 16) of one instance: int8 entries as they
// are, packed 4-bit entries times 16 and 2-bit entries times 64 (the same unpacking as the per-step kernels); rows
// >= n and columns beyond the row's storage are 0.
template <int JB>
__device__ __forceinline__ v4i afrag(const int8_t* Jb, int64_t ldj, int n, int row, int col) {
    const v4i z = {0, 0, 0, 0};
    if (row >= n) return z;
    const int8_t* rp = Jb + size_t(row) * size_t(ldj);
    if constexpr (JB == 8) {
        if (col >= ldj) return z;
        return *reinterpret_cast<const v4i*>(rp + col);
    } else if constexpr (JB == 4) {
        if (col / 2 >= ldj) return z;
        const uint32_t d0 = *reinterpret_cast<const uint32_t*>(rp + col / 2);
        const uint32_t d1 = *reinterpret_cast<const uint32_t*>(rp + col / 2 + 4);
        return v4i{int((d0 << 4) & 0xF0F0F0F0u), int(d0 & 0xF0F0F0F0u), int((d1 << 4) & 0xF0F0F0F0u), int(d1 & 0xF0F0F0F0u)};
    } else {
        if (col / 4 >= ldj) return z;
        const uint32_t d = *reinterpret_cast<const uint32_t*>(rp + col / 4);
        return v4i{int((d << 6) & 0xC0C0C0C0u), int((d << 4) & 0xC0C0C0C0u), int((d << 2) & 0xC0C0C0C0u), int(d & 0xC0C0C0C0u)};
    }
}

// The wave's 16 rows of one instance as MFMA A fragments (128 registers per lane, pinned to the VGPRs: with two waves
// per SIMD a wave has 256 registers, and an MFMA whose accumulators the compiler places in the AGPRs would push A
// fragments out to scratch).
template <int JB>
struct TeamRows {
    v4i A[TKB];

    __device__ __forceinline__ void pin() {
#pragma unroll
        for (int kb = 0; kb < TKB; ++kb) __asm__ volatile("" : "+v"(A[kb]));
    }
    __device__ __forceinline__ void load(const int8_t* Jb, int64_t ldj, int n, int c) {
        const int w = threadIdx.x / 64, L = threadIdx.x & 63, g = L >> 4, ml = L & 15;
#pragma unroll
        for (int kb = 0; kb < TKB; ++kb) A[kb] = afrag<JB>(Jb, ldj, n, TRW * c + 16 * w + ml, 64 * kb + 16 * g);
        pin();
    }
    // (J q) 2^-shift of this lane's epilogue row (lanes ml < 8; the value of other lanes is meaningless), rounded once
    // to FP32 exactly as the per-step kernels do. xl: the four digit rows (row stride lstride).
    __device__ __forceinline__ float product(const int8_t* xl, int lstride, int shift) {
        pin();
        const int L = threadIdx.x & 63, g = L >> 4, ml = L & 15;
        const int8_t* xb = xl + (ml & 3) * lstride + 16 * g;
        constexpr int AHEAD = 4;   // B fragments in flight
        v4i bq[AHEAD];
#pragma unroll
        for (int i = 0; i < AHEAD; ++i) bq[i] = *reinterpret_cast<const v4i*>(xb + 64 * i);
        v4i acc[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
#pragma unroll
        for (int h = 0; h < 2; ++h) __asm__ volatile("" : "+v"(acc[h]));   // accumulate in VGPRs
#pragma unroll
        for (int kb = 0; kb < TKB; ++kb) {
            const v4i bx = bq[kb % AHEAD];
            if (kb + AHEAD < TKB) bq[kb % AHEAD] = *reinterpret_cast<const v4i*>(xb + 64 * (kb + AHEAD));
            acc[kb & 1] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[kb], bx, acc[kb & 1], 0, 0, 0);
        }
        // digit d of row 4 g + e sits in element e of quad position d of every quad of the lane's 16-lane row
        int32_t ds[4] = {0, 0, 0, 0};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int x = acc[0][e] + acc[1][e];
            const int d0 = __builtin_amdgcn_update_dpp(0, x, 0x00, 0xF, 0xF, false);   // quad_perm [0,0,0,0]
            const int d1 = __builtin_amdgcn_update_dpp(0, x, 0x55, 0xF, 0xF, false);   // [1,1,1,1]
            const int d2 = __builtin_amdgcn_update_dpp(0, x, 0xAA, 0xF, 0xF, false);   // [2,2,2,2]
            const int d3 = __builtin_amdgcn_update_dpp(0, x, 0xFF, 0xF, 0xF, false);   // [3,3,3,3]
            const bool mine = (ml & 3) == e;
            ds[0] = mine ? d0 : ds[0];
            ds[1] = mine ? d1 : ds[1];
            ds[2] = mine ? d2 : ds[2];
            ds[3] = mine ? d3 : ds[3];
        }
        const int64_t tot = int64_t(ds[0]) + (int64_t(ds[1]) << 8) + (int64_t(ds[2]) << 16) + (int64_t(ds[3]) << 24);
        return float(ldexp(double(tot), -shift));
    }
};

// Team placement and the per-lane epilogue row.
struct TeamPos {
    int team, c, row;   // row of this lane's epilogue within the instance (lanes ml < 4)
    bool epi;
    __device__ TeamPos() {
        const int xcd = blockIdx.x % 8, idx = blockIdx.x / 8;
        team = (idx / TM) * 8 + xcd;
        c = idx % TM;
        const int w = threadIdx.x / 64, L = threadIdx.x & 63, g = L >> 4, ml = L & 15;
        epi = ml < 4;
        row = TRW * c + 16 * w + 4 * g + ml;
    }
};

// Polls this thread's two 16-byte chunks of exchange slot `slot` (columns 4 (tid + 256 j) .. + 3, j = 0, 1) until every
// word of a column < n is nonzero; returns false when the wait gave up (time-out or another workgroup's abort).
__device__ __forceinline__ bool poll_slot(const uint32_t* slot, int n, const MFTeamArgs& x, uint32_t (&wv)[RG][4]) {
    __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t*>(slot), 0, TNP * 4, 0x00020000);
    bool ok[RG];
    auto load = [&](int j) {
        const v4i v = __builtin_amdgcn_raw_buffer_load_b128(rs, 16 * (int(threadIdx.x) + THREADS * j), 0, 2);   // nt
#pragma unroll
        for (int e = 0; e < 4; ++e) wv[j][e] = uint32_t(v[e]);
    };
    auto current = [&](int j) {
        bool r = true;
#pragma unroll
        for (int e = 0; e < 4; ++e) r = r && (4 * (int(threadIdx.x) + THREADS * j) + e >= n || wv[j][e] != 0u);
        return r;
    };
#pragma unroll
    for (int j = 0; j < RG; ++j) load(j);
#pragma unroll
    for (int j = 0; j < RG; ++j) ok[j] = current(j);
    int polls = 0;
    uint64_t t_wait = 0;
    while (!__all(ok[0] && ok[1])) {
        if ((++polls & 63) == 0) {
            const uint64_t now = wall_clock64();
            if (polls == 64) t_wait = now;
            if (dev::aborted(x.status)) return false;
            if (now - t_wait > uint64_t(x.timeout_ticks)) {
                dev::raise_error(x.status, 1);
                return false;
            }
        }
        __builtin_amdgcn_s_sleep(1);
        __asm__ volatile("" ::: "memory");   // the re-poll loads are issued anew
#pragma unroll
        for (int j = 0; j < RG; ++j)
            if (!ok[j]) {
                load(j);
                ok[j] = current(j);
            }
    }
    return true;
}

// This thread's columns 4 (tid + 256 j) + e of one instance vector (index clamped to n - 1: the values of columns >= n
// are never used). No branches, so that loads issued ahead of their use stay in flight (the compiler waits for every
// outstanding load where the paths of a branch join).
__device__ __forceinline__ void load_cols(const float* base, int n, float (&d)[RG][4]) {
#pragma unroll
    for (int j = 0; j < RG; ++j)
#pragma unroll
        for (int e = 0; e < 4; ++e) d[j][e] = base[min(4 * (int(threadIdx.x) + THREADS * j) + e, n - 1)];
}

// Host-forced time-out (option "stall_test"): the bounded wait on status word 1, which nothing sets, gives up and
// records the error; the workgroup then exits.
__device__ __forceinline__ void stall(const MFTeamArgs& x) {
    const uint64_t t0 = wall_clock64();
    while (__hip_atomic_load(x.status + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
        if (wall_clock64() - t0 > uint64_t(x.timeout_ticks)) {
            dev::raise_error(x.status, 1);
            break;
        }
        __builtin_amdgcn_s_sleep(8);
    }
}

__device__ __forceinline__ void put(const MFTeamArgs& x, int slot, int B, int b, int row, float v) {
    x.xch[(size_t(slot) * B + b) * TNP + row] = enc(v);
}

// Forward over steps t0 .. t1 - 1 (one checkpoint segment): m_buf slot k + 1 = m_{t0+k+1} (slot 0 = m_{t0} is the
// caller's), u_tape slot k = u_{t0+k}, beta_tape[k B + b] = beta_{t0+k} (tapes optional). Exchange slots 1 .. t1 - t0 - 1
// must be zero.
template <bool COUNTED, int JB>
__global__ void __launch_bounds__(THREADS, 2) mf_team_fwd(MFArgs a, MFTeamArgs x, int t0, int t1, const float* __restrict__ m_init_prev,
                                                           const float* __restrict__ m_init, float* __restrict__ m_buf,
                                                           float* __restrict__ u_tape, float* __restrict__ beta_tape) {
    __shared__ __align__(16) int8_t xl[TLDS];
    __shared__ float red[4];
    __shared__ int bar_cell;
    const TeamPos P;
    const int n = a.n, B = a.B, len = t1 - t0;
    const int lstride = TLS;
    const size_t vec = size_t(B) * size_t(n);
    const int shift_extra = JB == 4 ? 4 : JB == 2 ? 6 : 0;
    const bool mine = P.epi && P.row < n;
    if (threadIdx.x == 0) bar_cell = 0;
    if (x.stall_test) {
        stall(x);
        return;
    }
    TeamRows<JB> rows;
    for (int b = P.team; b < B; b += 16) {
        __syncthreads();
        rows.load(a.J + size_t(b) * size_t(a.bstride), a.ldj, n, P.c);
        const size_t off = size_t(b) * n;
        float mp = 0.f, mc = 0.f;
        if (mine) {
            mp = m_init_prev[off + P.row];
            mc = m_init[off + P.row];
        }
        for (int k = 0; k < len; ++k) {
            const int t = t0 + k;
            // the step's schedule values, loaded before the exchange wait so that their latency is hidden by it
            const float coef_t = a.coef[t], T_t = a.T[t];
            float raw[RG][4][1];
            bool bail = false;
            if (k == 0) {
                float mv[RG][4];
                load_cols(m_init + off, n, mv);
#pragma unroll
                for (int j = 0; j < RG; ++j)
#pragma unroll
                    for (int e = 0; e < 4; ++e) raw[j][e][0] = mv[j][e];
            } else {
                uint32_t wv[RG][4];
                bail = !poll_slot(x.xch + (size_t(k) * B + b) * TNP, n, x, wv);
#pragma unroll
                for (int j = 0; j < RG; ++j)
#pragma unroll
                    for (int e = 0; e < 4; ++e) raw[j][e][0] = dec(wv[j][e]);
            }
            if (dev::barrier_any(bail, bar_cell)) return;
            float nsoft = 0.f;
            const int s = vector_regs<1>(raw, xl, lstride, n, red, [&](const float* r) { return r[0]; },
                                         [&](const float*, float v) { nsoft += 1.f - v * v; });
            float beta = coef_t;
            if (COUNTED) beta *= block_sum(nsoft, red);
            if (t == 0) beta = 0.f;
            __syncthreads();
            const float inv2T = 0.5f / T_t, q = a.q;
            const float hv = rows.product(xl, lstride, s + shift_extra);
            if (mine) {
                const float u = hv + q * mc - beta * mp;
                const float mn = tanhf(u * inv2T);
                if (k + 1 < len) put(x, k + 1, B, b, P.row, mn);
                m_buf[size_t(k + 1) * vec + off + P.row] = mn;
                if (u_tape) u_tape[size_t(k) * vec + off + P.row] = u;
                mp = mc;
                mc = mn;
            }
            if (beta_tape && P.c == 0 && threadIdx.x == 0) beta_tape[size_t(k) * B + b] = beta;
        }
    }
}

// Adjoint over steps t1 - 1 down to t0 of one segment, from the recomputed tapes: seg slot k = m_{t0+k} (k = 0 .. len),
// utape slot k = u_{t0+k}, btape[k B + b] = beta_{t0+k}, cp = m_{t0-1}. g_in = g_{t1} (whole vectors); delta_in =
// delta_{t1} and beta_in = beta_{t1} (NULL at the last step of the run). Writes g_{t0} into g_out and delta_{t0} into
// delta_out, and dT, dcoef of every step. Exchange slots 1 .. len - 1 must be zero. n must be a multiple of 4.
//
// Tapes. Step t reads m_{t+1}, m_t, m_{t-1} and u_t. They go to LDS by direct-to-LDS buffer copies issued at the end
// of the step before: a ring of four m vectors (m_i in ring slot i mod 4) and two u vectors (u_i in slot i mod 2). Each thread
// copies the 16-byte chunks of the columns it scans (4 (tid + 256 j) .. + 3), so after its own copies have landed
// (waiting for the exchange loads issued after them waits for them too: the memory counter retires in order) it can
// scan them; the epilogue reads other threads' rows only after the step's first workgroup barrier, which every wave
// reaches after its wait. The copies made at the end of step t go into the slots of m_{t-2} and u_{t-1}, last read in
// step t + 1 (m_{t+2} as its m_{t+1}, u_{t+1}) before this step's barrier.
template <bool COUNTED, int JB>
__global__ void __launch_bounds__(THREADS, 2) mf_team_bwd(MFArgs a, MFTeamArgs x, int t0, int t1, const float* __restrict__ seg,
                                                           const float* __restrict__ utape, const float* __restrict__ btape,
                                                           const float* __restrict__ cp, const float* __restrict__ g_in,
                                                           const float* __restrict__ delta_in, const float* __restrict__ beta_in,
                                                           float* __restrict__ g_out, float* __restrict__ delta_out,
                                                           double* __restrict__ dT, double* __restrict__ dcoef) {
    __shared__ __align__(16) int8_t xl[TLDS];
    __shared__ __align__(16) float gsh[TNP];
    __shared__ __align__(16) float mring[4][TNP];
    __shared__ __align__(16) float uring[2][TNP];
    __shared__ double red[16];
    __shared__ int bar_cell;
    const TeamPos P;
    const int n = a.n, B = a.B, len = t1 - t0;
    const int lstride = TLS;
    const size_t vec = size_t(B) * size_t(n);
    const int shift_extra = JB == 4 ? 4 : JB == 2 ? 6 : 0;
    const bool mine = P.epi && P.row < n;
    const int tid = threadIdx.x, w = tid / 64;
    if (tid == 0) bar_cell = 0;
    if (x.stall_test) {
        stall(x);
        return;
    }
    // this thread's two 16-byte chunks of instance vector `src` into LDS vector `dst` (chunks beyond n read as 0)
    auto copy_vec = [&](const float* src, float* dst) {
        __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(src), 0, 4 * n, 0x00020000);
#pragma unroll
        for (int j = 0; j < RG; ++j)
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (__attribute__((address_space(3))) void*)(dst + 256 * w + 1024 * j), 16,
                                                     16 * (tid + THREADS * j), 0, 0, 0);
    };
    auto mslot = [&](int i) { return &mring[(i + 4) & 3][0]; };   // m_{t0+i}, i >= -1
    TeamRows<JB> rows;
    for (int b = P.team; b < B; b += 16) {
        __syncthreads();   // the previous instance's LDS reads are complete
        const size_t off = size_t(b) * n;
        copy_vec(seg + size_t(len) * vec + off, mslot(len));
        copy_vec(seg + size_t(len - 1) * vec + off, mslot(len - 1));
        copy_vec(len >= 2 ? seg + size_t(len - 2) * vec + off : cp + off, mslot(len - 2));
        copy_vec(utape + size_t(len - 1) * vec + off, &uring[(len - 1) & 1][0]);
        rows.load(a.J + size_t(b) * size_t(a.bstride), a.ldj, n, P.c);
        float dn = 0.f;   // delta_{t+1} of this lane's row
        if (mine && delta_in) dn = delta_in[off + P.row];
        for (int kk = len - 1; kk >= 0; --kk) {
            const int t = t0 + kk;
            const float T_t = a.T[t];   // loaded before the exchange wait, which hides its latency
            const bool has_next = kk + 1 < len || delta_in != nullptr;
            const float bn = !has_next ? 0.f : (kk + 1 < len ? btape[size_t(kk + 1) * B + b] : beta_in[b]);
            const float* mnx = mslot(kk + 1);
            const float* mcu = mslot(kk);
            const float* mpr = mslot(kk - 1);
            const float* ut = &uring[kk & 1][0];
            // g_{t+1} of this thread's columns
            float gv[RG][4];
            bool bail = false;
            if (kk + 1 == len) {
#pragma unroll
                for (int j = 0; j < RG; ++j)
#pragma unroll
                    for (int e = 0; e < 4; ++e) gv[j][e] = g_in[off + min(4 * (tid + THREADS * j) + e, n - 1)];
            } else {
                uint32_t wv[RG][4];
                bail = !poll_slot(x.xch + (size_t(kk + 1) * B + b) * TNP, n, x, wv);
#pragma unroll
                for (int j = 0; j < RG; ++j)
#pragma unroll
                    for (int e = 0; e < 4; ++e) gv[j][e] = dec(wv[j][e]);
            }
            __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0): this thread's tape copies have landed too
            // the rows' g_{t+1} for the epilogue
#pragma unroll
            for (int j = 0; j < RG; ++j)
                *reinterpret_cast<float4*>(&gsh[4 * (tid + THREADS * j)]) = make_float4(gv[j][0], gv[j][1], gv[j][2], gv[j][3]);
            if (dev::barrier_any(bail, bar_cell)) return;
            const float inv2T = 0.5f / T_t;
            auto delta = [&](float gn, float mn) { return gn * (1.f - mn * mn) * inv2T; };
            // FP64 instance sums s1 = sum d u, s2 = sum d m_{t-1} (and nsoft): with the counted correction every row's
            // epilogue needs s2, so every member forms them over the whole vector (block_max_sum_d3). Otherwise they
            // only feed dT and dcoef, and only wave sw of member c = 8 sw + r forms anything: its lanes r, r + 8, ...,
            // r + 56, after the xor-32/16/8 butterfly levels, are one subtree of block_max_sum_d3's additions over the
            // vector pass's layout; it is stored after the step's publication, off the exchange's critical path, and
            // mf_team_sums adds the 32 subtrees in the order of the remaining levels (the same floats).
            const bool sub = !COUNTED && w == P.c / 8;
            double s1 = 0.0, s2 = 0.0, nsoft = 0.0;
            float tnx[RG][4], tcu[RG][4], tpr[RG][4], tu[RG][4];
#pragma unroll
            for (int j = 0; j < RG; ++j) {
                const int i0 = 4 * (tid + THREADS * j);
                const float4 a0 = *reinterpret_cast<const float4*>(mnx + i0), a1 = *reinterpret_cast<const float4*>(mpr + i0);
                const float4 a2 = *reinterpret_cast<const float4*>(ut + i0);
                tnx[j][0] = a0.x; tnx[j][1] = a0.y; tnx[j][2] = a0.z; tnx[j][3] = a0.w;
                tpr[j][0] = a1.x; tpr[j][1] = a1.y; tpr[j][2] = a1.z; tpr[j][3] = a1.w;
                tu[j][0] = a2.x; tu[j][1] = a2.y; tu[j][2] = a2.z; tu[j][3] = a2.w;
                if (COUNTED) {
                    const float4 a3 = *reinterpret_cast<const float4*>(mcu + i0);
                    tcu[j][0] = a3.x; tcu[j][1] = a3.y; tcu[j][2] = a3.z; tcu[j][3] = a3.w;
                }
            }
            float v[RG][4];
            float mx = vector_scan(n, v, [&](int j, int c) { return delta(gv[j][c], tnx[j][c]); },
                                   [&](int j, int c, float d) {
                                       if (COUNTED || sub) {
                                           s1 += double(d) * double(tu[j][c]);
                                           s2 += double(d) * double(tpr[j][c]);
                                       }
                                       if (COUNTED) { const double mcv = tcu[j][c]; nsoft += 1.0 - mcv * mcv; }
                                   });
            if constexpr (COUNTED) block_max_sum_d3(mx, s1, s2, nsoft, red);
            else mx = block_max(mx, reinterpret_cast<float*>(red));
            const int s = block_shift(mx);
            vector_write(v, xl, lstride, n, s);
            if constexpr (COUNTED) {
                if (t == 0) s2 = 0.0;   // beta_0 = 0: no dependence on m_{-1} or coef_0
                if (P.c == 0 && tid == 0) {
                    dT[size_t(t) * B + b] = -s1 / double(a.T[t]);   // (T_t here would cost the counted kernels a spill)
                    dcoef[size_t(t) * B + b] = -s2 * nsoft;
                }
            }
            __syncthreads();
            const float q = a.q;
            const float cm = COUNTED ? float(2.0 * double(a.coef[t]) * s2) : 0.f;
            const float vr = rows.product(xl, lstride, s + shift_extra);
            if (mine) {
                const float d = delta(gsh[P.row], mnx[P.row]);
                float g = vr + q * d;
                if (COUNTED) g += cm * mcu[P.row];
                if (has_next) g -= bn * dn;
                if (kk > 0) {
                    put(x, kk, B, b, P.row, g);
                } else {
                    g_out[off + P.row] = g;
                    delta_out[off + P.row] = d;
                }
                dn = d;
            }
            if (sub) {
                s1 = wave_sum_d_hi(s1);
                s2 = wave_sum_d_hi(s2);
                if ((tid & 63) == (P.c & 7))
                    reinterpret_cast<double2*>(x.part)[(size_t(kk) * B + b) * TM + P.c] = make_double2(s1, s2);
            }
            // the next step's two new tapes, into the slots read one step ago (m_{t+2}, u_{t+1}): issued after this
            // step's last LDS read, since the compiler cannot tell the ring slots apart and makes every LDS access
            // after a copy wait for it; they land during the next step's exchange wait (its vmcnt(0) and barrier)
            if (kk > 0) {
                copy_vec(kk >= 2 ? seg + size_t(kk - 2) * vec + off : cp + off, mslot(kk - 2));
                copy_vec(utape + size_t(kk - 1) * vec + off, &uring[(kk - 1) & 1][0]);
            }
        }
    }
    __builtin_amdgcn_s_waitcnt(0);   // no LDS copy outlives the workgroup
}

// dT and dcoef of the segment's steps t0 .. t0 + len - 1 (uncounted correction) from the 32 subtree sums per (step,
// instance) that mf_team_bwd stored: within member group 8 w .. 8 w + 7 the xor-4/2/1 butterfly levels, then the
// four groups left to right, as block_max_sum_d3 adds them.
__global__ void mf_team_sums(const double2* __restrict__ part, int t0, int len, int B, const float* __restrict__ T,
                             double* __restrict__ dT, double* __restrict__ dcoef) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len * B) return;
    const int t = t0 + i / B, b = i % B;
    const double2* p = part + size_t(i) * TM;
    double s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const double2* q = p + 8 * g;
        const double w1 = ((q[0].x + q[4].x) + (q[2].x + q[6].x)) + ((q[1].x + q[5].x) + (q[3].x + q[7].x));
        const double w2 = ((q[0].y + q[4].y) + (q[2].y + q[6].y)) + ((q[1].y + q[5].y) + (q[3].y + q[7].y));
        s1 = g == 0 ? w1 : s1 + w1;
        s2 = g == 0 ? w2 : s2 + w2;
    }
    if (t == 0) s2 = 0.0;   // beta_0 = 0: no dependence on m_{-1} or coef_0
    dT[size_t(t) * B + b] = -s1 / double(T[t]);
    dcoef[size_t(t) * B + b] = -s2;
}

}  // namespace mf

// Two workgroups per CU: a dynamic-LDS pad above a third of the CU's 160 KiB keeps a third one off (512 workgroups
// on 256 CUs, all co-resident).
template <class K>
static int team_pad(K kern, size_t& pad) {
    hipFuncAttributes fa;
    if (hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(kern)) != hipSuccess) return -1;
    const size_t third = 160 * 1024 / 3 + 1024;
    pad = fa.sharedSizeBytes < third ? third - fa.sharedSizeBytes : 0;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kern), mf::THREADS, pad) != hipSuccess) return -1;
    return per_cu >= 2 ? 0 : -2;
}

bool mf_team_fits(int n, int cus) { return n <= mf::TNP && n % 4 == 0 && cus >= 256; }

size_t mf_team_slot_words(int B) { return size_t(B) * mf::TNP; }

size_t mf_team_part_doubles(int B) { return size_t(B) * 2 * mf::TM; }

template <bool COUNTED, int JB>
static int team_fwd_t(const MFArgs& a, const MFTeamArgs& x, int t0, int t1, const float* mp, const float* mc, float* m_buf,
                      float* ut, float* bt, hipStream_t st) {
    const auto kern = mf::mf_team_fwd<COUNTED, JB>;
    static size_t pad = 0;
    static int rc = 1;
    if (rc == 1) rc = team_pad(kern, pad);
    if (rc) return rc;
    hipLaunchKernelGGL(kern, dim3(512), dim3(mf::THREADS), pad, st, a, x, t0, t1, mp, mc, m_buf, ut, bt);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <bool COUNTED, int JB>
static int team_bwd_t(const MFArgs& a, const MFTeamArgs& x, int t0, int t1, const float* seg, const float* utape, const float* btape,
                      const float* cp, const float* g_in, const float* delta_in, const float* beta_in, float* g_out, float* delta_out,
                      double* dT, double* dcoef, hipStream_t st) {
    const auto kern = mf::mf_team_bwd<COUNTED, JB>;
    static size_t pad = 0;
    static int rc = 1;
    if (rc == 1) rc = team_pad(kern, pad);
    if (rc) return rc;
    hipLaunchKernelGGL(kern, dim3(512), dim3(mf::THREADS), pad, st, a, x, t0, t1, seg, utape, btape, cp, g_in, delta_in, beta_in,
                       g_out, delta_out, dT, dcoef);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_mf_team_forward(const MFArgs& a, const MFTeamArgs& x, const float* m_init_prev, const float* m_init, int t0, int t1,
                           float* m_buf, float* u_tape, float* beta_tape, hipStream_t st) {
    const size_t vec = size_t(a.B) * a.n;
    if (m_buf != m_init && hipMemcpyAsync(m_buf, m_init, vec * sizeof(float), hipMemcpyDeviceToDevice, st) != hipSuccess) return -1;
    FillRanges f{};
    f.add(x.status, 1);   // the abort flag (the error word is kept)
    if (t1 - t0 > 1) f.add(x.xch + mf_team_slot_words(a.B), int64_t(size_t(t1 - t0 - 1) * mf_team_slot_words(a.B) / 2));
    if (launch_fill_zero_ranges(f, st)) return -1;
#define ANC_MF_FWD(C, J) team_fwd_t<C, J>(a, x, t0, t1, m_init_prev, m_init, m_buf, u_tape, beta_tape, st)
    if (a.jbits == 4) return a.counted ? ANC_MF_FWD(true, 4) : ANC_MF_FWD(false, 4);
    if (a.jbits == 2) return a.counted ? ANC_MF_FWD(true, 2) : ANC_MF_FWD(false, 2);
    return a.counted ? ANC_MF_FWD(true, 8) : ANC_MF_FWD(false, 8);
#undef ANC_MF_FWD
}

int launch_mf_team_backward(const MFArgs& a, const MFTeamArgs& x, int t0, int t1, const float* seg, const float* utape,
                            const float* btape, const float* cp, const float* g_in, const float* delta_in, const float* beta_in,
                            float* g_out, float* delta_out, double* dT, double* dcoef, hipStream_t st) {
    FillRanges f{};
    f.add(x.status, 1);
    if (t1 - t0 > 1) f.add(x.xch + mf_team_slot_words(a.B), int64_t(size_t(t1 - t0 - 1) * mf_team_slot_words(a.B) / 2));
    if (launch_fill_zero_ranges(f, st)) return -1;
#define ANC_MF_BWD(C, J) team_bwd_t<C, J>(a, x, t0, t1, seg, utape, btape, cp, g_in, delta_in, beta_in, g_out, delta_out, dT, dcoef, st)
    int rc;
    if (a.jbits == 4) rc = a.counted ? ANC_MF_BWD(true, 4) : ANC_MF_BWD(false, 4);
    else if (a.jbits == 2) rc = a.counted ? ANC_MF_BWD(true, 2) : ANC_MF_BWD(false, 2);
    else rc = a.counted ? ANC_MF_BWD(true, 8) : ANC_MF_BWD(false, 8);
#undef ANC_MF_BWD
    if (rc || a.counted) return rc;
    const int items = (t1 - t0) * a.B;
    hipLaunchKernelGGL(mf::mf_team_sums, dim3((items + 255) / 256), dim3(256), 0, st, reinterpret_cast<const double2*>(x.part), t0,
                       t1 - t0, a.B, a.T, dT, dcoef);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace anc
