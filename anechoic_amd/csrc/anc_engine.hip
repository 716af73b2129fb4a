// Device side of the MI355X SCA/Onsager engine: instance preparation kernels and the persistent team kernel.
//
// Arithmetic contract (reference fpga/v80_sca/src/sca_ref_bias.hpp:25-75, generalised NP):
//   per anneal, 256 xoshiro128** lanes seeded from (seed, trial, lane) (sca_ref.hpp:71-91); slot i = k*256 + l is drawn
//   by lane l in round k; padding draws are consumed; s_i(0) = +1 iff the draw's top bit is set;
//   h(0) = b + J s(0); each step t: corr = ((n_lin(t-1) kcorr[t]) >> 8) + kconst[t];
//   z_i = s_i h_i 2^16 + q[t] - (s_i == s'_i ? corr : -corr); r_i = ((u16 - 32768) fourT[t]) >> 16;
//   n_lin counts -2T < z < 2T (twoT = fourT >> 1); slot i flips iff z_i < r_i; then h += J (s_new - s_old).
//
// Engine layout (MI355X, 8 XCDs x 32 CUs):
//   * A team is 32 workgroups on one XCD (workgroup b runs on XCD b % 8). Team member c owns RNG lanes 8c..8c+7, i.e.
//     the 64 slots x = k*256 + l of those lanes, renumbered x' = l*8 + k so that every member owns one contiguous
//     64-slot block [64c, 64c+64) of the permuted index.
//   * Each member keeps its 64 rows of the permuted coupling matrix J'[y'][x'] resident in VGPRs as int8 MFMA
//     A-fragments (wave w holds the K-quarter x' in [512w, 512w+512)), and its part of the field H[64 x RT] as four
//     per-wave int32 MFMA accumulators (H = bias + sum_w P_w).
//   * One step = decide (256 threads = RT replicas x 8 lanes; thread owns its lane's RNG stream and 8 slots) ->
//     publish the member's flip matrix slice D' = D + tau (int8, D in {-2,0,2}, tau = step parity bit in bit 0 of
//     every byte) directly in MFMA B-fragment layout plus a tagged n_lin partial per replica -> every member reads the
//     D' fragments of its team (self-validating: a byte is current iff its bit 0 equals tau) ->
//     v_mfma_i32_16x16x64_i8 accumulation P_w += J'_w D'_w -> the tau contribution tau*rowsum(J) is removed exactly
//     in the decision threads.
//   * No fence or barrier crosses workgroups: every exchanged word carries its own validity, so relaxed loads and
//     stores suffice (the exchange never relies on ordering between different words).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "anc_device.hpp"
#include "anc_kernels.hpp"

namespace anc {
namespace dev {

// ---------------------------------------------------------------------------------------------------------- prep kernels
// Instance check (jmat.hpp:35-55 rules): J symmetric, zero diagonal; field bound max_y sum_x |J_xy| and max |b|.
// stats: [0] asymmetric entries, [1] nonzero diagonal entries, [2] max row abs sum, [3] max |b|.
__global__ void check_instance(const int8_t* __restrict__ J, int64_t ldj, int n, const int32_t* __restrict__ bias,
                               unsigned long long* stats) {
    const int y = blockIdx.x;
    if (y >= n) return;
    const int8_t* row = J + size_t(y) * size_t(ldj);
    unsigned long long asym = 0, diag = 0, rabs = 0;
    for (int x = threadIdx.x; x < n; x += blockDim.x) {
        const int v = row[x];
        rabs += unsigned(v < 0 ? -v : v);
        if (x == y) diag += v != 0;
        else if (v != J[size_t(x) * size_t(ldj) + y]) asym += 1;
    }
    for (int o = 32; o > 0; o >>= 1) {
        asym += __shfl_xor(asym, o); diag += __shfl_xor(diag, o); rabs += __shfl_xor(rabs, o);
    }
    __shared__ unsigned long long red[3][16];
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) { red[0][w] = asym; red[1][w] = diag; red[2][w] = rabs; }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long a = 0, d = 0, r = 0;
        for (int i = 0; i < int(blockDim.x / 64); ++i) { a += red[0][i]; d += red[1][i]; r += red[2][i]; }
        if (a) atomicAdd(stats + 0, a);
        if (d) atomicAdd(stats + 1, d);
        atomicMax(stats + 2, r);
        if (bias) {
            const int b = bias[y];
            atomicMax(stats + 3, (unsigned long long)(b < 0 ? -(long long)b : (long long)b));
        }
    }
}

// J'[y'][x'] = J[y][x] with x' = l*krk + k for x = k*256 + l (zero outside n). krk = the kernel's compile-time round
// count; rounds k >= NP(n)/256 are phantom slots (x >= n), so they are zero as well. One block per permuted row.
__global__ void permute_instance(const int8_t* __restrict__ J, int64_t ldj, int n, int npk, int krk, int8_t* __restrict__ Jp) {
    const int yp = blockIdx.x;
    const int y = (yp % krk) * LANES + yp / krk;
    int8_t* out = Jp + size_t(yp) * size_t(npk);
    for (int xp = threadIdx.x; xp < npk; xp += blockDim.x) {
        const int x = (xp % krk) * LANES + xp / krk;
        out[xp] = (x < n && y < n) ? J[size_t(y) * size_t(ldj) + x] : int8_t(0);
    }
}

// gen_ising random (gpu/multibit_bias_20261008/src/gen_ising.cpp:118-135) on the device: pair p = (x < y) in
// row-major order uses draws 2p and 2p + 1 of the splitmix64 stream seeded with seed ^ 0xA5A5F00D12345678 (the
// stream is counter-based: draw k is mix(state + (k + 1) gamma)); |J| = 1 + (r1 >> 1) % M, negative iff r1 & 1,
// zero iff r0 < zthr. Row x of the grid writes J[x][y] and J[y][x] for y > x, zeros on the diagonal and in the
// padding columns n .. ldj - 1.
__device__ __forceinline__ uint64_t splitmix_at(uint64_t state, uint64_t k) {
    uint64_t x = state + k * 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__global__ void gen_ising_pairs(uint64_t seed, int n, int M, uint64_t zthr, int8_t* __restrict__ J, int64_t ldj) {
    const int x = blockIdx.y;
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= ldj) return;
    int8_t* row = J + size_t(x) * size_t(ldj);
    if (y <= x || y >= n) {
        if (y == x || y >= n) row[y] = 0;
        return;
    }
    const uint64_t st = seed ^ 0xA5A5F00D12345678ull;
    const int64_t p = int64_t(x) * (n - 1) - int64_t(x) * (x - 1) / 2 + (y - x - 1);
    const uint64_t r0 = splitmix_at(st, uint64_t(2 * p + 1)), r1 = splitmix_at(st, uint64_t(2 * p + 2));
    int v = 0;
    if (!(zthr && r0 < zthr)) {
        const int mag = 1 + int((r1 >> 1) % uint64_t(M));
        v = (r1 & 1) ? -mag : mag;
    }
    row[y] = int8_t(v);
    J[size_t(y) * size_t(ldj) + x] = int8_t(v);
}

// Biases of gen_ising random (pbig >= 1): b_i = draw_i % (2 bmax + 1) - bmax from seed ^ 0x5B1A5B1A5B1A5B1A.
__global__ void gen_ising_bias(uint64_t seed, int n, int bmax, int32_t* __restrict__ b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    b[i] = bmax > 0 ? int32_t(int64_t(splitmix_at(seed ^ 0x5B1A5B1A5B1A5B1Aull, uint64_t(i + 1)) % uint64_t(2 * bmax + 1)) - bmax)
                    : 0;
}

// Zeroes n 64-bit words (exchange workspace, status and the accumulated outputs) before a team launch.
__global__ void fill_zero(uint64_t* __restrict__ p, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) p[i] = 0;
}

// ---------------------------------------------------------------------------------------------------------- team kernel
// One team member's whole run. TEAM = workgroups per team (32: one XCD; 64: two XCDs), KR = rounds (the kernel's
// NP/256), NT = 16-replica N-tiles per team (RT = 16 NT replicas per work item), WIDE = int64 decision thresholds,
// XFLAVOR = exchange load flavour; `team` and member index `c` come from the launching kernel's placement.
template <int TEAM, int KR, int NT, bool WIDE, int XFLAVOR>
__device__ __forceinline__ void team_body(const TeamArgs& a, int team, int c) {
    constexpr int LPC = LANES / TEAM;    // RNG lanes per member
    constexpr int RT = 16 * NT;          // replicas per item
    constexpr int SPC = LPC * KR;        // slots per member (64 for 32 x 8, 48 for 64 x 12)
    constexpr int MT = SPC / 16;         // M-tiles (and 16-slot k-groups) per member
    constexpr int KBW = KR;              // K-blocks of 64 per wave: each wave owns a quarter of the NP columns
    constexpr int KG = 16 * KR;          // k-groups of 16 slots in NP
    constexpr int NQ = TEAM / LPC;       // n_lin words gathered per decision thread
    constexpr int PPITCH = SPC + 4;      // LDS partial row pitch (int32), padded against bank conflicts
    static_assert(SPC % 16 == 0 && KR % 4 == 0 && THREADS >= RT * LPC, "unsupported team shape");
    using wide_t = typename std::conditional<WIDE, int64_t, int32_t>::type;

    __shared__ int32_t P[2][4][RT][PPITCH];   // per-wave partial fields, double-buffered by step parity
    __shared__ int32_t RS[4][SPC];            // per-wave row sums of J' (tau correction)

    if (team >= a.n_teams) return;
    const int tid = threadIdx.x, w = tid / 64, L = tid % 64;
    const int g = L / 16, ml = L % 16;   // MFMA lane coordinates

    // ---- exchange addressing (per team; reused by every work item of the team)
    // fragments by k-group: [parity][k-group][n-tile][16 replicas][16 B]
    const size_t team_frag = size_t(team) * 2 * KG * NT * 32;   // uint64 units
    const size_t team_nlin = size_t(team) * 2 * TEAM * RT;
    __amdgpu_buffer_rsrc_t frag_rsrc = __builtin_amdgcn_make_buffer_rsrc(a.xfrag + team_frag, 0, 0x7fffffff, 0x00020000);
    __amdgpu_buffer_rsrc_t nlin_rsrc = __builtin_amdgcn_make_buffer_rsrc(a.xnlin + team_nlin, 0, 0x7fffffff, 0x00020000);

    // ---- decision thread identity: replica n (in the item), lane j (in member)
    const int n = tid / LPC, j = tid % LPC;
    const bool dthread = n < RT;
    const int lane = LPC * c + j;        // global RNG lane
    auto nlin_off = [&](int u, int q4) { return int(((size_t(u & 1) * TEAM + (NQ * j + q4)) * RT + n) * 4); };
    const uint64_t t_limit = uint64_t(a.timeout_ticks);   // per wait, in wall-clock ticks
    uint64_t t_wait = 0;
    bool bail = false;

    // optional per-phase cycle counters of thread 0 (the reference's SCA_PROFILE, sca_gpu.cu:163-168)
    const bool prof = a.prof != nullptr && tid == 0;
    uint64_t pacc[6] = {0, 0, 0, 0, 0, 0};
    uint64_t ptick = prof ? __builtin_amdgcn_s_memtime() : 0;
    auto mark = [&](int i) {
        if (prof) {
            const uint64_t now = __builtin_amdgcn_s_memtime();
            pacc[i] += now - ptick;
            ptick = now;
        }
    };

    // ---- work items: RT-replica slices of one instance (a.items == null) or of the instances of a descriptor table.
    // Team `team` takes items team, team + n_teams, ...; the exchange step counter u continues across the items of a
    // team, so parity and tags never repeat between consecutive items.
    const int n_items = a.items ? a.n_items : (a.R + RT - 1) / RT;
    int seq = 0;
    for (int item = team; item < n_items; item += a.n_teams, ++seq) {
    __syncthreads();   // the previous item's LDS reads are complete
    const int8_t* Jp = a.Jp;
    const int32_t* ibias = a.bias;
    int in_n = a.n, r0 = item * RT, cnt = min(RT, a.R - item * RT), kr = a.kr;
    if (a.items) {
        const TeamItem it = a.items[item];
        Jp = it.Jp; ibias = it.bias; in_n = it.n; r0 = it.r0; cnt = it.count; kr = it.kr;
    }
    const int ubase = seq * (a.S + 1);
    const int r = r0 + n;
    const bool active = dthread && n < cnt;

    // ---- coupling fragments: A[mt][i] = J'[y' = 64c + 16mt + ml][x' = 64(8w + i) + 16g .. +16]
    v4i A[MT][KBW];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int i = 0; i < KBW; ++i) {
            const int8_t* src = Jp + size_t(SPC * c + 16 * mt + ml) * size_t(a.np) + size_t(64 * (KBW * w + i) + 16 * g);
            A[mt][i] = *reinterpret_cast<const v4i*>(src);
        }
    // ---- row sums of J' over this wave's K-quarter, by MFMA with an all-ones B
    {
        const v4i ones = {0x01010101, 0x01010101, 0x01010101, 0x01010101};
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            v4i acc = {0, 0, 0, 0};
#pragma unroll
            for (int i = 0; i < KBW; ++i) acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[mt][i], ones, acc, 0, 0, 0);
            if (ml == 0)
#pragma unroll
                for (int e = 0; e < 4; ++e) RS[w][16 * mt + 4 * g + e] = acc[e];
        }
    }

    Rng rng;
    uint32_t sbits = ~0u, spbits = ~0u;   // bit k: spin of slot k*256+lane is +1 (all KR rounds start at +1)
    int32_t bias_k[KR], rowsum_k[KR];
    int32_t nflips = 0;
    if (active) rng.seed(a.seed, a.trials ? a.trials[r] : a.trial0 + uint32_t(r), uint32_t(lane));
    // initial spins (padding draws consumed: sca_ref_bias.hpp:31-36)
#pragma unroll
    for (int k = 0; k < KR; ++k) {
        const int x = k * LANES + lane;
        bias_k[k] = 0;
        if (active) {
            const uint32_t rv = k < kr ? rng.next() : 0u;   // rounds k >= kr are phantom slots: no draw
            if (x < in_n && !(rv >> 31)) sbits &= ~(1u << k);
            if (x < in_n && ibias) bias_k[k] = ibias[x];
        }
    }
    // publish D_0 = s(0) - 1 (0 or -2 for x < n, 0 for padding) with tau(0)
    auto publish = [&](int u, const int (&d)[KR]) {
        if (!dthread) return;
        const uint32_t tau = tau_of(u);
        uint32_t* base = reinterpret_cast<uint32_t*>(a.xfrag + team_frag);
        // member-local byte position of slot k of this thread: p = KR j + k; k-group p / 16, byte p % 16
#pragma unroll
        for (int q4 = 0; q4 < KR / 4; ++q4) {
            uint32_t word = 0;
#pragma unroll
            for (int e = 0; e < 4; ++e) word |= dbyte(d[4 * q4 + e], tau) << (8 * e);
            const int pos = KR * j + 4 * q4;
            const int kg = MT * c + pos / 16;
            const size_t off = ((((size_t(u & 1) * KG + kg) * NT + n / 16) * 16 + (n % 16)) * 16 + (pos % 16)) / 4;
            xstore4<XFLAVOR>(base + off, word);
        }
    };
    {
        int d0[KR];
#pragma unroll
        for (int k = 0; k < KR; ++k) {
            const int x = k * LANES + lane;
            // D_0 = s(0) - 1 (even, so the tau bit can be ORed in); H adds back J 1 = rowsum
            d0[k] = (active && x < in_n && !((sbits >> k) & 1)) ? -2 : 0;
        }
        publish(ubase, d0);
    }
    __syncthreads();   // RS visible
    if (dthread) {
#pragma unroll
        for (int k = 0; k < KR; ++k) {
            const int m = KR * j + k;
            rowsum_k[k] = RS[0][m] + RS[1][m] + RS[2][m] + RS[3][m];
        }
    }

    v4i acc[MT][NT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = v4i{0, 0, 0, 0};

    int32_t nlin_prev = 0;
    int32_t H[KR];
    int32_t rr_k[KR];

    for (int t = 0; t <= a.S; ++t) {
        const int u = ubase + t;   // tag of the D' fragments consumed this iteration
        const bool want_nlin = t >= 1 && dthread;
        // step-t tables, loaded before the exchange wait so that their latency is hidden
        int64_t kc_t = 0;
        int32_t kk_t = 0, f4_t = 1, q_t = 0;
        if (t < a.S) { kc_t = a.kcorr[t]; kk_t = a.kconst[t]; f4_t = a.fourT[t]; q_t = a.q[t]; }
        // ===== MFMA phase: P_w += J'_w (D_u + tau_u)
        {
            const uint32_t tau = tau_of(u);
            const uint32_t want = tau ? 0x01010101u : 0u;
            auto frag_off = [&](int i, int nt) {   // B-fragment lane (g, ml) = k-group 4 kb + g, replica ml
                return int((((size_t(u & 1) * KG + 4 * (KBW * w + i) + g) * NT + nt) * 16 + ml) * 16);
            };
            auto current = [&](const v4i (&b)[NT]) {
                uint32_t all_and = 0xFFFFFFFFu, all_or = 0;
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
#pragma unroll
                    for (int e = 0; e < 4; ++e) { all_and &= uint32_t(b[nt][e]); all_or |= uint32_t(b[nt][e]); }
                return __all(((tau ? all_and : all_or) & 0x01010101u) == want);
            };
            // first poll: every fragment load issued back to back, validated and consumed in issue order
            v4i B[KBW][NT];
#pragma unroll
            for (int i = 0; i < KBW; ++i)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) B[i][nt] = xload16<XFLAVOR>(frag_rsrc, frag_off(i, nt));
            // while the loads are in flight: this step's random thresholds r_k = floor((u16 - 32768) fourT / 2^16)
            // (they depend only on the lane's stream and the table, not on the exchanged data)
            if (active && t < a.S) {
                const int32_t f_hi = f4_t >> 16, f_lo = f4_t & 0xFFFF;
#pragma unroll
                for (int k = 0; k < KR; ++k) {
                    if (k >= kr) break;   // phantom rounds draw nothing
                    const int32_t av = int32_t(rng.next() >> 16) - 32768;
                    rr_k[k] = av * f_hi + ((av * f_lo) >> 16);   // exact: av fourT = av f_hi 2^16 + av f_lo
                }
            }
            uint32_t pending = 0;
#pragma unroll
            for (int i = 0; i < KBW; ++i) {
                if (current(B[i])) {
#pragma unroll
                    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                        for (int nt = 0; nt < NT; ++nt)
                            acc[mt][nt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[mt][i], B[i][nt], acc[mt][nt], 0, 0, 0);
                } else {
                    pending |= 1u << i;
                }
            }
            // re-polls of the fragments that were not yet current
            int polls = 0;
            while (pending && !bail) {
                if ((++polls & 63) == 0) {
                    const uint64_t now = wall_clock64();
                    if (polls == 64) t_wait = now;
                    if (aborted(a.status)) { bail = true; break; }
                    if (now - t_wait > t_limit) { raise_error(a.status, 1); bail = true; break; }
                }
                __builtin_amdgcn_s_sleep(1);
#pragma unroll
                for (int i = 0; i < KBW; ++i) {
                    if (!((pending >> i) & 1u)) continue;
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt) B[i][nt] = xload16<XFLAVOR>(frag_rsrc, frag_off(i, nt));
                    if (current(B[i])) {
#pragma unroll
                        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                            for (int nt = 0; nt < NT; ++nt)
                                acc[mt][nt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[mt][i], B[i][nt], acc[mt][nt], 0, 0, 0);
                        pending &= ~(1u << i);
                    }
                }
            }
        }
        mark(0);
        // ===== n_lin of step t-1: published just before the D_u fragments, so it is (almost always) visible once the
        // fragments have been consumed; team total for replica n
        if (want_nlin && !bail) {
            uint32_t nv[NQ];
#pragma unroll
            for (int q4 = 0; q4 < NQ; ++q4) nv[q4] = xload4<XFLAVOR>(nlin_rsrc, nlin_off(u, q4));
            const uint32_t want = uint32_t(u + 1) & 0xFFFFFu;
            int polls = 0;
            while (true) {
                bool ok = true;
#pragma unroll
                for (int q4 = 0; q4 < NQ; ++q4) ok = ok && (nv[q4] >> 12) == want;
                if (ok) break;
                if ((++polls & 63) == 0) {
                    const uint64_t now = wall_clock64();
                    if (polls == 64) t_wait = now;
                    if (aborted(a.status)) { bail = true; break; }
                    if (now - t_wait > t_limit) { raise_error(a.status, 1); bail = true; break; }
                }
                __builtin_amdgcn_s_sleep(1);
#pragma unroll
                for (int q4 = 0; q4 < NQ; ++q4)
                    if ((nv[q4] >> 12) != want) nv[q4] = xload4<XFLAVOR>(nlin_rsrc, nlin_off(u, q4));
            }
            int32_t part = 0;
#pragma unroll
            for (int q4 = 0; q4 < NQ; ++q4) part += int32_t(nv[q4] & 0xFFFu);
#pragma unroll
            for (int o = 1; o < LPC; o <<= 1) part += __shfl_xor(part, o, LPC);
            nlin_prev = part;
            if (active && c == 0 && j == 0 && a.nlin) a.nlin[size_t(r) * size_t(a.S) + (t - 1)] = part;
        }
        mark(1);
        // ===== partial fields to LDS, then H = bias + sum_w P_w - tau_count * rowsum
        {
            const int buf = t & 1;
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    *reinterpret_cast<v4i*>(&P[buf][w][16 * nt + ml][16 * mt + 4 * g]) = acc[mt][nt];
        }
        if (__syncthreads_or(bail ? 1 : 0)) { bail = true; break; }
        mark(2);
        if (dthread) {
            const int buf = t & 1;
            // acc = J'(s(0) - 1) + sum_u J' D_u + c rowsum, c = number of exchange steps so far with tau = 1;
            // with tc = c - 1: H = bias + acc - tc rowsum
            const int tc = tau_count(u) - tau_count(ubase - 1) - 1;
#pragma unroll
            for (int k = 0; k < KR; ++k) {
                const int m = KR * j + k;
                H[k] = bias_k[k] + P[buf][0][n][m] + P[buf][1][n][m] + P[buf][2][n][m] + P[buf][3][n][m] - tc * rowsum_k[k];
            }
        }
        mark(3);
        if (t == a.S) break;
        // ===== decision phase of step t
        int d[KR];
        int32_t nlin_part = 0;
        if (active) {
            const int64_t corr = ((int64_t(nlin_prev) * kc_t) >> 8) + int64_t(kk_t);
            const int32_t twoT = f4_t >> 1;
            const int64_t c_same = int64_t(q_t) - corr, c_diff = int64_t(q_t) + corr;
            // per-step, per-class constants: flip <=> v <= (r - (c+1)) >> 16 ; band <=> lo(c) <= v <= hi(c)
            wide_t c1[2] = {wide_t(c_diff + 1), wide_t(c_same + 1)};
            wide_t lo[2] = {wide_t(((-int64_t(twoT) - c_diff) >> 16) + 1), wide_t(((-int64_t(twoT) - c_same) >> 16) + 1)};
            wide_t hi[2] = {wide_t((int64_t(twoT) - c_diff - 1) >> 16), wide_t((int64_t(twoT) - c_same - 1) >> 16)};
            const uint32_t sold = sbits;
            const uint32_t same_bits = ~(sold ^ spbits);
#pragma unroll
            for (int k = 0; k < KR; ++k) {
                const int x = k * LANES + lane;
                const int32_t rr = rr_k[k];
                const bool s_up = (sold >> k) & 1;
                const int cls = (same_bits >> k) & 1;
                const int32_t v = s_up ? H[k] : -H[k];
                // bitwise & (not &&): no short-circuit branches in the unrolled loop
                const bool flip = wide_t(v) <= ((wide_t(rr) - c1[cls]) >> 16);
                const bool band = (wide_t(v) >= lo[cls]) & (wide_t(v) <= hi[cls]);
                const bool valid = x < in_n;
                nlin_part += int(valid & band);
                const bool fl = valid & flip;
                d[k] = fl ? (s_up ? -2 : 2) : 0;
                if (fl) sbits ^= (1u << k);
                nflips += fl ? 1 : 0;
            }
            spbits = sold | ~((1u << KR) - 1u);
            // padding slots keep s = s' = +1 (their bits never change)
        } else {
#pragma unroll
            for (int k = 0; k < KR; ++k) d[k] = 0;
        }
        mark(4);
        // the n_lin partial goes out first, so that it is visible no later than the fragments it travels with
        if (dthread) {
#pragma unroll
            for (int o = 1; o < LPC; o <<= 1) nlin_part += __shfl_xor(nlin_part, o, LPC);
            if (j == 0) {
                uint32_t* dst = a.xnlin + team_nlin + (size_t((u + 1) & 1) * TEAM + c) * RT + n;
                xstore4<XFLAVOR>(dst, (((uint32_t(u + 2)) & 0xFFFFFu) << 12) | uint32_t(nlin_part));
            }
        }
        publish(u + 1, d);
        mark(5);
    }
    // ===== outputs (sca_ref_bias.hpp:69-73): spins, sum s h, sum b s, flips, optional fields
    if (active && !bail) {
        int64_t sh = 0, bs = 0;
#pragma unroll
        for (int k = 0; k < KR; ++k) {
            const int x = k * LANES + lane;
            if (x >= in_n) continue;
            const bool s_up = (sbits >> k) & 1;
            sh += s_up ? int64_t(H[k]) : -int64_t(H[k]);
            bs += s_up ? int64_t(bias_k[k]) : -int64_t(bias_k[k]);
            if (s_up) atomicOr(reinterpret_cast<unsigned long long*>(a.spins + size_t(r) * size_t(a.spin_stride) + (x >> 6)),
                               (unsigned long long)1 << (x & 63));
            if (a.fields) a.fields[size_t(r) * size_t(a.field_stride) + x] = H[k];
        }
        int64_t fl = nflips;
#pragma unroll
        for (int o = 1; o < LPC; o <<= 1) {
            sh += __shfl_xor(sh, o, LPC); bs += __shfl_xor(bs, o, LPC); fl += __shfl_xor(fl, o, LPC);
        }
        if (j == 0) {
            atomicAdd(reinterpret_cast<unsigned long long*>(a.sum_sh + r), (unsigned long long)sh);
            atomicAdd(reinterpret_cast<unsigned long long*>(a.sum_bs + r), (unsigned long long)bs);
            atomicAdd(reinterpret_cast<unsigned long long*>(a.flips + r), (unsigned long long)fl);
        }
    }
    if (bail) break;
    }   // work items
    if (prof) {
#pragma unroll
        for (int i = 0; i < 6; ++i) a.prof[size_t(blockIdx.x) * 8 + i] = (long long)pacc[i];
        a.prof[size_t(blockIdx.x) * 8 + 6] = int64_t(seq) * a.S;
        a.prof[size_t(blockIdx.x) * 8 + 7] = team * 1000 + c;
    }
}

// Single-shape launch. TEAM = 32: XCD-local teams (workgroup b runs on XCD b % 8, so team members are the workgroups
// with equal b % 8); TEAM = 64: consecutive workgroups (device-scope exchange).
template <int TEAM, int KR, int NT, bool WIDE, int XFLAVOR>
__global__ void __launch_bounds__(THREADS, 1) team_kernel(TeamArgs a) {
    int team, c;
    if (TEAM == 32) {
        const int xcd = blockIdx.x % 8, idx = blockIdx.x / 8;
        team = (idx / TEAM) * 8 + xcd;
        c = idx % TEAM;
    } else {
        team = blockIdx.x / TEAM;
        c = blockIdx.x % TEAM;
    }
    team_body<TEAM, KR, NT, WIDE, XFLAVOR>(a, team, c);
}

// Both team shapes in one persistent launch (instances of uneven size): one workgroup per CU, grid 8 XCDs x 32.
// XCDs 0 .. xcd2-1 each host one 32-workgroup team of the NP = 2048 items (a2, exchange flavour XFLAVOR); the
// remaining XCDs, in pairs, host 64-workgroup teams of the NP = 2304..3072 items (a3, device-scope exchange).
template <int NT, bool WIDE, int XFLAVOR>
__global__ void __launch_bounds__(THREADS, 1) team_kernel_mixed(TeamArgs a2, TeamArgs a3, int xcd2) {
    const int xcd = blockIdx.x % 8, idx = blockIdx.x / 8;
    if (idx >= 32) return;
    if (xcd < xcd2) team_body<32, 8, NT, WIDE, XFLAVOR>(a2, xcd, idx);
    else team_body<64, 12, NT, WIDE, 1>(a3, (xcd - xcd2) / 2, ((xcd - xcd2) % 2) * 32 + idx);
}

}  // namespace dev

// ---------------------------------------------------------------------------------------------------------- launchers
int launch_check_instance(const int8_t* J, int64_t ldj, int n, const int32_t* bias, unsigned long long* stats, hipStream_t st) {
    hipLaunchKernelGGL(dev::check_instance, dim3(n), dim3(256), 0, st, J, ldj, n, bias, stats);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_permute(const int8_t* J, int64_t ldj, int n, int npk, int8_t* Jp, hipStream_t st) {
    hipLaunchKernelGGL(dev::permute_instance, dim3(npk), dim3(256), 0, st, J, ldj, n, npk, npk / 256, Jp);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_gen_ising(uint64_t seed, int n, int M, uint64_t zthr, int bmax, int8_t* J, int64_t ldj, int32_t* bias, hipStream_t st) {
    hipLaunchKernelGGL(dev::gen_ising_pairs, dim3(unsigned((ldj + 255) / 256), unsigned(n)), dim3(256), 0, st, seed, n, M, zthr, J, ldj);
    if (bias) hipLaunchKernelGGL(dev::gen_ising_bias, dim3((n + 255) / 256), dim3(256), 0, st, seed, n, bmax, bias);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_fill_zero(uint64_t* p, int64_t words, hipStream_t st) {
    if (words <= 0) return 0;
    int blocks = int((words + 255) / 256);
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(dev::fill_zero, dim3(blocks), dim3(256), 0, st, p, words);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Dynamic LDS that pads a workgroup above half of the CU's 160 KiB LDS, so that no CU hosts two team members, and a
// check that one such workgroup fits a CU. Cached per kernel (the first call must not be under stream capture).
static int member_pad(const void* kern, size_t& pad) {
    hipFuncAttributes fa;
    if (hipFuncGetAttributes(&fa, kern) != hipSuccess) return -1;
    const size_t half = 80 * 1024 + 1024;
    pad = fa.sharedSizeBytes < half ? half - fa.sharedSizeBytes : 0;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, dev::THREADS, pad) != hipSuccess) return -1;
    return per_cu >= 1 ? 0 : -2;
}

template <typename K>
static int cached_pad(K kern, size_t& pad) {
    static int rc = 1;
    static size_t cached = 0;
    if (rc == 1) rc = member_pad((const void*)kern, cached);
    pad = cached;
    return rc;
}

template <int TEAM, int KR, int NT, bool WIDE, int XF>
static int launch_team_t(const TeamArgs& a, int cus, bool dry, hipStream_t st) {
    const auto kern = dev::team_kernel<TEAM, KR, NT, WIDE, XF>;
    const int grid = TEAM == 32 ? ((a.n_teams + 7) / 8) * 8 * TEAM : a.n_teams * TEAM;
    size_t pad = 0;
    if (int rc = cached_pad(kern, pad)) return rc;
    if (a.n_teams * TEAM > cus || grid > cus) return -2;
    if (dry) return 0;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(dev::THREADS), pad, st, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int TEAM, int KR, int NT, bool WIDE>
static int launch_flavour(const TeamArgs& a, int xflavor, int cus, bool dry, hipStream_t st) {
    if (xflavor == 2) return launch_team_t<TEAM, KR, NT, WIDE, 2>(a, cus, dry, st);
    if (xflavor == 3) return launch_team_t<TEAM, KR, NT, WIDE, 3>(a, cus, dry, st);
    return launch_team_t<TEAM, KR, NT, WIDE, 1>(a, cus, dry, st);
}

template <int TEAM, int KR>
static int launch_shape(const TeamArgs& a, int nt, bool wide, int xflavor, int cus, bool dry, hipStream_t st) {
    if (nt == 1) return wide ? launch_flavour<TEAM, KR, 1, true>(a, xflavor, cus, dry, st) : launch_flavour<TEAM, KR, 1, false>(a, xflavor, cus, dry, st);
    return wide ? launch_flavour<TEAM, KR, 2, true>(a, xflavor, cus, dry, st) : launch_flavour<TEAM, KR, 2, false>(a, xflavor, cus, dry, st);
}

int kernel_np_for(int np) { return np == 2048 ? 2048 : (np > 2048 && np <= 3072) ? 3072 : 0; }
int team_size_for(int np) { const int k = kernel_np_for(np); return k == 2048 ? 32 : k == 3072 ? 64 : 0; }

int launch_team(const TeamArgs& a, int nt, bool wide, int xflavor, int cus, bool dry, hipStream_t st) {
    if (a.np == 2048) return launch_shape<32, 8>(a, nt, wide, xflavor, cus, dry, st);
    if (a.np == 3072) return launch_shape<64, 12>(a, nt, wide, 1, cus, dry, st);   // two XCDs per team: device scope
    return -3;
}

template <int NT, bool WIDE, int XF>
static int launch_mixed_t(const TeamArgs& a2, const TeamArgs& a3, int xcd2, int cus, bool dry, hipStream_t st) {
    const auto kern = dev::team_kernel_mixed<NT, WIDE, XF>;
    size_t pad = 0;
    if (int rc = cached_pad(kern, pad)) return rc;
    if (cus < 256 || a2.n_teams > xcd2 || 2 * a3.n_teams > 8 - xcd2) return -2;
    if (dry) return 0;
    hipLaunchKernelGGL(kern, dim3(256), dim3(dev::THREADS), pad, st, a2, a3, xcd2);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_team_mixed(const TeamArgs& a2, const TeamArgs& a3, int xcd2, int nt, bool wide, int xflavor, int cus, bool dry,
                      hipStream_t st) {
    if (a2.np != 2048 || a3.np != 3072 || xcd2 < 1 || xcd2 > 7 || (8 - xcd2) % 2) return -3;
    const bool x3 = xflavor != 1;
    if (nt == 1) {
        if (wide) return x3 ? launch_mixed_t<1, true, 3>(a2, a3, xcd2, cus, dry, st) : launch_mixed_t<1, true, 1>(a2, a3, xcd2, cus, dry, st);
        return x3 ? launch_mixed_t<1, false, 3>(a2, a3, xcd2, cus, dry, st) : launch_mixed_t<1, false, 1>(a2, a3, xcd2, cus, dry, st);
    }
    if (wide) return x3 ? launch_mixed_t<2, true, 3>(a2, a3, xcd2, cus, dry, st) : launch_mixed_t<2, true, 1>(a2, a3, xcd2, cus, dry, st);
    return x3 ? launch_mixed_t<2, false, 3>(a2, a3, xcd2, cus, dry, st) : launch_mixed_t<2, false, 1>(a2, a3, xcd2, cus, dry, st);
}

}  // namespace anc
