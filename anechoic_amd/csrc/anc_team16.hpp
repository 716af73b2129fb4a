// The 16-member team kernel (n <= 2048, spin crumbs): 16 workgroups on one XCD anneal 16 replicas per work item.
// anc_team16.hip instantiates it.
//
// Arithmetic contract (reference fpga/v80_sca/src/sca_ref_bias.hpp:25-75): see anc_engine.hip. Exchange encoding and
// tags as in anc_team.hpp (mode 3): every slot is a 2-bit crumb N = s + 1 + tau(u), and the field of each step is
// recomputed as H = b + J' N - (1 + tau) rowsum(J') with int32 accumulators (exact).
//
// Shape. Member c (one workgroup, one CU) owns the 128 slots of RNG lanes L = 16 c .. 16 c + 15 (8 rounds each). Wave w
// holds the J' rows of lanes L = 16 c + 4 w + g (g = 0..3) over all 2048 columns as int8 MFMA A fragments (two 16-row
// M-tiles, 256 registers per lane). The rows are ordered so that the accumulator element e of M-tile mt in MFMA lane
// (g, ml) is round k = 4 mt + e of RNG lane 16 c + 4 w + g for replica ml: every lane decides the 8 slots of one
// (replica, RNG lane) straight from its accumulators, with no cross-wave reduction of partial fields.
//
// Exchange per step: the team's crumbs of 16 replicas (8 KiB) and the per-(member, wave) n_lin bytes (1 KiB) are
// loaded once per workgroup (two 16-byte chunks per thread), validated by their tags, re-polled until current and
// staged in LDS (double-buffered by step parity); after one workgroup barrier every wave reads its B fragments from LDS.
// Measured on MI355X with a probe of the same exchange (tagged 16-byte chunks published with plain stores and polled
// with non-temporal loads, one workgroup barrier per step, 32 workgroups of one XCD): about 0.5 us per step plus about
// 30 ns per KiB each workgroup reads, so 16-replica items (8 KiB) save about 0.3 us per step against 32-replica items
// (16 KiB).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "anc_device.hpp"
#include "anc_kernels.hpp"

namespace anc {
namespace dev {

// n_lin byte tag of exchange step u: ((u mod 3) + 1) << 6. It differs from the tag two steps earlier (the previous
// use of the same parity buffer) and is never 0, so zeroed memory never reads as current.
__device__ __forceinline__ uint32_t nlin_tag(int u) { return uint32_t((u % 3) + 1) << 6; }

// Sum of v over the four 16-lane rows of the wave (lanes l, l ^ 16, l ^ 32, l ^ 48), in every lane: two cross-row
// swaps on the VALU (gfx950 v_permlane16_swap / v_permlane32_swap), no LDS round trip.
__device__ __forceinline__ int32_t rows4_sum(int32_t v) {
    const auto a = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    const int32_t s = int32_t(a[0]) + int32_t(a[1]);
    const auto b = __builtin_amdgcn_permlane32_swap(s, s, false, false);
    return int32_t(b[0]) + int32_t(b[1]);
}
__device__ __forceinline__ int64_t rows4_sum64(int64_t v) {
    const int32_t lo = int32_t(uint32_t(v)), hi = int32_t(uint64_t(v) >> 32);
    const auto a = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto ah = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    const int64_t s = (int64_t(uint32_t(a[0])) | (int64_t(ah[0]) << 32)) + (int64_t(uint32_t(a[1])) | (int64_t(ah[1]) << 32));
    const int32_t slo = int32_t(uint32_t(s)), shi = int32_t(uint64_t(s) >> 32);
    const auto b = __builtin_amdgcn_permlane32_swap(slo, slo, false, false);
    const auto bh = __builtin_amdgcn_permlane32_swap(shi, shi, false, false);
    return (int64_t(uint32_t(b[0])) | (int64_t(bh[0]) << 32)) + (int64_t(uint32_t(b[1])) | (int64_t(bh[1]) << 32));
}

template <int XFLAVOR>
__device__ __forceinline__ void xstore1(uint8_t* p, uint8_t v) {
    if (XFLAVOR == 3) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    else __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Team-workspace sizes of one 16-member team: crumb chunks [parity][k-group 32][replica 16][16 B] and n_lin bytes
// [parity][replica 16][member 16][wave 4].
constexpr int T16_FRAG_U64 = 2 * 32 * 16 * 2;   // uint64 units
constexpr int T16_NLIN_U32 = 2 * 16 * 64 / 4;   // uint32 units

// CORR24: the counted correction's product n_lin kcorr >> 8 as two 24-bit multiplies (the host's range plan proves it
// exact; int32 thresholds only). PROF: record the per-phase cycle counters of thread 0 (a.prof).
template <bool WIDE, int XFLAVOR, bool CORR24, bool PROF>
__device__ __forceinline__ void team16_body(const TeamArgs& a, int team, int c) {
    static_assert(!(WIDE && CORR24), "the 24-bit correction split is an int32-threshold form");
    constexpr int KR = 8, RT = 16, KG = 32, KBW = 8, MT = 2;
    constexpr uint32_t TAGM = 0x55555555u;
    constexpr bool PADH = !WIDE;               // int32 decisions: padding slots carry the field 2^30 (see anc_team.hpp)
    constexpr int32_t PAD_H = int32_t(1) << 30;
    using wide_t = typename std::conditional<WIDE, int64_t, int32_t>::type;

    __shared__ v4i XS[2][KG * RT];              // staged crumb chunks [parity][k-group * 16 + replica]
    __shared__ int32_t NLSUM[2][RT];            // staged n_lin of the previous step per replica [parity][replica]
    __shared__ int bar_cells[2];

    if (team >= a.n_teams) return;
    const int tid = threadIdx.x, w = tid >> 6, g = (tid >> 4) & 3, ml = tid & 15;
    const int lane = 16 * c + 4 * w + g;        // RNG lane of this MFMA lane's 8 slots
    const uint64_t t_limit = uint64_t(a.timeout_ticks);
    uint64_t t_wait = 0;
    bool bail = false;
    if (a.stall_test) {   // host-forced time-out: the bounded wait on a word nothing sets gives up, the workgroup exits
        const uint64_t t0 = wall_clock64();
        while (__hip_atomic_load(a.status + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
            if (wall_clock64() - t0 > t_limit) { raise_error(a.status, 1); break; }
            __builtin_amdgcn_s_sleep(8);
        }
        return;
    }

    uint64_t* xfrag = a.xfrag + size_t(team) * T16_FRAG_U64;
    uint32_t* xnlin = a.xnlin + size_t(team) * T16_NLIN_U32;
    __amdgpu_buffer_rsrc_t frag_rsrc = __builtin_amdgcn_make_buffer_rsrc(xfrag, 0, 0x7fffffff, 0x00020000);
    __amdgpu_buffer_rsrc_t nlin_rsrc = __builtin_amdgcn_make_buffer_rsrc(xnlin, 0, 0x7fffffff, 0x00020000);

    const bool prof = PROF && tid == 0;
    uint64_t pacc[6] = {0, 0, 0, 0, 0, 0};
    uint64_t ptick = prof ? __builtin_amdgcn_s_memtime() : 0;
    auto mark = [&](int i) {
        if (prof) {
            const uint64_t now = __builtin_amdgcn_s_memtime();
            pacc[i] += now - ptick;
            ptick = now;
        }
    };

    const int n_items = a.items ? a.n_items : (a.R + RT - 1) / RT;
    if (tid < 2) bar_cells[tid] = 0;
    int seq = 0;
    for (int item = team; item < n_items; item += a.n_teams, ++seq) {
    __syncthreads();   // the previous item's LDS reads are complete
    const int8_t* Jp = a.Jp;
    const int32_t* ibias = a.bias;
    int in_n = a.n, r0 = item * RT, cnt = min(RT, a.R - item * RT);
    if (a.items) {
        const TeamItem it = a.items[item];
        Jp = it.Jp; ibias = it.bias; in_n = it.n; r0 = it.r0; cnt = it.count;
    }
    const int ubase = seq * (a.S + 1);
    const int r = r0 + ml;
    const bool active = ml < cnt;

    // ---- A fragments: M-tile mt, load block kb (columns 256 kb .. 256 kb + 255), crumb position q. A lane (g, ml)
    // supplies row ml of the M-tile, i.e. round 4 mt + (ml & 3) of RNG lane 16 c + 4 w + (ml >> 2): J' row (x' order,
    // x' = 8 L + k) 128 c + 32 w + 8 (ml >> 2) + 4 mt + (ml & 3). Its 64 columns x' = 256 kb + 64 g + 4 B + q go to byte B
    // of fragment q (the order in which crumb position q of an exchange chunk unpacks).
    v4i A[MT][KBW][4];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int kb = 0; kb < KBW; ++kb) {
            const size_t row = size_t(128 * c + 32 * w + 8 * (ml >> 2) + 4 * mt + (ml & 3));
            const v4i* src = reinterpret_cast<const v4i*>(Jp + row * 2048 + size_t(256 * kb + 64 * g));
            const v4i w4[4] = {src[0], src[1], src[2], src[3]};
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int d = 0; d < 4; ++d) {
                    uint32_t v = 0;
#pragma unroll
                    for (int e = 0; e < 4; ++e) v |= ((uint32_t(w4[d][e]) >> (8 * q)) & 0xFFu) << (8 * e);
                    A[mt][kb][q][d] = int(v);
                }
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (mt == 0 || kb < KBW - 1) __asm__ volatile("" : "+a"(A[mt][kb][q]));   // to the AGPRs (see the step loop)
        }
    // ---- row sums of this lane's 8 rows (MFMA with an all-ones B): the offset correction of the field
    int32_t rowsum_k[KR];
    {
        const v4i ones = {0x01010101, 0x01010101, 0x01010101, 0x01010101};
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            v4i acc = {0, 0, 0, 0};
#pragma unroll
            for (int kb = 0; kb < KBW; ++kb)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[mt][kb][q], ones, acc, 0, 0, 0);
#pragma unroll
            for (int e = 0; e < 4; ++e) rowsum_k[4 * mt + e] = acc[e];
        }
    }

    Rng rng;
    uint32_t sbits = ~0u, spbits = ~0u;   // bit k: spin of slot k * 256 + lane is +1
    int32_t bias_k[KR];
    int32_t base0_k[KR], base1_k[KR];     // the field without J' N: b - (1 + tau) rowsum for tau = 0 and tau = 1
    int32_t nflips = 0;
    if (active) rng.seed(a.seed, a.trials ? a.trials[r] : a.trial0 + uint32_t(r), uint32_t(lane));
#pragma unroll
    for (int k = 0; k < KR; ++k) {   // initial spins (sca_ref_bias.hpp:31-36)
        const int x = k * LANES + lane;
        bias_k[k] = 0;
        if (active) {
            const uint32_t rv = rng.next();
            if (x < in_n && !(rv >> 31)) sbits &= ~(1u << k);
            if (x < in_n && ibias) bias_k[k] = ibias[x];
        }
        if (PADH && x >= in_n) bias_k[k] = PAD_H;
        base0_k[k] = bias_k[k] - rowsum_k[k];
        base1_k[k] = bias_k[k] - 2 * rowsum_k[k];
    }
    // this lane's 8 slots x' = 8 lane + k as one 16-bit crumb half (crumb k = tau + 2 s_up_k at bits 2k, 2k + 1) of
    // k-group 2 c + w / 2, replica ml, byte 8 (w & 1) + 2 g
    auto publish_spins = [&](int u, uint32_t sb) {
        if (!active) return;
        const uint32_t half = (tau_of(u) ? 0x5555u : 0u) | (spread_crumbs(sb & 0xFFu) << 1);
        uint16_t* base = reinterpret_cast<uint16_t*>(xfrag);
        const size_t byte = ((size_t(u & 1) * KG + 2 * c + (w >> 1)) * RT + ml) * 16 + 8 * (w & 1) + 2 * g;
        xstore2<XFLAVOR>(base + byte / 2, uint16_t(half));
    };
    publish_spins(ubase, sbits);
    int32_t nlin_prev = 0;
    int32_t H[KR];
    int32_t rr_k[KR] = {0, 0, 0, 0, 0, 0, 0, 0};   // lanes of replicas outside the item keep 0 (their results are unused)

    typedef const __attribute__((address_space(4))) int32_t* ctab32;
    typedef const __attribute__((address_space(4))) int64_t* ctab64;
    int64_t kc_n = 0;
    int32_t kk_n = 0, f4_n = 1, q_n = 0;
    auto load_tables = [&](int t) {
        const int tt = min(t, a.S - 1);
        kc_n = ((ctab64)a.kcorr)[tt]; kk_n = ((ctab32)a.kconst)[tt]; f4_n = ((ctab32)a.fourT)[tt]; q_n = ((ctab32)a.q)[tt];
    };
    int64_t kc_t;
    int32_t kk_t, f4_t, q_t;
    auto take_tables = [&]() { kc_t = kc_n; kk_t = kk_n; f4_t = f4_n; q_t = q_n; };
    load_tables(0);
    take_tables();

    // staging assignment: thread tid loads crumb chunks ci = tid and tid + 256 (ci = k-group * 16 + replica) and, in
    // wave 0, the n_lin quarter tid (replica tid / 4)
    const int ci0 = tid, ci1 = tid + 256;
    const bool need0 = (ci0 & 15) < cnt, need1 = (ci1 & 15) < cnt;
    const bool nl_thread = tid < 64 && (tid >> 2) < cnt;

    for (int t = 0; t <= a.S; ++t) {
        // J' fragments live in the accumulation registers (AGPRs), which the MFMA reads directly as its A operand:
        // the 256 registers of A do not compete with the step's working values for the VGPRs (an empty asm statement
        // with the AGPR constraint; no instruction is emitted)
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int kb = 0; kb < KBW; ++kb)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (mt == 0 || kb < KBW - 1) __asm__ volatile("" : "+a"(A[mt][kb][q]));
#pragma unroll
        for (int k = 0; k < KR; ++k) __asm__ volatile("" : "+v"(base0_k[k]), "+v"(base1_k[k]));   // both kept: one select per slot
        const int u = ubase + t;
        const int par = u & 1;
        const uint32_t tau = tau_of(u);
        const uint32_t want = tau ? TAGM : 0u;
        const uint32_t ntag = nlin_tag(u) * 0x01010101u;
        const bool want_nlin = t >= 1 && nl_thread;
        // ===== exchange: issue the loads, draw this step's thresholds while they are in flight, re-poll stale chunks
        v4i b0 = {0, 0, 0, 0}, b1 = {0, 0, 0, 0}, nb = {0, 0, 0, 0};
        auto spin_off = [&](int ci) { return (par * KG * RT + ci) * 16; };
        const int nl_off = (par * RT * 4 + tid) * 16;
        if (need0) b0 = xload16<XFLAVOR>(frag_rsrc, spin_off(ci0));
        if (need1) b1 = xload16<XFLAVOR>(frag_rsrc, spin_off(ci1));
        if (want_nlin) nb = xload16<XFLAVOR>(nlin_rsrc, nl_off);
        load_tables(t + 1);
        if (active && t < a.S) {
            const int32_t f_hi = f4_t >> 16, f_lo = f4_t & 0xFFFF;
#pragma unroll
            for (int k = 0; k < KR; ++k) {
                const int32_t av = int32_t(rng.next() >> 16) - 32768;
                rr_k[k] = av * f_hi + ((av * f_lo) >> 16);   // floor((u16 - 32768) fourT / 2^16), exact in int32
            }
        }
        auto cur_spin = [&](const v4i& b) {
            const uint32_t an = uint32_t(b[0]) & uint32_t(b[1]) & uint32_t(b[2]) & uint32_t(b[3]);
            const uint32_t orr = uint32_t(b[0]) | uint32_t(b[1]) | uint32_t(b[2]) | uint32_t(b[3]);
            return ((tau ? an : orr) & TAGM) == want;
        };
        auto cur_nlin = [&](const v4i& b) {
            return ((uint32_t(b[0]) & 0xC0C0C0C0u) == ntag) & ((uint32_t(b[1]) & 0xC0C0C0C0u) == ntag) &
                   ((uint32_t(b[2]) & 0xC0C0C0C0u) == ntag) & ((uint32_t(b[3]) & 0xC0C0C0C0u) == ntag);
        };
        bool ok0 = !need0 || cur_spin(b0), ok1 = !need1 || cur_spin(b1), okn = !want_nlin || cur_nlin(nb);
        int polls = 0;
        if (prof && !__all(ok0 & ok1 & okn)) pacc[5] += 1;   // steps whose first poll was stale (wave 0)
        while (!__all(ok0 & ok1 & okn)) {
            if ((++polls & 63) == 0) {
                const uint64_t now = wall_clock64();
                if (polls == 64) t_wait = now;
                if (aborted(a.status)) { bail = true; break; }
                if (now - t_wait > t_limit) { raise_error(a.status, 1); bail = true; break; }
            }
            __builtin_amdgcn_s_sleep(1);
            __asm__ volatile("" ::: "memory");   // the re-poll loads are issued anew (no reuse of the loads above)
            if (!ok0) b0 = xload16<XFLAVOR>(frag_rsrc, spin_off(ci0));
            if (!ok1) b1 = xload16<XFLAVOR>(frag_rsrc, spin_off(ci1));
            if (!okn) nb = xload16<XFLAVOR>(nlin_rsrc, nl_off);
            ok0 = ok0 || cur_spin(b0);
            ok1 = ok1 || cur_spin(b1);
            okn = okn || cur_nlin(nb);
        }
        XS[par][ci0] = b0;
        XS[par][ci1] = b1;
        if (tid < 64) {   // wave 0: the replica's 64 (member, wave) bytes, count in the low 6 bits, over its 4 lanes
            int32_t s = 0;
#pragma unroll
            for (int d = 0; d < 4; ++d) s = __builtin_amdgcn_udot4(uint32_t(nb[d]) & 0x3F3F3F3Fu, 0x01010101u, s, false);
            s = group_sum<4>(s);
            if ((tid & 3) == 0) NLSUM[par][tid >> 2] = s;
        }
        mark(0);
        if (barrier_any(bail, bar_cells[t & 1])) { bail = true; break; }
        mark(1);
        // ===== field: acc = J' N over all 2048 columns from the staged chunks. Four independent accumulator chains
        // (two M-tiles x even / odd crumb positions), so that consecutive MFMAs never wait for each other's results; the
        // crumb unpacking of block kb + 1 is written next to the MFMAs of block kb and left to the compiler's scheduler
        // (explicit one-MFMA-three-VALU sched_group_barrier groups measured 1.3 % slower per launch, and forcing all eight
        // LDS reads of the chunks to complete together 0.4 % slower). From the barrier to the decisions the code has no
        // branch, so that the work that does not need the field (n_lin, the per-slot decision constants) can issue beside
        // the MFMAs.
        v4i bs[KBW];   // this lane's eight chunks (the compiler places the LDS reads and their waits)
#pragma unroll
        for (int kb = 0; kb < KBW; ++kb) bs[kb] = XS[par][(4 * kb + g) * RT + ml];
        // n_lin of step t - 1 for replica ml, summed by wave 0 at staging (zero at t = 0 and for replicas outside the item)
        nlin_prev = NLSUM[par][ml];
        // int32 decisions of step t (sca_ref.hpp:119-133), constants per slot. With m = 0 for s = +1 and -1 for s = -1,
        // s h = (h ^ m) - m, so: flip <=> s h <= T <=> (h ^ m) + ~(T + m) < 0, and band <=> s h - lo in [0, w) <=>
        // (h ^ m) + (-m - lo) < w unsigned. The range plan proves |q| + |corr| + fourT / 2 + 2 < 2^31 for every reachable
        // n_lin (so every threshold below is exact in int32); with CORR24 the product n_lin kcorr >> 8 is split as
        // n_lin (kcorr >> 8) + ((n_lin (kcorr & 255)) >> 8), both 24-bit multiplies (the plan proves |kcorr >> 8| < 2^23
        // and n |kcorr >> 8| < 2^31). Padding slots carry the field 2^30: never a flip, never in the band.
        int32_t hb_k[KR], m_k[KR], fo_k[KR], bo_k[KR], bw_k[KR];
        if constexpr (PADH) {
            const int32_t corr = CORR24 ? __mul24(nlin_prev, int32_t(kc_t >> 8)) +
                                              int32_t(__umul24(uint32_t(nlin_prev), uint32_t(kc_t) & 0xFFu) >> 8) + kk_t
                                        : int32_t(((int64_t(nlin_prev) * kc_t) >> 8) + int64_t(kk_t));
            const int32_t twoT = f4_t >> 1;
            const int32_t c_same = q_t - corr, c_diff = q_t + corr;
            const int32_t c1d = c_diff + 1, c1s = c_same + 1;
            const int32_t lod = ((-twoT - c_diff) >> 16) + 1, los = ((-twoT - c_same) >> 16) + 1;
            const int32_t lod1 = lod - 1, los1 = los - 1;
            const int32_t hid = (twoT - c_diff - 1) >> 16, his = (twoT - c_same - 1) >> 16;
            const int32_t wd = hid >= lod ? hid - lod + 1 : 0, ws = his >= los ? his - los + 1 : 0;
            const uint32_t same_bits = ~(sbits ^ spbits);
#pragma unroll
            for (int k = 0; k < KR; ++k) {
                const bool same = (same_bits >> k) & 1;
                const int32_t nm = __builtin_amdgcn_sbfe(int32_t(sbits), k, 1);   // ~m: -1 for s = +1, 0 for s = -1
                const int32_t x = rr_k[k] - (same ? c1s : c1d);   // T = x >> 16
                hb_k[k] = tau ? base1_k[k] : base0_k[k];
                m_k[k] = ~nm;
                fo_k[k] = nm - (x >> 16);                 // ~(T + m) = -T - m - 1 = nm - T
                bo_k[k] = nm - (same ? los1 : lod1);      // -m - lo = nm + 1 - lo
                bw_k[k] = same ? ws : wd;
            }
        }
        // the chains start from the MFMA's inline zero (no accumulator registers to clear: 2 % per launch against
        // zeroed registers, now that the VALU issue bounds this phase)
        v4i acc2[MT][2] = {{{0, 0, 0, 0}, {0, 0, 0, 0}}, {{0, 0, 0, 0}, {0, 0, 0, 0}}};
        // crumb position q of every byte as an int8 B fragment: even positions shifted down (values 0..3), odd ones
        // masked in place (values 0, 4, 8, 12: the odd-position chain accumulates 4 J' N exactly and is divided by 4
        // at the end; the host admits instances with 12 x field bound < 2^31 only, so no partial sum can overflow)
        auto unpack = [&](const v4i& b, v4i (&bq)[4]) {
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                const uint32_t x = uint32_t(b[d]), y = x >> 4;
                bq[0][d] = int(x & 0x03030303u);
                bq[1][d] = int(x & 0x0C0C0C0Cu);
                bq[2][d] = int(y & 0x03030303u);
                bq[3][d] = int(y & 0x0C0C0C0Cu);
            }
        };
        v4i bq[2][4];
        unpack(bs[0], bq[0]);
#pragma unroll
        for (int kb = 0; kb < KBW; ++kb) {
            const int cur = kb & 1;
            if (kb + 1 < KBW) unpack(bs[kb + 1], bq[cur ^ 1]);
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
                    acc2[mt][q & 1] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[mt][kb][q], bq[cur][q], acc2[mt][q & 1], 0, 0, 0);
        }
        uint32_t fm = 0;
        int32_t nband = 0;
#pragma unroll
        for (int k = 0; k < KR; ++k) {
            const int32_t acc = acc2[k >> 2][0][k & 3] + (acc2[k >> 2][1][k & 3] >> 2);
            if constexpr (PADH) {
                H[k] = hb_k[k] + acc;
                const int32_t hm = H[k] ^ m_k[k];
                fm |= (uint32_t(hm + fo_k[k]) >> 31) << k;
                nband += uint32_t(hm + bo_k[k]) < uint32_t(bw_k[k]) ? 1 : 0;
            } else {
                H[k] = (tau ? base1_k[k] : base0_k[k]) + acc;
            }
        }
        // the decisions are formed here, before the branches below (otherwise the compiler sinks them, and the
        // constants they need, past the loop exit, away from the MFMAs they can issue beside)
        if constexpr (PADH) __asm__ volatile("" ::"v"(fm), "v"(nband));
        mark(2);
        if (t >= 1 && active && c == 0 && tid < 16 && a.nlin) a.nlin[size_t(r) * size_t(a.S) + (t - 1)] = nlin_prev;
        if (t == a.S) break;
        int32_t nlin_part = 0;
        if constexpr (PADH) {
            const uint32_t sold = sbits;
            sbits = sold ^ fm;
            nflips += __builtin_popcount(fm);
            nlin_part = nband;
            spbits = sold | ~((1u << KR) - 1u);
        } else if (active) {
            // ===== int64 decisions of step t (sca_ref.hpp:119-133)
            const int64_t corr = ((int64_t(nlin_prev) * kc_t) >> 8) + int64_t(kk_t);
            const int32_t twoT = f4_t >> 1;
            const int64_t c_same = int64_t(q_t) - corr, c_diff = int64_t(q_t) + corr;
            wide_t c1[2] = {wide_t(c_diff + 1), wide_t(c_same + 1)};
            wide_t lo[2] = {wide_t(((-int64_t(twoT) - c_diff) >> 16) + 1), wide_t(((-int64_t(twoT) - c_same) >> 16) + 1)};
            wide_t hi[2] = {wide_t((int64_t(twoT) - c_diff - 1) >> 16), wide_t((int64_t(twoT) - c_same - 1) >> 16)};
            const uint32_t sold = sbits;
            const uint32_t same_bits = ~(sold ^ spbits);
#pragma unroll
            for (int k = 0; k < KR; ++k) {
                const int x = k * LANES + lane;
                const bool s_up = (sold >> k) & 1;
                const int cls = (same_bits >> k) & 1;
                const int32_t v = s_up ? H[k] : -H[k];
                const bool flip = wide_t(v) <= ((wide_t(rr_k[k]) - c1[cls]) >> 16);
                const bool band = (wide_t(v) >= lo[cls]) & (wide_t(v) <= hi[cls]);
                const bool valid = x < in_n;
                nlin_part += int(valid & band);
                const bool fl = valid & flip;
                if (fl) sbits ^= (1u << k);
                nflips += fl ? 1 : 0;
            }
            spbits = sold | ~((1u << KR) - 1u);
        }
        take_tables();
        mark(3);
        // ===== publish: this wave's n_lin byte per replica (sum over the 4 lanes g of the replica), then the spins
        nlin_part = rows4_sum(nlin_part);
        if (active && g == 0) {
            uint8_t* nlb = reinterpret_cast<uint8_t*>(xnlin);
            xstore1<XFLAVOR>(nlb + (size_t((u + 1) & 1) * RT + ml) * 64 + 4 * c + w, uint8_t(nlin_tag(u + 1) | uint32_t(nlin_part)));
        }
        publish_spins(u + 1, sbits);
        mark(4);
    }
    // ===== outputs (sca_ref_bias.hpp:69-73): spins, sum s h, sum b s, flips, optional fields
    if (!bail) {
        int64_t sh = 0, bs = 0;
        if (active) {
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
        }
        const int64_t fl = rows4_sum64(active ? nflips : 0);
        sh = rows4_sum64(sh);
        bs = rows4_sum64(bs);
        if (active && g == 0) {
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

// Workgroup b runs on XCD b % 8; the 16 members of a team are the workgroups with equal b % 8 and equal (b / 8) / 16:
// teams 0..7 use XCD-local workgroups 0..15, teams 8..15 workgroups 16..31.
template <bool WIDE, int XFLAVOR, bool CORR24, bool PROF>
__global__ void __launch_bounds__(THREADS, 1) team16_kernel(TeamArgs a) {
    const int xcd = blockIdx.x % 8, idx = blockIdx.x / 8;
    team16_body<WIDE, XFLAVOR, CORR24, PROF>(a, (idx / 16) * 8 + xcd, idx % 16);
}

}  // namespace dev

// 16-member teams: n_teams <= 16 (two per XCD), 16-replica work items; NP = 2048, spin crumbs.
int launch_team16(const TeamArgs& a, bool wide, int xflavor, int cus, bool dry, hipStream_t st);

}  // namespace anc
