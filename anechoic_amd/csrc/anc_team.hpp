// The persistent team kernels (n <= 3072): the kernel templates and their launch helpers. anc_team_k32.hip,
// anc_team_k64.hip and anc_team_mixed.hip instantiate them (parallel compilation).
//
// Arithmetic contract (reference fpga/v80_sca/src/sca_ref_bias.hpp:25-75, generalised NP): see anc_engine.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "anc_device.hpp"
#include "anc_kernels.hpp"

namespace anc {
namespace dev {

// ---------------------------------------------------------------------------------------------------------- team kernel
// One team member's whole run. TEAM = workgroups per team (32: one XCD; 64: two XCDs), KR = rounds (the kernel's
// NP/256: 8 or 12), NT = 16-replica N-tiles per team (RT = 16 NT replicas per work item), WIDE = int64 decision thresholds,
// XFLAVOR = exchange load flavour; `team` and member index `c` come from the launching kernel's placement.
//
// Exchange modes (XM):
//   0 (option "exchange" 0): the flip matrix, D' = D + tau bytes (D in {-2, 0, 2}, tag tau in bit 0), accumulated over
//     the steps, P_w += J'_w D'_u (the incremental update h += J ds), with the tau contributions removed in the field
//     (tau_count rowsum);
//   1 (int8, option "exchange" 1): the spins, one nibble per slot, N = s + 1 + tau (tag in bit 0, s + 1 in bit 1); a
//     16-byte load holds 32 slots, unpacked by masks into two int8 B fragments (even and odd slots, matched by a byte
//     shuffle of the J' fragments at load), and every step computes the field afresh with int32 accumulators,
//     H = b + J'_w N_u - (1 + tau) rowsum: half the exchange bytes of mode 0;
//   3 (int8, the default, option "exchange" 2; 32-workgroup teams with 8 rounds): the spins as 2-bit crumbs, tag in
//     the low bit and s + 1 in the high bit; a 16-byte load holds 64 slots, unpacked by shifts and masks into the int8
//     B fragments of the four crumb positions (J' fragments byte-transposed to match), H as in mode 1: half the
//     exchange loads of mode 1.
// LDS of one team member, carved from the kernel's dynamic LDS, so that a kernel hosting both team shapes (each
// workgroup runs one) needs only the larger layout: the per-wave partial fields [parity][wave][replica][slot] (int32,
// double-buffered by step parity), the per-wave row sums of J', the step barrier cells, the single-anneal helper
// table and, in the 12-round kernels with 32-replica items, the J' fragments of every wave's last load block (the
// registers hold the other 240 of the 288; the second N-tile's fragments and accumulators need the rest).
template <int TEAM, int KR, int NT, int XM>
struct TeamLds {
    static constexpr int LPC = LANES / TEAM, RT = 16 * NT, SPC = LPC * KR, PPITCH = SPC + 4, MT = SPC / 16;
    static constexpr int AH = XM == 3 ? 4 : XM == 1 ? 2 : 1;
    static constexpr bool HELP = NT == 1 && KR == 8 && XM != 0;
    static constexpr bool ALDS = TEAM == 32 && KR == 12 && NT == 2;
    static constexpr size_t P_OFF = 0;
    static constexpr size_t RS_OFF = P_OFF + size_t(2) * 4 * RT * PPITCH * 4;
    static constexpr size_t BAR_OFF = RS_OFF + size_t(4) * SPC * 4;
    static constexpr size_t PL_OFF = BAR_OFF + 16;
    static constexpr size_t AL_OFF = (PL_OFF + size_t(HELP ? 8 * 28 : 4) * 4 + 15) / 16 * 16;
    static constexpr size_t BYTES = AL_OFF + (ALDS ? size_t(4) * MT * AH * 64 * 16 : 0);
};

template <int TEAM, int KR, int NT, bool WIDE, int XFLAVOR, int XMODE = 1>
__device__ __forceinline__ void team_body(const TeamArgs& a, int team, int c) {
    constexpr int LPC = LANES / TEAM;    // RNG lanes per member
    constexpr int RT = 16 * NT;          // replicas per item
    constexpr int SPC = LPC * KR;        // slots per member (64 for 32 x 8, 48 for 64 x 12)
    constexpr int MT = SPC / 16;         // M-tiles (and 16-slot k-groups) per member
    constexpr int XM = XMODE;            // exchange mode (above)
    constexpr int KPL = XM == 3 ? 64 : XM ? 32 : 16;   // slots per 16-byte exchange lane chunk (crumbs / nibbles / bytes)
    constexpr int KBW = 16 * KR / KPL;   // load blocks (4 lane chunks) per wave: each wave owns a quarter of the NP columns
    constexpr int AH = XM == 3 ? 4 : XM == 1 ? 2 : 1;   // int8 MFMA K-blocks per load block (mode 1: even and odd
                                                        // slots; mode 3: the four crumb positions)
    constexpr int KG = 256 * KR / KPL;   // exchange k-groups (lane chunks) in NP
    constexpr uint32_t TAGM = XM == 3 ? 0x55555555u : XM == 1 ? 0x11111111u : 0x01010101u;   // tag bits
    constexpr int NQ = TEAM / LPC;       // n_lin words gathered per decision thread
    constexpr int PPITCH = SPC + 4;      // LDS partial row pitch (int32), padded against bank conflicts
    static_assert(SPC % 16 == 0 && KR % 4 == 0 && THREADS >= RT * LPC, "unsupported team shape");
    // nibble exchange: 8 rounds (one word per thread), or 12 rounds with the int8 MFMA (three 16-bit
    // halves); 32 workgroups x 12 rounds hold 96 J' rows per member in registers, which fit with 16-replica items only
    static_assert(!XM || (TEAM == 32 && KR == 8) || (XM == 1 && KR == 12) || (TEAM == 32 && XM == 1 && KR == 4),
                  "nibble exchange shape");
    static_assert(XM != 3 || (TEAM == 32 && KR == 8), "crumb exchange: 32 workgroups x 8 rounds");
    static_assert(!(TEAM == 32 && KR == 12) || NT <= 2, "32 workgroups x 12 rounds: 16- or 32-replica work items");
    using acc_t = v4i;
    // rounds k >= kr (phantom slots, no draws) exist only in the 12-round kernels (NP(n) < 3072); with 8 rounds
    // kr = KR, and the per-round tests are left out
    constexpr bool PHANTOM = KR == 12;
    // draws per lane and step: KR, except for the 4-round kernels of n <= 1024 instances, whose slot layout is the
    // NP = 2048 one with rounds 4 .. 7 all padding: zero J' rows and columns, spins fixed at +1, outside n_lin and
    // the outputs; only their draws remain, and they are drawn and dropped (sca_ref_bias.hpp:31-36, :45-47)
    constexpr int DRAWS = KR == 4 ? 8 : KR;
    using wide_t = typename std::conditional<WIDE, int64_t, int32_t>::type;
    // int32 decisions with 8 rounds: padding slots carry the field PAD_H (bias 2^30, zero J' rows). The int32 range
    // plan bounds |h| <= 2^30 - 1 and every threshold (lo, hi, (r - c - 1) >> 16) by 2^15 + 1 in magnitude, so a
    // padding slot (spin +1, v = 2^30) never flips and is never in the band: the per-slot validity tests go away
    constexpr bool PADH = !WIDE && !PHANTOM;
    constexpr int32_t PAD_H = int32_t(1) << 30;

    using Lay = TeamLds<TEAM, KR, NT, XM>;
    extern __shared__ __align__(16) uint8_t team_lds[];
    auto& P = *reinterpret_cast<int32_t (*)[2][4][RT][PPITCH]>(team_lds + Lay::P_OFF);   // per-wave partial fields
    auto& RS = *reinterpret_cast<int32_t (*)[4][SPC]>(team_lds + Lay::RS_OFF);           // per-wave row sums of J'
    int* const bar_cells = reinterpret_cast<int*>(team_lds + Lay::BAR_OFF);
    // single-anneal items (NT = 1, 8 rounds, spin exchanges): the decisions of replica 0's 64 slots are spread over
    // wave 0's 64 lanes (lane 8 j + k decides round k of RNG lane j and assembles its field); per owner j: bias[0..7],
    // r[0..7], s, s', rowsum[0..7] at 20..27
    constexpr bool HELP = Lay::HELP;
    auto& PL = *reinterpret_cast<int32_t (*)[HELP ? 8 : 1][HELP ? 28 : 1]>(team_lds + Lay::PL_OFF);
    // 12 rounds, 32-replica items: the last load block's J' fragments of wave w, [w][mt][h][lane]
    constexpr bool ALDS = Lay::ALDS;
    v4i* const AL = reinterpret_cast<v4i*>(team_lds + Lay::AL_OFF);
    int32_t bias_h = 0, rowsum_h = 0;   // the helper lane's slot (single-anneal items)

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
    // n_lin words [parity][replica][member]: thread (n, j) reads the NQ members NQ j .. NQ j + NQ - 1 of its replica
    // with NQ / 4 sixteen-byte loads
    static_assert(NQ % 4 == 0, "n_lin words are read four at a time");
    auto nlin_off = [&](int u) { return int(((size_t(u & 1) * RT + min(n, RT - 1)) * TEAM + NQ * j) * 4); };
    const uint64_t t_limit = uint64_t(a.timeout_ticks);   // per wait, in wall-clock ticks
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
    if (tid < 2) bar_cells[tid] = 0;   // visible after the first item's barrier; step t uses cell t % 2
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

    // ---- coupling fragments of load block kb(i) = KBW w + (i + c) % KBW (the members of a team start their
    // K-quarters at different blocks, so that they do not all read the same exchange lines at the same moment;
    // integer sums: the order does not matter), row y' = SPC c + 16mt + ml, lane group g:
    //   mode 0: A[mt][i][0] = J'[y'][64 kb + 16g .. +16];
    //   mode 1: the 32 columns x' = 128 kb + 32g + p: A[mt][i][0] byte 4d + e = column p = 8d + 2e (even slots),
    //           A[mt][i][1] the odd ones p = 8d + 2e + 1 (the order of the unpacked nibbles).
    auto kb = [&](int i) { return KBW * w + (i + c) % KBW; };
    v4i A[MT][KBW][AH];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int i = 0; i < KBW; ++i) {
            const size_t row = size_t(SPC * c + 16 * mt + ml);
            if constexpr (XM == 0) {
                A[mt][i][0] = *reinterpret_cast<const v4i*>(Jp + row * size_t(a.np) + size_t(64 * kb(i) + 16 * g));
            } else if constexpr (XM == 3) {
                // the 64 columns x' = 256 kb + 64 g + 4 B + q: A[mt][i][q] byte B = column 4 B + q (the order in
                // which crumb position q of the exchange chunk's bytes unpacks)
                const v4i* src = reinterpret_cast<const v4i*>(Jp + row * size_t(a.np) + size_t(256 * kb(i) + 64 * g));
                const v4i w4[4] = {src[0], src[1], src[2], src[3]};
#pragma unroll
                for (int q = 0; q < 4; ++q)
#pragma unroll
                    for (int d = 0; d < 4; ++d) {
                        uint32_t v = 0;
#pragma unroll
                        for (int e = 0; e < 4; ++e) v |= ((uint32_t(w4[d][e]) >> (8 * q)) & 0xFFu) << (8 * e);
                        A[mt][i][q][d] = int(v);
                    }
            } else {
                const v4i* src = reinterpret_cast<const v4i*>(Jp + row * size_t(a.np) + size_t(128 * kb(i) + 32 * g));
                const v4i w0 = src[0], w1 = src[1];
                v4i f0, f1;
#pragma unroll
                for (int d = 0; d < 4; ++d) {
                    const uint32_t lo = uint32_t(d < 2 ? w0[2 * d] : w1[2 * d - 4]);
                    const uint32_t hi = uint32_t(d < 2 ? w0[2 * d + 1] : w1[2 * d - 3]);
                    f0[d] = int(__builtin_amdgcn_perm(hi, lo, 0x06040200u));
                    f1[d] = int(__builtin_amdgcn_perm(hi, lo, 0x07050301u));
                }
                if (ALDS && i == KBW - 1) {   // this wave's own LDS slots (only this wave reads them: no barrier)
                    AL[((w * MT + mt) * AH + 0) * 64 + L] = f0;
                    AL[((w * MT + mt) * AH + 1) * 64 + L] = f1;
                } else {
                    A[mt][i][0] = f0;
                    A[mt][i][1] = f1;
                }
                // 12 rounds: the 288 registers of A exceed the 256 accumulation registers; the first 240 go there
                // as soon as they are formed (an empty asm statement with the AGPR constraint; no instruction), so
                // that the prologue does not hold them in VGPRs next to its temporaries and spill (with 32-replica
                // items exactly the 240 that stay in registers)
                if constexpr (KR == 12 && TEAM == 32)
                    if (ALDS ? i < KBW - 1 : mt * KBW + i < 30) {
                        __asm__ volatile("" : "+a"(A[mt][i][0]));
                        __asm__ volatile("" : "+a"(A[mt][i][1]));
                    }
            }
        }
    // fragment (mt, i, h): from the registers, or for the 12-round 32-replica kernels' last load block from LDS
    auto afr = [&](int mt, int i, int h) -> v4i {
        if (ALDS && i == KBW - 1) return AL[((w * MT + mt) * AH + h) * 64 + L];
        return A[mt][i][h];
    };
    // ---- row sums of J' over this wave's K-quarter, by MFMA with an all-ones B (the tau / offset corrections of the
    // field)
    {
        const v4i ones = {0x01010101, 0x01010101, 0x01010101, 0x01010101};
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            v4i acc = {0, 0, 0, 0};
#pragma unroll
            for (int i = 0; i < KBW; ++i)
#pragma unroll
                for (int h = 0; h < AH; ++h) acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(afr(mt, i, h), ones, acc, 0, 0, 0);
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
            const uint32_t rv = (!PHANTOM || k < kr) ? rng.next() : 0u;   // rounds k >= kr are phantom slots: no draw
            if (x < in_n && !(rv >> 31)) sbits &= ~(1u << k);
            if (x < in_n && ibias) bias_k[k] = ibias[x];
        }
        if (PADH && x >= in_n) bias_k[k] = PAD_H;
    }
    if (active)
#pragma unroll
        for (int k = KR; k < DRAWS; ++k) rng.next();
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
    // mode 1: this thread's KR slots x' = SPC c + KR j + k (from an even x'), nibble k = s_k + 1 + tau: KR = 8 as one
    // word, KR = 12 as three 16-bit halves (even byte offsets, so that none crosses a 16-byte lane chunk); mode 3: the
    // 8 slots as 2-bit crumbs in one 16-bit half
    auto publish_spins = [&](int u, uint32_t sb) {
        if (!dthread) return;
        const uint32_t tau = tau_of(u);
        auto nib8 = [&](uint32_t bits8) { return (tau ? 0x11111111u : 0u) | (spread_nibbles(bits8 & 0xFFu) << 1); };
        auto chunk_byte = [&](int b) {   // byte b of the member-ordered nibble row -> offset in the exchange
            return ((((size_t(u & 1) * KG + b / 16) * NT + n / 16) * 16 + (n % 16)) * 16 + b % 16);
        };
        if constexpr (XM == 3) {
            // crumb k = tau + 2 s_up_k at bits 2k, 2k + 1 of the 16-bit half of the member-ordered crumb row that holds
            // slots SPC c + 8 j .. + 7 (4 slots per byte, byte offset (SPC c + 8 j) / 4)
            const int b0 = (SPC * c + KR * j) / 4;
            const uint32_t word = (tau ? 0x5555u : 0u) | (spread_crumbs(sb & 0xFFu) << 1);
            uint16_t* base = reinterpret_cast<uint16_t*>(a.xfrag + team_frag);
            xstore2<XFLAVOR>(base + chunk_byte(b0) / 2, uint16_t(word));
            return;
        }
        const int b0 = (SPC * c + KR * j) / 2;
        if constexpr (KR == 8) {
            xstore4<XFLAVOR>(reinterpret_cast<uint32_t*>(a.xfrag + team_frag) + chunk_byte(b0) / 4, nib8(sb));
        } else if constexpr (KR == 4) {
            xstore2<XFLAVOR>(reinterpret_cast<uint16_t*>(a.xfrag + team_frag) + chunk_byte(b0) / 2, uint16_t(nib8(sb)));
        } else {
            const uint32_t lo = nib8(sb), hi = nib8(sb >> 8) & 0xFFFFu;   // nibbles 0..7 and 8..11
            uint16_t* base = reinterpret_cast<uint16_t*>(a.xfrag + team_frag);
            xstore2<XFLAVOR>(base + chunk_byte(b0) / 2, uint16_t(lo));
            xstore2<XFLAVOR>(base + chunk_byte(b0 + 2) / 2, uint16_t(lo >> 16));
            xstore2<XFLAVOR>(base + chunk_byte(b0 + 4) / 2, uint16_t(hi));
        }
    };
    if constexpr (XM != 0) {
        publish_spins(ubase, sbits);
    } else {
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
    if constexpr (HELP) {
        if (cnt == 1 && tid < 64) {   // wave 0: the owners' per-item constants to their helper lanes
            if (active)
#pragma unroll
                for (int k = 0; k < KR; ++k) { PL[j][k] = bias_k[k]; PL[j][20 + k] = rowsum_k[k]; }
            __builtin_amdgcn_wave_barrier();
            bias_h = PL[tid >> 3][tid & 7];
            rowsum_h = PL[tid >> 3][20 + (tid & 7)];
        }
    }

    // mode 0: P_w accumulates over all steps; modes 1 and 3: P_w of the current step only (`fresh` = the step's
    // first load block starts from zero)
    acc_t acc[MT][NT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = acc_t{0, 0, 0, 0};
    // the MFMAs of load block i with its exchange chunks b[nt]
    auto mma_block = [&](int i, const v4i (&b)[NT], bool fresh) {
        if constexpr (XM == 1) {
            v4i lo[NT], hi[NT];   // even and odd slots of the 32 nibbles, one byte each
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
#pragma unroll
                for (int d = 0; d < 4; ++d) {
                    lo[nt][d] = int(uint32_t(b[nt][d]) & 0x0F0F0F0Fu);
                    hi[nt][d] = int((uint32_t(b[nt][d]) >> 4) & 0x0F0F0F0Fu);
                }
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                const v4i a0 = afr(mt, i, 0), a1 = afr(mt, i, 1);
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) {
                    acc[mt][nt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a0, lo[nt], fresh ? v4i{0, 0, 0, 0} : acc[mt][nt], 0, 0, 0);
                    acc[mt][nt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a1, hi[nt], acc[mt][nt], 0, 0, 0);
                }
            }
            return;
        }
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            if constexpr (XM == 3) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {   // crumb position q of the 16 bytes: the N of slots 4 B + q
                    v4i bq;
#pragma unroll
                    for (int d = 0; d < 4; ++d) bq[d] = int((uint32_t(b[nt][d]) >> (2 * q)) & 0x03030303u);
#pragma unroll
                    for (int mt = 0; mt < MT; ++mt)
                        acc[mt][nt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[mt][i][q], bq, (fresh && q == 0) ? v4i{0, 0, 0, 0} : acc[mt][nt], 0, 0, 0);
                }
            } else if constexpr (XM == 1) {
                // (handled below: the A fragments are shared by the N-tiles)
            } else {
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
                    acc[mt][nt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[mt][i][0], b[nt], acc[mt][nt], 0, 0, 0);
            }
        }
    };

    int32_t nlin_prev = 0;
    int32_t H[KR];
    int32_t rr_k[KR];

    // step tables, loaded one step ahead through the scalar cache (the constant address space gives scalar loads:
    // no vector memory instruction, which the CU's four waves would otherwise queue behind the exchange loads);
    // take_tables() moves the prefetched ones into this step's registers at the end of the step's decisions
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
    for (int t = 0; t <= a.S; ++t) {
        const int u = ubase + t;   // tag of the D' fragments consumed this iteration
        const bool want_nlin = t >= 1 && dthread;
        // n_lin words of step t-1 (published just before the D_u fragments): first read right after the first
        // fragment poll, stale ones re-read after every fragment re-poll, so their latency overlaps the exchange wait
        // and a word is read no earlier than the fragments it travels with
        uint32_t nv[NQ];
        const uint32_t nwant = uint32_t(u + 1) & 0xFFFFFu;
        auto nlin_stale = [&]() {
            bool st_ = false;
#pragma unroll
            for (int q4 = 0; q4 < NQ; ++q4) st_ |= (nv[q4] >> 12) != nwant;
            return st_;
        };
        auto nlin_load = [&]() {
#pragma unroll
            for (int q = 0; q < NQ / 4; ++q) {
                const v4i w4 = xload16<XFLAVOR>(nlin_rsrc, nlin_off(u) + 16 * q);
#pragma unroll
                for (int e = 0; e < 4; ++e) nv[4 * q + e] = uint32_t(w4[e]);
            }
        };
        auto nlin_reload = [&]() {
            if (nlin_stale()) nlin_load();
        };
        // ===== MFMA phase: P_w += J'_w D'_u (mode 0), P_w = J'_w N_u (modes 1 and 3)
        {
            const uint32_t tau = tau_of(u);
            const uint32_t want = tau ? TAGM : 0u;
            auto frag_off = [&](int i, int nt) {   // B lane (g, ml) = exchange k-group 4 kb + g, replica ml
                return int((((size_t(u & 1) * KG + 4 * kb(i) + g) * NT + nt) * 16 + ml) * 16);
            };
            auto current = [&](const v4i (&b)[NT]) {
                uint32_t all_and = 0xFFFFFFFFu, all_or = 0;
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
#pragma unroll
                    for (int e = 0; e < 4; ++e) { all_and &= uint32_t(b[nt][e]); all_or |= uint32_t(b[nt][e]); }
                return __all(((tau ? all_and : all_or) & TAGM) == want);
            };
            // first poll: every fragment load issued back to back, validated and consumed in issue order
            v4i B[KBW][NT];
#pragma unroll
            for (int i = 0; i < KBW; ++i)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) B[i][nt] = xload16<XFLAVOR>(frag_rsrc, frag_off(i, nt));
            if (want_nlin) nlin_load();
            load_tables(t + 1);
            // while the loads are in flight: this step's random thresholds r_k = floor((u16 - 32768) fourT / 2^16)
            // (they depend only on the lane's stream and the table, not on the exchanged data)
            if (active && t < a.S) {
                const int32_t f_hi = f4_t >> 16, f_lo = f4_t & 0xFFFF;
#pragma unroll
                for (int k = 0; k < KR; ++k) {
                    if (PHANTOM && k >= kr) break;   // phantom rounds draw nothing
                    const int32_t av = int32_t(rng.next() >> 16) - 32768;
                    rr_k[k] = av * f_hi + ((av * f_lo) >> 16);   // exact: av fourT = av f_hi 2^16 + av f_lo
                }
#pragma unroll
                for (int k = KR; k < DRAWS; ++k) rng.next();
            }
            uint32_t pending = 0;
#pragma unroll
            for (int i = 0; i < KBW; ++i) {
                if (current(B[i])) {
                    mma_block(i, B[i], XM != 0 && i == 0);
                } else {
                    pending |= 1u << i;
                    if (XM != 0 && i == 0)
#pragma unroll
                        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                            for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = acc_t{0, 0, 0, 0};
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
                __asm__ volatile("" ::: "memory");   // the re-poll loads are issued anew
#pragma unroll
                for (int i = 0; i < KBW; ++i)
                    if ((pending >> i) & 1u)
#pragma unroll
                        for (int nt = 0; nt < NT; ++nt) B[i][nt] = xload16<XFLAVOR>(frag_rsrc, frag_off(i, nt));
                if (want_nlin) nlin_reload();
#pragma unroll
                for (int i = 0; i < KBW; ++i) {
                    if (!((pending >> i) & 1u)) continue;
                    if (current(B[i])) {
                        mma_block(i, B[i], false);
                        pending &= ~(1u << i);
                    }
                }
            }
        }
        mark(0);
        // ===== n_lin of step t-1: team total for replica n (the words are usually current by now)
        if (want_nlin && !bail) {
            int polls = 0;
            while (nlin_stale()) {
                if ((++polls & 63) == 0) {
                    const uint64_t now = wall_clock64();
                    if (polls == 64) t_wait = now;
                    if (aborted(a.status)) { bail = true; break; }
                    if (now - t_wait > t_limit) { raise_error(a.status, 1); bail = true; break; }
                }
                __builtin_amdgcn_s_sleep(1);
                __asm__ volatile("" ::: "memory");   // the re-poll loads are issued anew
                nlin_reload();
            }
            int32_t part = 0;
#pragma unroll
            for (int q4 = 0; q4 < NQ; ++q4) part += int32_t(nv[q4] & 0xFFFu);
            part = group_sum<LPC>(part);
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
                    *reinterpret_cast<v4i*>(&P[buf][w][16 * nt + ml][16 * mt + 4 * g]) = __builtin_bit_cast(v4i, acc[mt][nt]);
        }
        if (barrier_any(bail, bar_cells[t & 1])) { bail = true; break; }
        mark(2);
        // single-anneal items: the helper lanes assemble their own fields (the owners' H is needed at t = S only)
        const bool helper_fields = HELP && cnt == 1 && tid < 64 && t < a.S;
        if (dthread && !helper_fields) {
            const int buf = t & 1;
            // acc = J'(s(0) - 1) + sum_u J' D_u + c rowsum, c = number of exchange steps so far with tau = 1;
            // with tc = c - 1: H = bias + acc - tc rowsum
            if constexpr (XM == 1 || XM == 3) {
                // acc = J' (s + 1 + tau) over the whole row: H = b + acc - (1 + tau) rowsum
                const int32_t off = 1 + int32_t(tau_of(u));
#pragma unroll
                for (int k = 0; k < KR; ++k) {
                    const int m = KR * j + k;
                    H[k] = bias_k[k] + P[buf][0][n][m] + P[buf][1][n][m] + P[buf][2][n][m] + P[buf][3][n][m] - off * rowsum_k[k];
                }
            } else {
                const int tc = tau_count(u) - tau_count(ubase - 1) - 1;
#pragma unroll
                for (int k = 0; k < KR; ++k) {
                    const int m = KR * j + k;
                    H[k] = bias_k[k] + P[buf][0][n][m] + P[buf][1][n][m] + P[buf][2][n][m] + P[buf][3][n][m] - tc * rowsum_k[k];
                }
            }
        }
        mark(3);
        if (t == a.S) break;
        if (!dthread) continue;
        // ===== decision phase of step t
        int d[KR];
        int32_t nlin_part = 0;
        bool helped = false;
        if constexpr (HELP) {
            if (cnt == 1 && tid < 64) {   // wave 0 holds the 8 owners (tid < 8) and serves as their 64 helpers
                helped = true;
                if (active) {
#pragma unroll
                    for (int k = 0; k < KR; ++k) PL[j][8 + k] = rr_k[k];
                    PL[j][16] = int32_t(sbits);
                    PL[j][17] = int32_t(spbits);
                }
                __builtin_amdgcn_wave_barrier();   // the wave's LDS writes precede its reads (one wave, in order)
                const int jo = tid >> 3, ko = tid & 7;
                const int buf = t & 1, m = KR * jo + ko;
                const int32_t hv = bias_h + P[buf][0][0][m] + P[buf][1][0][m] + P[buf][2][0][m] + P[buf][3][0][m] -
                                   (1 + int32_t(tau_of(u))) * rowsum_h;
                const int32_t rr = PL[jo][8 + ko];
                const uint32_t so = uint32_t(PL[jo][16]), sp = uint32_t(PL[jo][17]);
                const int32_t nl = __builtin_amdgcn_readfirstlane(nlin_prev);   // replica 0's team total
                const int64_t corr = ((int64_t(nl) * kc_t) >> 8) + int64_t(kk_t);
                const int32_t twoT = f4_t >> 1;
                const int64_t c_same = int64_t(q_t) - corr, c_diff = int64_t(q_t) + corr;
                const bool s_up = (so >> ko) & 1;
                const bool same = !(((so ^ sp) >> ko) & 1);
                const int64_t cc = same ? c_same : c_diff;
                const wide_t c1 = wide_t(cc + 1);
                const wide_t lo = wide_t(((-int64_t(twoT) - cc) >> 16) + 1), hi = wide_t((int64_t(twoT) - cc - 1) >> 16);
                const int32_t v = s_up ? hv : -hv;
                const bool valid = ko * LANES + LPC * c + jo < in_n;
                const bool flip = valid & (wide_t(v) <= ((wide_t(rr) - c1) >> 16));
                const bool band = valid & (wide_t(v) >= lo) & (wide_t(v) <= hi);
                const uint64_t fm = __ballot(flip), bm = __ballot(band);
                if (active) {
                    const uint32_t nf = uint32_t(fm >> (8 * j)) & 0xFFu;
                    sbits = uint32_t(PL[j][16]) ^ nf;
                    nflips += __builtin_popcount(nf);
                    nlin_part = __builtin_popcount(uint32_t(bm >> (8 * j)) & 0xFFu);
                    spbits = uint32_t(PL[j][16]) | ~((1u << KR) - 1u);
                }
#pragma unroll
                for (int k = 0; k < KR; ++k) d[k] = 0;
            }
        }
        if (helped) {
        } else if (PADH && active) {
            const int64_t corr = ((int64_t(nlin_prev) * kc_t) >> 8) + int64_t(kk_t);
            const int32_t twoT = f4_t >> 1;
            const int64_t c_same = int64_t(q_t) - corr, c_diff = int64_t(q_t) + corr;
            // flip <=> v <= (r - (c+1)) >> 16 ; band <=> lo <= v <= hi <=> uint32(v - lo) < width, width = hi - lo + 1
            // (0 for an empty band); all in int32 by the range plan
            const int32_t c1d = int32_t(c_diff + 1), c1s = int32_t(c_same + 1);
            const int32_t lod = int32_t(((-int64_t(twoT) - c_diff) >> 16) + 1), los = int32_t(((-int64_t(twoT) - c_same) >> 16) + 1);
            const int32_t hid = int32_t((int64_t(twoT) - c_diff - 1) >> 16), his = int32_t((int64_t(twoT) - c_same - 1) >> 16);
            const uint32_t wd = hid >= lod ? uint32_t(hid - lod + 1) : 0u, ws = his >= los ? uint32_t(his - los + 1) : 0u;
            const uint32_t sold = sbits;
            const uint32_t same_bits = ~(sold ^ spbits);
            uint32_t fm = 0, bm = 0;
#pragma unroll
            for (int k = 0; k < KR; ++k) {
                const bool s_up = (sold >> k) & 1;
                const bool same = (same_bits >> k) & 1;
                const int32_t v = s_up ? H[k] : -H[k];
                fm |= uint32_t(v <= ((rr_k[k] - (same ? c1s : c1d)) >> 16)) << k;
                bm |= uint32_t((uint32_t(v) - uint32_t(same ? los : lod)) < (same ? ws : wd)) << k;
            }
            sbits = sold ^ fm;
            nflips += __builtin_popcount(fm);
            nlin_part = __builtin_popcount(bm);
            spbits = sold | ~((1u << KR) - 1u);
#pragma unroll
            for (int k = 0; k < KR; ++k) d[k] = ((fm >> k) & 1) ? (((sold >> k) & 1) ? -2 : 2) : 0;
        } else if (active) {
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
        take_tables();
        mark(4);
        // the n_lin partial goes out first, so that it is visible no later than the fragments it travels with
        nlin_part = group_sum<LPC>(nlin_part);
        if (j == 0) {
            uint32_t* dst = a.xnlin + team_nlin + (size_t((u + 1) & 1) * RT + n) * TEAM + c;
            xstore4<XFLAVOR>(dst, (((uint32_t(u + 2)) & 0xFFFFFu) << 12) | uint32_t(nlin_part));
        }
        if constexpr (XM != 0) publish_spins(u + 1, sbits);
        else publish(u + 1, d);
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
template <int TEAM, int KR, int NT, bool WIDE, int XFLAVOR, int XMODE = 1>
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
    team_body<TEAM, KR, NT, WIDE, XFLAVOR, XMODE>(a, team, c);
}

// Both team shapes in one persistent launch (instances of uneven size): one workgroup per CU, grid 8 XCDs x 32.
// XCDs 0 .. xcd2-1 each host one 32-workgroup team of the NP = 2048 items (a2, exchange flavour XFLAVOR). The
// NP = 2304..3072 items (a3) run, with the spin exchanges (XMODE 1, 3), on one 32-workgroup 12-round team per
// remaining XCD (16 NT3-replica items, flavour XFLAVOR); with the flip exchange (XMODE 0), on 64-workgroup teams of
// two remaining XCDs each (16 NT-replica items, device-scope exchange). xcd2 = 8 (spin exchanges): every XCD's team
// runs its share of the NP 2048 items and then its share of the NP 3072 items (the two classes' exchange areas are
// separate, so the second body starts on a zeroed area and its step tags may restart).
// rot (xcd2 = 8): XCD x runs team (x + rot) mod 8 of the NP 2048 items and team (x + (rot >> 3)) mod 8 of the NP 3072
// items, so that the teams with an extra item of one class (items go round-robin from team 0) land on the XCDs with
// fewer items of the classes before it.
template <int NT, int NT3, bool WIDE, int XFLAVOR, int XMODE>
__global__ void __launch_bounds__(THREADS, 1) team_kernel_mixed(TeamArgs a1, TeamArgs a2, TeamArgs a3, int xcd2, int rot) {
    const int xcd = blockIdx.x % 8, idx = blockIdx.x / 8;
    if (idx >= 32) return;
    if constexpr (XMODE != 0) {
        // the 4-round items of n <= 1024 instances (a1) run, in turn, before the NP 2048 items on the same teams
        if (xcd < xcd2) {
            team_body<32, 4, NT, WIDE, XFLAVOR, 1>(a1, xcd, idx);
            __syncthreads();   // the bodies' LDS layouts overlap
        }
        if (xcd2 == 8) {
            team_body<32, 8, NT, WIDE, XFLAVOR, XMODE>(a2, (xcd + rot) & 7, idx);
            __syncthreads();
            team_body<32, 12, NT3, WIDE, XFLAVOR, 1>(a3, (xcd + (rot >> 3)) & 7, idx);
            return;
        }
    }
    if (xcd < xcd2) team_body<32, 8, NT, WIDE, XFLAVOR, XMODE>(a2, xcd, idx);
    else if constexpr (XMODE != 0) team_body<32, 12, NT3, WIDE, XFLAVOR, 1>(a3, xcd - xcd2, idx);
    else team_body<64, 12, NT, WIDE, 1, XMODE>(a3, (xcd - xcd2) / 2, ((xcd - xcd2) % 2) * 32 + idx);
}
// dynamic LDS of the mixed kernel: the largest of its bodies' layouts
template <int NT, int NT3, int XMODE>
constexpr size_t mixed_lds_bytes() {
    const size_t b1 = XMODE != 0 ? TeamLds<32, 4, NT, 1>::BYTES : 0;
    const size_t b2 = TeamLds<32, 8, NT, XMODE>::BYTES;
    const size_t b3 = XMODE != 0 ? TeamLds<32, 12, NT3, 1>::BYTES : TeamLds<64, 12, NT, XMODE>::BYTES;
    const size_t b = b2 > b3 ? b2 : b3;
    return b1 > b ? b1 : b;
}

}  // namespace dev

// Dynamic LDS of a team launch: at least `need` bytes (the team body's LDS layout), and enough to put the workgroup
// above half of the CU's 160 KiB LDS, so that no CU hosts two team members; and a check that one such workgroup fits
// a CU.
inline int member_pad(const void* kern, size_t need, size_t& pad) {
    hipFuncAttributes fa;
    if (hipFuncGetAttributes(&fa, kern) != hipSuccess) return -1;
    const size_t half = 80 * 1024 + 1024;
    pad = fa.sharedSizeBytes < half ? half - fa.sharedSizeBytes : 0;
    if (pad < need) pad = need;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, dev::THREADS, pad) != hipSuccess) return -1;
    return per_cu >= 1 ? 0 : -2;
}

// member_pad, cached per kernel (the first call for a kernel must not be under stream capture)
inline int cached_pad(const void* kern, size_t& pad, size_t need = 0) {
    struct Entry {
        const void* kern;
        int rc;
        size_t pad;
    };
    static Entry cache[128];
    static int count = 0;
    for (int i = 0; i < count; ++i)
        if (cache[i].kern == kern) {
            pad = cache[i].pad;
            return cache[i].rc;
        }
    const int rc = member_pad(kern, need, pad);
    if (count < 128) cache[count++] = Entry{kern, rc, pad};
    return rc;
}

template <int TEAM, int KR, int NT, bool WIDE, int XF, int XMODE = 1>
inline int launch_team_t(const TeamArgs& a, int cus, bool dry, hipStream_t st) {
    const auto kern = dev::team_kernel<TEAM, KR, NT, WIDE, XF, XMODE>;
    const int grid = TEAM == 32 ? ((a.n_teams + 7) / 8) * 8 * TEAM : a.n_teams * TEAM;
    size_t pad = 0;
    if (int rc = cached_pad((const void*)kern, pad, dev::TeamLds<TEAM, KR, NT, XMODE>::BYTES)) return rc;
    if (a.n_teams * TEAM > cus || grid > cus) return -2;
    if (dry) return 0;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(dev::THREADS), pad, st, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int NT, int NT3, bool WIDE, int XF, int XMODE = 1>
inline int launch_mixed_t(const TeamArgs& a1, const TeamArgs& a2, const TeamArgs& a3, int xcd2, int rot, int cus, bool dry,
                          hipStream_t st) {
    const auto kern = dev::team_kernel_mixed<NT, NT3, WIDE, XF, XMODE>;
    size_t pad = 0;
    if (int rc = cached_pad((const void*)kern, pad, dev::mixed_lds_bytes<NT, NT3, XMODE>())) return rc;
    if (cus < 256 || (XMODE == 0 && a1.n_teams > 0) || a1.n_teams > (xcd2 == 8 ? 8 : xcd2)) return -2;
    if (xcd2 == 8 ? XMODE == 0 || a2.n_teams > 8 || a3.n_teams > 8 : a2.n_teams > xcd2 || (XMODE != 0 ? 1 : 2) * a3.n_teams > 8 - xcd2)
        return -2;
    if (dry) return 0;
    hipLaunchKernelGGL(kern, dim3(256), dim3(dev::THREADS), pad, st, a1, a2, a3, xcd2, rot);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Per-translation-unit entry points.
int launch_team32(const TeamArgs& a, int nt, bool wide, int xflavor, int cus, bool dry, hipStream_t st);
// NP = 3072 teams: 32 workgroups x 12 rounds with the spin exchange (nt 1 or 2), 64 workgroups with the flip one
int launch_team3072(const TeamArgs& a, int nt, bool wide, int xflavor, int cus, bool dry, hipStream_t st);
// the flip-matrix exchange (a.xmode == 0) of the three launches
int launch_team32_flip(const TeamArgs& a, int nt, bool wide, int xflavor, int cus, bool dry, hipStream_t st);
int launch_team32_crumb(const TeamArgs& a, int nt, bool wide, int xflavor, int cus, bool dry, hipStream_t st);
int launch_team_mixed_crumb(const TeamArgs& a1, const TeamArgs& a2, const TeamArgs& a3, int xcd2, int rot, int nt, int nt3,
                            bool wide, int xflavor, int cus, bool dry, hipStream_t st);
int launch_team64_flip(const TeamArgs& a, int nt, bool wide, int cus, bool dry, hipStream_t st);
int launch_team_mixed_flip(const TeamArgs& a1, const TeamArgs& a2, const TeamArgs& a3, int xcd2, int nt, bool wide, int xflavor,
                           int cus, bool dry, hipStream_t st);

}  // namespace anc
