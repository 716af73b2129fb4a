// Large instances (NP(n) > 3072, up to 16384 spins): the field H = b + J s as a streamed int8 MFMA GEMM every step.
// The kernel template; anc_large_{k,c}{16,32,64}.hip instantiate it per kernel slot count and correction mode
// (parallel compilation).
//
// Same per-anneal arithmetic as the team kernels (fpga/v80_sca/src/sca_ref_bias.hpp:25-75 with NP = NP(n)). With the
// counted correction (or an n_lin trace) the band count of every replica is reduced over the whole instance every
// step through device-scope 64-bit counters (see ncnt below); the temperature-scaled correction skips it.
//
// Layout (kernel slot count npk = 256 KR, KR in {16, 32, 64}; F = 64 / KR RNG lanes per lane group g):
//   * 256 workgroups, one per CU: KR row tiles x (256 / KR) replica tiles of RPT = 128 or 64 replicas (a replica tile
//     is a work item; R > RPT (256 / KR) runs as several rounds of items; 64-replica items keep every team busy when
//     R is small, at twice the J' stream per replica). Row tile c holds 256 permuted rows
//     x' = 256 c + rho, rho = 16 mt + 4 g + e (mt 0..15, g 0..3, e 0..3): RNG lane 4F c + F g + (j / KR), round
//     j % KR with j = 4 mt + e, i.e. slot x = 256 (j % KR) + lane.
//   * Four waves in a 2 x 2 arrangement: wave (wr, wc) handles rows 128 wr .. 128 wr + 127 (M-tiles 8 wr ..
//     8 wr + 7) of its row tile and replicas RPT/2 wc .. RPT/2 wc + RPT/2 - 1 (NTW = RPT / 32 MFMA N-tiles). Its
//     accumulators (8 x NTW tiles, 32 NTW int32 per lane) hold the field H of those rows and replicas across all
//     steps; element (mt, e) of lane (g, ml) is row rho, so a lane owns rounds of whole RNG streams for its NTW
//     replicas and decides them straight
//     from its accumulators (at KR = 64 the two row halves split one 64-round stream: each wave draws all 64
//     rounds and keeps its 32).
//   * Workgroup b runs on XCD b % 8; row tile c lives on XCD c % 8, so the replica tiles sharing a row tile share it
//     through that XCD's L2.
//   * J' (npk^2 int8, 256 MiB at npk = 16384) is stored fragment-major, [row tile][K-block of 64][mt][lane][16 B]:
//     every 128-column stage (one K-pair) of a row tile is 32 KiB of contiguous bytes, copied into a ring of three
//     LDS slots by direct-to-LDS loads two stages ahead of its use (the first two stages of a step during the previous
//     step's decisions) and read by the waves as MFMA A operands; within a K-pair the columns follow the nibble order
//     of the exchange (anc_large.hip, permute_large). A reads from LDS run four MFMA groups ahead.
//   * The spins are exchanged as nibbles N = s + 1 + tau (tau in bit 0 of every nibble, as in the team kernels)
//     through a global buffer [team][parity][replica half][K-pair][n-tile][lane][16 B]: lane (g, ml) of a producer
//     wave holds exactly the 32 rows of one 16-byte lane chunk per replica and stores it with one device-scope store.
//     The 8 chunks-of-64-lanes of a stage (2 replica halves x 4 N-tiles) are copied with device-scope direct-to-LDS
//     loads (lane L's 16 bytes land at slot + 16 L) into a ring of three LDS slots, two stages ahead of their use,
//     2 per wave; the two row-half waves of a replica half read the same copy, validate it, and unpack each chunk by
//     masks into the int8 B fragments of the K-pair's even and odd nibbles. While a chunk is not yet current the
//     wave copies its half's chunks again. Every nibble carries its validity, so no barrier crosses workgroups.
//   * Every step recomputes the field: the accumulators start at zero and collect J' N over all K-pairs, and
//     H = b + acc - (1 + tau) rowsum(J').
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "anc_device.hpp"
#include "anc_kernels.hpp"

namespace anc {
namespace dev {

// PHANTOM: kr < KR possible (rounds k >= kr of the kernel's layout draw nothing). COUNTED: the band counts are
// reduced over the instance every step (counted correction or n_lin trace).
template <int KR, bool WIDE, bool PHANTOM, bool COUNTED, int RPT>
__global__ void __launch_bounds__(THREADS, 1) large_kernel(LargeArgs a) {
    constexpr int F = 64 / KR;          // RNG lanes per row-tile lane group g
    constexpr int NRT = KR;             // row tiles = workgroups per team
    constexpr int NK = 4 * KR;          // K-blocks of 64 (npk / 64)
    constexpr int NST = NK / 2;         // LDS stages of two K-blocks
    constexpr int PER = NRT / 8;        // row tiles per XCD
    static_assert(RPT == 128 || RPT == 64, "replicas per work item");
    constexpr int NTW = RPT / 32;       // MFMA N-tiles per wave (16 replicas each): a wave handles half the item's replicas
    constexpr int HALF = RPT / 2;       // replicas per replica half
    constexpr int WC_XD = NK * NTW * 512;   // bytes of one replica half's spin chunks per parity (NK / 2 K-pairs x NTW KiB)
    constexpr int RPW = NTW / 2;        // B chunk rows each wave copies per stage (2 NTW rows per stage, 4 waves)
    // s_waitcnt immediates (lgkmcnt(0), expcnt untouched): at most this stage's copies (RPW B + 8 A) in flight, at
    // most RPW (the prologue's stage-1 B copies), none
    constexpr int WAIT_STAGE = 0x0070 | (8 + RPW), WAIT_PRO = 0x0070 | RPW, WAIT_ALL = 0x0070;
    constexpr int FS = KR >= 32 ? 1 : 32 / KR;   // RNG streams per lane and replica
    using wide_t = typename std::conditional<WIDE, int64_t, int32_t>::type;
    static_assert(KR == 16 || KR == 32 || KR == 64, "unsupported large shape");

    // Rings of NR slots for both operands, filled by direct-to-LDS copies NR - 1 stages ahead of their use: A, one stage
    // of J' (2 K-blocks x 16 M-tiles x 64 lanes, 32 KiB per slot), and B, the stage's spin chunks [replica half][N-tile]
    // [lane] (8 KiB per slot). Separate arrays per slot, addressed by compile-time indices, so that the compiler sees
    // that a copy into one slot does not touch the slot being read (no wait is inserted for it). 120 KiB in all.
    constexpr int NR = 3;
    __shared__ v4i As0[2048], As1[2048], As2[2048];
    __shared__ v4i Bs0[2][NTW][64], Bs1[2][NTW][64], Bs2[2][NTW][64];
    auto aslot = [&](auto k) -> v4i (&)[2048] {
        constexpr int K = decltype(k)::value;
        if constexpr (K == 0) return As0;
        else if constexpr (K == 1) return As1;
        else return As2;
    };
    auto bslot = [&](auto k) -> v4i (&)[2][NTW][64] {
        constexpr int K = decltype(k)::value;
        if constexpr (K == 0) return Bs0;
        else if constexpr (K == 1) return Bs1;
        else return Bs2;
    };
    __shared__ int32_t Lb[256], Lr[256];
    __shared__ int bar_cells[3];

    const int xcd = blockIdx.x % 8, idx = blockIdx.x / 8;
    const int crow = 8 * (idx % PER) + xcd, team = idx / PER;
    if (team >= a.n_teams) return;
    const int tid = threadIdx.x, L = tid % 64, g = L / 16, ml = L % 16;
    const int w = __builtin_amdgcn_readfirstlane(tid / 64);
    const int wr = w >> 1, wc = w & 1;   // row half (128 rows) and replica half (RPT / 2 replicas) of this wave

    __amdgpu_buffer_rsrc_t jf_rsrc = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<int8_t*>(a.Jf) + size_t(crow) * NK * 16384, 0, 0x7fffffff, 0x00020000);
    constexpr int TEAM_XD = 4 * WC_XD;   // spins of a team: 2 parities x 2 replica halves
    constexpr int TEAMS = 256 / NRT;
    __amdgpu_buffer_rsrc_t xd_rsrc =
        __builtin_amdgcn_make_buffer_rsrc(a.xd + size_t(team) * TEAM_XD, 0, 0x7fffffff, 0x00020000);
    // n_lin counters [team][parity][replica] after all teams' spins (counted correction / trace): 64-bit words,
    // contributor count in bits 44..63 (modulo 2^20), running band count in bits 0..43; every (row tile, row half)
    // adds (1 << 44) + its partial
    unsigned long long* const ncnt =
        reinterpret_cast<unsigned long long*>(a.xd + size_t(TEAMS) * TEAM_XD + size_t(team) * 4096);
    constexpr unsigned long long NC_ONE = 1ull << 44;
    constexpr int CONTRIB = 2 * NRT;
    // per-lane part of the exchange addresses (the rest is wave-uniform and goes into scalar offsets): lane L's chunk
    const int v_lane = L * 16;
    // the 32 KiB of J' stage s into A slot `slot`: wave w copies bytes 8192 w .. 8192 w + 8191 (8 copies of 64 lanes x
    // 16 bytes; lane L's bytes land at 16 L after the copy's base)
    auto issue_a = [&](int s, auto slot) {
#pragma unroll
        for (int i = 0; i < 8; ++i)
            __builtin_amdgcn_raw_ptr_buffer_load_lds(jf_rsrc, (__attribute__((address_space(3))) void*)&aslot(slot)[w * 512 + i * 64], 16,
                                                     v_lane, s * 32768 + w * 8192 + i * 1024, 0, 0);
    };

    Lb[tid] = a.bp ? a.bp[256 * crow + tid] : 0;
    Lr[tid] = a.rs[256 * crow + tid];
    // barrier flags: cell 0 after each step's K loop, cell 2 for the per-step n_lin barrier of the counted kernels
    // (the stage barriers in between are plain)
    if (tid < 3) bar_cells[tid] = 0;   // visible after the first item's barrier

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

    const uint64_t t_limit = uint64_t(a.timeout_ticks);
    uint64_t t_wait = 0;
    bool bail = false;
    const int st0 = (xcd * NST) / 8;   // first stage of every step's K loop (the same on all workgroups of an XCD)
    // row j' = 4 mt + e (0..31) of this lane is row rho = 128 wr + 16 mt + 4 g + e of the tile: tile-row index
    // j = 32 wr + j' selects sub-lane j / KR and round j % KR. sl0 = sub-lane of this lane's first stream.
    const int sl0 = KR == 64 ? 0 : (32 * wr) / KR;
    const int lane0 = 4 * F * crow + F * g;   // RNG lane of sub-lane 0

    // the spin chunk of this lane's 32 rows of replica (wc, nt) for exchange step u: nibble 4 mt + e = row
    // 16 mt + 4 g + e of this row half, N = s + 1 + tau (bit j' of sbits: row j' = 4 mt + e has s = +1)
    auto store_spins = [&](int u, int nt, uint32_t sbits) {
        const uint32_t tw = tau_of(u) ? 0x11111111u : 0u;
        v4i chunk;
#pragma unroll
        for (int d = 0; d < 4; ++d) chunk[d] = int(tw | (spread_nibbles((sbits >> (8 * d)) & 0xFFu) << 1));
        const int soff = (u & 1) * 2 * WC_XD + wc * WC_XD + ((2 * crow + wr) * NTW + nt) * 1024;
        __builtin_amdgcn_raw_buffer_store_b128(chunk, xd_rsrc, v_lane, soff, 16);   // device scope (sc1)
    };

    // last band-count sum read per (step parity, N-tile): n_lin(v) = sum after step v - sum after step v - 2
    uint32_t nsum[2][NTW] = {};
    // waits until every contribution of decision step v is in replica rl's counter; returns n_lin(v)
    auto read_nlin = [&](int v, int nt, int rl) -> int32_t {
        unsigned long long* wp = ncnt + (v & 1) * 128 + rl;
        const uint32_t want = uint32_t(CONTRIB) * uint32_t(v / 2 + 1) & 0xFFFFFu;
        unsigned long long wv = __hip_atomic_load(wp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int polls = 0;
        while (uint32_t(wv >> 44) != want && !bail) {
            if ((++polls & 63) == 0) {
                const uint64_t now = wall_clock64();
                if (polls == 64) t_wait = now;
                if (aborted(a.status)) { bail = true; break; }
                if (now - t_wait > t_limit) { raise_error(a.status, 1); bail = true; break; }
            }
            __builtin_amdgcn_s_sleep(1);
            wv = __hip_atomic_load(wp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        const uint32_t sum = uint32_t(wv);
        const int32_t nl = int32_t(sum - nsum[v & 1][nt]);
        nsum[v & 1][nt] = sum;
        return nl;
    };
    const int n_items = (a.R + RPT - 1) / RPT;
    int seq = 0;
    for (int item = team; item < n_items; item += a.n_teams, ++seq) {
    __syncthreads();
    const int ubase = seq * (a.S + 1);
    const int r0 = item * RPT, cnt = min(RPT, a.R - r0);
    Rng rng[NTW][FS];
    uint32_t sb[NTW], spb[NTW];   // bit j': spin of row j' of this lane (replica nt) is +1
    uint32_t am[NTW];             // bit j': row j' is a real slot (x < n) of an active replica
    int32_t nfl[NTW];
    // rows of this lane that are real slots, and rounds that draw (k < kr; the rest are phantom rounds)
    uint32_t vmask = 0;
#pragma unroll
    for (int jj = 0; jj < 32; ++jj) {
        const int j = 32 * wr + jj;
        if (256 * (j % KR) + lane0 + j / KR < a.n) vmask |= 1u << jj;
    }
    // draws of one step for the streams of replica nt; use(jj, rv) for this lane's rows, in round order. Round k
    // draws iff k < kr; kr is passed through an empty asm so that the per-round tests are not hoisted out of the
    // step loop (they would pin scalar registers for the whole loop).
    auto draw_step = [&](int nt, auto use) {
        int kr = a.kr;
        if (PHANTOM) asm volatile("" : "+s"(kr));
        auto draws = [&](int k) { return !PHANTOM || k < kr; };
        if (KR == 64) {
            // one stream of 64 rounds (kr > 32): this wave's rows are rounds 32 wr .. 32 wr + 31, the others are
            // drawn and dropped (one copy of the row code for both row halves)
            if (wr == 1) {
#pragma unroll 1
                for (int k = 0; k < 32; ++k) rng[nt][0].next();
            }
#pragma unroll
            for (int k = 0; k < 32; ++k) use(k, draws(32 * wr + k) ? rng[nt][0].next() : 0u);
            if (wr == 0) {
#pragma unroll 1
                for (int k = 32; k < 64; ++k)
                    if (draws(k)) rng[nt][0].next();
            }
        } else {
#pragma unroll
            for (int jj = 0; jj < 32; ++jj) {
                const int k = (32 * wr + jj) % KR;
                use(jj, draws(k) ? rng[nt][jj / KR].next() : 0u);
            }
        }
    };
#pragma unroll
    for (int nt = 0; nt < NTW; ++nt) {
        const int rl = HALF * wc + 16 * nt + ml;
        const bool active = rl < cnt;
        const uint32_t trial = active ? (a.trials ? a.trials[r0 + rl] : a.trial0 + uint32_t(r0 + rl)) : 0u;
        sb[nt] = ~0u;
        spb[nt] = ~0u;
        am[nt] = active ? vmask : 0u;
        nfl[nt] = 0;
#pragma unroll
        for (int f = 0; f < FS; ++f) rng[nt][f].seed(a.seed, trial, uint32_t(lane0 + sl0 + f));
        // initial spins (padding draws consumed; phantom rounds k >= kr draw nothing)
        draw_step(nt, [&](int jj, uint32_t rv) {
            if (((am[nt] >> jj) & 1) && !(rv >> 31)) sb[nt] &= ~(1u << jj);
        });
    }
    // s(0) with tau(ubase)
#pragma unroll
    for (int nt = 0; nt < NTW; ++nt) store_spins(ubase, nt, sb[nt]);

    v4i acc[8][NTW];
#pragma unroll
    for (int mt = 0; mt < 8; ++mt)
#pragma unroll
        for (int nt = 0; nt < NTW; ++nt) acc[mt][nt] = v4i{0, 0, 0, 0};

    int tc = 0;   // H = b + acc - tc rowsum, tc = 1 + tau of the step's exchange
    for (int t = 0; t <= a.S; ++t) {
        const int u = ubase + t;
        // ===== acc = J' N_u over all K (stages of one K-pair; A through LDS, B through the LDS ring)
        {
#pragma unroll
            for (int mt = 0; mt < 8; ++mt)
#pragma unroll
                for (int nt = 0; nt < NTW; ++nt) acc[mt][nt] = v4i{0, 0, 0, 0};
            const uint32_t tau = tau_of(u);
            const uint32_t want = tau ? 0x11111111u : 0u;
            auto current = [&](const v4i (&b)[NTW]) {
                uint32_t all_and = 0xFFFFFFFFu, all_or = 0;
#pragma unroll
                for (int i = 0; i < NTW; ++i)
#pragma unroll
                    for (int e = 0; e < 4; ++e) { all_and &= uint32_t(b[i][e]); all_or |= uint32_t(b[i][e]); }
                return __all(((tau ? all_and : all_or) & 0x11111111u) == want);
            };
            // device-scope copies (sc1). A read through the XCD's L2 could hit a line cached four steps earlier, whose
            // tag bits are current again; device scope returns the producers' stores.
            auto copy_frag = [&](auto slot, int half, int i, int s) {
                __builtin_amdgcn_raw_ptr_buffer_load_lds(xd_rsrc, (__attribute__((address_space(3))) void*)&bslot(slot)[half][i][0], 16,
                                                         v_lane, (u & 1) * 2 * WC_XD + half * WC_XD + (NTW * s + i) * 1024, 0, 16);
            };
            // the 2 NTW chunk rows (replica half, N-tile) of stage index si_ into ring slot `slot` (= si_ % NR): wave w
            // copies rows RPW w .. RPW w + RPW - 1
            auto issue_b = [&](int si_, auto slot) {
                const int s_ = (st0 + si_) % NST;
#pragma unroll
                for (int k = 0; k < RPW; ++k) copy_frag(slot, (RPW * w + k) / NTW, (RPW * w + k) % NTW, s_);
            };
            auto read_b = [&](v4i (&b)[NTW], auto slot) {
#pragma unroll
                for (int i = 0; i < NTW; ++i) b[i] = bslot(slot)[wc][i][L];
            };
            using ic0 = std::integral_constant<int, 0>;
            using ic1 = std::integral_constant<int, 1>;
            using ic2 = std::integral_constant<int, 2>;
            // Pipeline: stage si computes from A and B slot si % NR; during stage si the copies of stage si + NR - 1
            // (B, then A) go into slot (si + NR - 1) % NR, whose last reads were in stage si - 1, before the barrier that
            // ended it. Each wave's copies of stage si + 1 were issued one stage earlier (or in the prologue); before the
            // barrier that ends stage si the wave waits until only the copies it issued during stage si are in flight
            // (vector memory counters retire in order), so the data of stage si + 1 is in LDS for every wave after it.
            // The first two A stages of a step were copied during the previous step's decisions (J' is the same every
            // step), except in the first step of a work item.
            static_assert(NR == 3, "three ring slots");
            if (t == 0) {
                issue_b(0, ic0{});
                issue_a(st0, ic0{});
                issue_b(1, ic1{});
                issue_a((st0 + 1) % NST, ic1{});
                copy_barrier<WAIT_STAGE>();   // stage 0's copies have landed (stage 1's may not)
            } else {
                issue_b(0, ic0{});
                issue_b(1, ic1{});
                copy_barrier<WAIT_PRO>();     // stage 0's copies, and everything issued before them
            }
            // stage si (si % NR = slot): reads slot `slot`, copies stage si + 2 into slot (slot + 2) % NR
            auto stage = [&](int si, auto slot) {
                constexpr int K = decltype(slot)::value;
                using next = std::integral_constant<int, (K + NR - 1) % NR>;
                const int s = (st0 + si) % NST;
                const bool ahead = si + NR - 1 < NST;
                if (ahead) {
                    issue_b(si + NR - 1, next{});
                    issue_a((s + NR - 1) % NST, next{});
                }
                v4i Bp[NTW];
                read_b(Bp, slot);
                // this stage's spin chunks must carry tag tau (copied again with device scope until they do)
                if (!bail && !current(Bp)) {
                    int polls = 0;
                    while (true) {
                        if ((++polls & 63) == 0) {
                            const uint64_t now = wall_clock64();
                            if (polls == 64) t_wait = now;
                            if (aborted(a.status)) { bail = true; break; }
                            if (now - t_wait > t_limit) { raise_error(a.status, 1); bail = true; break; }
                        }
                        __builtin_amdgcn_s_sleep(1);
#pragma unroll
                        for (int i = 0; i < NTW; ++i) copy_frag(slot, wc, i, s);
                        __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0): the copies have landed in LDS
                        read_b(Bp, slot);
                        if (current(Bp)) break;
                    }
                }
                if (!bail) {
                    // B fragments of the K-pair's even (Bc[nt]) and odd (Bc[NTW + nt]) nibbles
                    v4i Bc[2 * NTW];
#pragma unroll
                    for (int nt = 0; nt < NTW; ++nt)
#pragma unroll
                        for (int d = 0; d < 4; ++d) {
                            Bc[nt][d] = int(uint32_t(Bp[nt][d]) & 0x0F0F0F0Fu);
                            Bc[NTW + nt][d] = int((uint32_t(Bp[nt][d]) >> 4) & 0x0F0F0F0Fu);
                        }
                    v4i af[2][8];
#pragma unroll
                    for (int kl = 0; kl < 2; ++kl)
#pragma unroll
                        for (int mt = 0; mt < 8; ++mt) af[kl][mt] = aslot(slot)[kl * 1024 + (8 * wr + mt) * 64 + L];
#pragma unroll
                    for (int kl = 0; kl < 2; ++kl)
#pragma unroll
                        for (int mt = 0; mt < 8; ++mt)
#pragma unroll
                            for (int nt = 0; nt < NTW; ++nt)
                                acc[mt][nt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[kl][mt], Bc[NTW * kl + nt], acc[mt][nt], 0, 0, 0);
                    // schedule: four A reads from LDS ahead, then the NTW MFMAs of each read
                    __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);
#pragma unroll
                    for (int q = 0; q < 12; ++q) {
                        __builtin_amdgcn_sched_group_barrier(0x008, NTW, 0);
                        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
                    }
                    __builtin_amdgcn_sched_group_barrier(0x008, 4 * NTW, 0);
                }
                // this wave's copies of stage si + 1 have landed: at most the copies of this stage (RPW B, 8 A) remain
                // in flight, and stay in flight across the barrier
                if (ahead) copy_barrier<WAIT_STAGE>();
                else copy_barrier<WAIT_ALL>();
            };
            // every wave runs every stage (a wave that gave up a wait skips the work but keeps the barrier count);
            // whether any wave gave up is read once, after the K loop
            constexpr int NFULL = NST - NST % NR;
            for (int si = 0; si < NFULL; si += NR) {
                stage(si, ic0{});
                stage(si + 1, ic1{});
                stage(si + 2, ic2{});
            }
            if constexpr (NST % NR >= 1) stage(NFULL, ic0{});
            if constexpr (NST % NR >= 2) stage(NFULL + 1, ic1{});
        }
        if (barrier_any(bail, bar_cells[0])) bail = true;
        if (bail) break;
        // the first two A stages of step t + 1 (J' does not change), copied during this step's decisions; slots 0 and
        // 1 are free since the barrier above
        if (t < a.S) {
            issue_a(st0, std::integral_constant<int, 0>{});
            issue_a((st0 + 1) % NST, std::integral_constant<int, 1>{});
        }
        mark(0);
        // acc = J' (s + 1 + tau) over the whole row: H = b + acc - (1 + tau) rowsum
        tc = 1 + int(tau_of(u));
        // n_lin of decision step t - 1 (counted correction; also the trace, incl. the last step at t = S)
        int32_t nl_prev[NTW] = {};
        const int vstep = seq * a.S + t;   // decision-step index of this launch (parity and counts of the counters)
        if (COUNTED && t >= 1) {
#pragma unroll
            for (int nt = 0; nt < NTW; ++nt) nl_prev[nt] = read_nlin(vstep - 1, nt, HALF * wc + 16 * nt + ml);
            if (a.nlin && crow == 0 && wr == 0 && g == 0)
#pragma unroll
                for (int nt = 0; nt < NTW; ++nt) {
                    const int rl = HALF * wc + 16 * nt + ml;
                    if (rl < cnt) a.nlin[size_t(r0 + rl) * size_t(a.S) + (t - 1)] = nl_prev[nt];
                }
        }
        if (COUNTED && barrier_any(bail, bar_cells[2])) { bail = true; break; }
        if (t == a.S) break;
        // ===== decisions of step t (sca_ref.hpp:119-133), replica by replica in round order
        {
            const int32_t f4_t = a.fourT[t], q_t = a.q[t];
            const int64_t kc_t = COUNTED ? a.kcorr[t] : 0;
            const int32_t twoT = f4_t >> 1;
            const int32_t f_hi = f4_t >> 16, f_lo = f4_t & 0xFFFF;
#pragma unroll
            for (int nt = 0; nt < NTW; ++nt) {
                const int64_t corr = ((int64_t(nl_prev[nt]) * kc_t) >> 8) + int64_t(a.kconst[t]);
                const int64_t c_same = int64_t(q_t) - corr, c_diff = int64_t(q_t) + corr;
                const wide_t c1_diff = wide_t(c_diff + 1), c1_same = wide_t(c_same + 1);
                // band <=> lo(c) <= v <= hi(c)  (-2T < z < 2T)
                const wide_t lo_d = wide_t(((-int64_t(twoT) - c_diff) >> 16) + 1), lo_s = wide_t(((-int64_t(twoT) - c_same) >> 16) + 1);
                const wide_t hi_d = wide_t((int64_t(twoT) - c_diff - 1) >> 16), hi_s = wide_t((int64_t(twoT) - c_same - 1) >> 16);
                int32_t nband = 0;
                const uint32_t same = ~(sb[nt] ^ spb[nt]);
                uint32_t snew = sb[nt];
                draw_step(nt, [&](int jj, uint32_t rv) {
                    const int mt = jj >> 2, e = jj & 3;
                    const int rho = 128 * wr + 16 * mt + 4 * g + e;
                    const int32_t H = acc[mt][nt][e] + Lb[rho] - tc * Lr[rho];
                    const int32_t av = int32_t(rv >> 16) - 32768;
                    const int32_t rr = av * f_hi + ((av * f_lo) >> 16);
                    const bool up = (sb[nt] >> jj) & 1;
                    const bool cls = (same >> jj) & 1;
                    const int32_t v = up ? H : -H;
                    const bool flip = wide_t(v) <= ((wide_t(rr) - (cls ? c1_same : c1_diff)) >> 16);
                    const bool valid = (am[nt] >> jj) & 1;
                    const bool fl = valid & flip;
                    if (COUNTED)
                        nband += int32_t(valid & (wide_t(v) >= (cls ? lo_s : lo_d)) & (wide_t(v) <= (cls ? hi_s : hi_d)));
                    snew ^= fl ? (1u << jj) : 0u;
                    nfl[nt] += fl ? 1 : 0;
                });
                store_spins(u + 1, nt, snew);
                spb[nt] = sb[nt];
                sb[nt] = snew;
                if (COUNTED) {
                    // this wave's band count of replica (nt, ml) over its 128 rows, added to the counter
                    nband += __shfl_xor(nband, 16);
                    nband += __shfl_xor(nband, 32);
                    if (g == 0)
                        atomicAdd(ncnt + (vstep & 1) * 128 + HALF * wc + 16 * nt + ml, NC_ONE + (unsigned long long)nband);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
            mark(4);
        }
    }
    if (bail) break;
    // ===== outputs (sca_ref_bias.hpp:69-73). Each replica's fields go through LDS first (A slot 0 is idle after the
    // K loop; the next item's first copies into it follow the barrier at the top of the item loop), so that the
    // per-row loops run rolled with few live registers: unrolled, this once-per-item code would need every
    // accumulator in VGPRs at once and spill.
    int32_t* const Hs = reinterpret_cast<int32_t*>(&As0[0]) + w * 2048;   // this wave's 32 rows x 64 lanes
#pragma unroll
    for (int nt = 0; nt < NTW; ++nt) {
        const int rl = HALF * wc + 16 * nt + ml;
#pragma unroll
        for (int mt = 0; mt < 8; ++mt)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int rho = 128 * wr + 16 * mt + 4 * g + e;
                Hs[(4 * mt + e) * 64 + L] = acc[mt][nt][e] + Lb[rho] - tc * Lr[rho];
            }
        int64_t sh = 0, bs = 0;
        const uint32_t amn = am[nt], sbn = sb[nt];
#pragma unroll 1
        for (int jj = 0; jj < 32; ++jj) {
            const int j = 32 * wr + jj;
            const int rho = 128 * wr + 16 * (jj >> 2) + 4 * g + (jj & 3);
            const int32_t H = Hs[jj * 64 + L];
            const bool up = (sbn >> jj) & 1;
            if ((amn >> jj) & 1) {
                sh += up ? int64_t(H) : -int64_t(H);
                bs += up ? int64_t(Lb[rho]) : -int64_t(Lb[rho]);
                if (a.fields) {
                    const int x = 256 * (j % KR) + lane0 + j / KR;
                    a.fields[size_t(r0 + rl) * size_t(a.field_stride) + x] = H;
                }
            }
        }
        int64_t fl = nfl[nt];
#pragma unroll
        for (int o = 16; o < 64; o <<= 1) {
            sh += __shfl_xor(sh, o); bs += __shfl_xor(bs, o); fl += __shfl_xor(fl, o);
        }
        // packed spins: for round k, the 4F lanes of this row tile are 4F consecutive bits of one word
#pragma unroll 1
        for (int jj = 0; jj < 32; ++jj) {
            const int j = 32 * wr + jj, x = 256 * (j % KR) + lane0 + j / KR;
            uint64_t m = (((amn & sbn) >> jj) & 1) ? (1ull << (x & 63)) : 0ull;
            m |= __shfl_xor(m, 16);
            m |= __shfl_xor(m, 32);
            // lanes g = 0..3 hold the same (round, sub-lane) at bits 4F crow + F g + sl; one atomic per word and round
            if (g == 0 && rl < cnt && m)
                atomicOr(reinterpret_cast<unsigned long long*>(a.spins + size_t(r0 + rl) * size_t(a.spin_stride) + (x >> 6)),
                         (unsigned long long)m);
        }
        if (g == 0 && rl < cnt) {
            atomicAdd(reinterpret_cast<unsigned long long*>(a.sum_sh + r0 + rl), (unsigned long long)sh);
            atomicAdd(reinterpret_cast<unsigned long long*>(a.sum_bs + r0 + rl), (unsigned long long)bs);
            atomicAdd(reinterpret_cast<unsigned long long*>(a.flips + r0 + rl), (unsigned long long)fl);
        }
    }
    }   // work items
    if (prof) {
#pragma unroll
        for (int i = 0; i < 6; ++i) a.prof[size_t(blockIdx.x) * 8 + i] = (long long)pacc[i];
        a.prof[size_t(blockIdx.x) * 8 + 6] = int64_t(seq) * a.S;
        a.prof[size_t(blockIdx.x) * 8 + 7] = team * 1000 + crow;
    }
    __builtin_amdgcn_s_waitcnt(0);   // no LDS copy outlives the workgroup (a time-out can leave one in flight)
}

}  // namespace dev

template <int KR, bool WIDE, bool PHANTOM, bool COUNTED, int RPT>
inline int launch_large_t(const LargeArgs& a, int cus, bool dry, hipStream_t st) {
    const auto kern = dev::large_kernel<KR, WIDE, PHANTOM, COUNTED, RPT>;
    static int rc = 1;
    static size_t pad = 0;
    if (rc == 1) {
        hipFuncAttributes fa;
        rc = -1;
        if (hipFuncGetAttributes(&fa, (const void*)kern) == hipSuccess) {
            const size_t half = 80 * 1024 + 1024;
            pad = fa.sharedSizeBytes < half ? half - fa.sharedSizeBytes : 0;
            int per_cu = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)kern, dev::THREADS, pad) == hipSuccess)
                rc = per_cu >= 1 ? 0 : -2;
        }
    }
    if (rc) return rc;
    if (cus < 256 || a.n_teams * KR > 256) return -2;
    if (dry) return 0;
    hipLaunchKernelGGL(kern, dim3(256), dim3(dev::THREADS), pad, st, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int KR, bool COUNTED, int RPT>
inline int launch_large_r(const LargeArgs& a, bool wide, int cus, bool dry, hipStream_t st) {
    const bool ph = a.kr < KR;
    if (wide) return ph ? launch_large_t<KR, true, true, COUNTED, RPT>(a, cus, dry, st) : launch_large_t<KR, true, false, COUNTED, RPT>(a, cus, dry, st);
    return ph ? launch_large_t<KR, false, true, COUNTED, RPT>(a, cus, dry, st) : launch_large_t<KR, false, false, COUNTED, RPT>(a, cus, dry, st);
}

template <int KR, bool COUNTED>
inline int launch_large_k(const LargeArgs& a, bool wide, int cus, bool dry, hipStream_t st) {
    if (a.rpt == 128) return launch_large_r<KR, COUNTED, 128>(a, wide, cus, dry, st);
    if (a.rpt == 64) return launch_large_r<KR, COUNTED, 64>(a, wide, cus, dry, st);
    return -3;
}

// Per-translation-unit entry points: kernel slot count 256 KR, temperature-scaled (k) or counted (c) correction.
int launch_large_k16(const LargeArgs& a, bool wide, int cus, bool dry, hipStream_t st);
int launch_large_k32(const LargeArgs& a, bool wide, int cus, bool dry, hipStream_t st);
int launch_large_k64(const LargeArgs& a, bool wide, int cus, bool dry, hipStream_t st);
int launch_large_c16(const LargeArgs& a, bool wide, int cus, bool dry, hipStream_t st);
int launch_large_c32(const LargeArgs& a, bool wide, int cus, bool dry, hipStream_t st);
int launch_large_c64(const LargeArgs& a, bool wide, int cus, bool dry, hipStream_t st);

}  // namespace anc
