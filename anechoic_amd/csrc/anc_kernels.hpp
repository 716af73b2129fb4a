// Declarations shared by the HIP kernels (anc_engine.hip) and the host side (anc_host.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace anc {

// Returns ANC_ERR_HIP (with the message recorded for anc_last_error) from the enclosing function on a HIP error.
#define ANC_HIP(call)                                                                                       \
    do {                                                                                                    \
        hipError_t e_ = (call);                                                                             \
        if (e_ != hipSuccess) return fail(ANC_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));       \
    } while (0)

// One work item of the team kernel: `count` replicas (<= 16 NT) of one instance, output rows r0 .. r0 + count - 1,
// trial ids trials[r0 + i] (or trial0 + r0 + i); kr = NP(n)/256 rounds drawn per lane and step.
struct TeamItem {
    const int8_t* Jp;
    const int32_t* bias;
    int n;
    int r0;
    int count;
    int kr;
};

// Arguments of the persistent team kernel (one struct, passed by value).
struct TeamArgs {
    const int8_t* Jp;          // permuted J' (np x np, row-major), J'[y'][x'] with x' = l*(np/256) + k for slot x = k*256 + l
    int xmode;                 // exchange: 2 = spin crumbs (the default), 1 = spin nibbles, 0 = flip matrix (bytes)
    int np;                    // the kernel's padded slot count (2048 or 3072)
    int kr;                    // rounds drawn per lane and step: NP(n)/256 <= np/256 (the rest are phantom slots)
    int n;
    const int32_t* bias;       // original numbering, n entries, or null
    int S;
    const int32_t* fourT;
    const int32_t* q;
    const int64_t* kcorr;
    const int32_t* kconst;
    uint64_t seed;
    const uint32_t* trials;    // R trial ids, or null: trial = trial0 + r
    uint32_t trial0;
    int R;                     // replicas of the single-instance mode (items == null)
    int n_teams;
    const TeamItem* items;     // work items of the multi-instance mode, or null
    int n_items;
    uint64_t* xfrag;           // exchange: D' fragments [team][parity][member][nt][64 lanes][16 B]
    uint32_t* xnlin;           // exchange: tagged n_lin partials [team][parity][RT replicas][member]
    int* status;               // status block: [0] abort flag, [2] first error code (1 = exchange time-out)
    long long timeout_ticks;   // wall-clock ticks before a wait gives up
    int corr24;                // 1: the range plan proves the 24-bit split of the counted correction (16-member teams)
    int stall_test;            // 1: every workgroup first waits for a word that is never set, until the time-out (tests
                               // of the error path: the kernel's own bounded wait gives up and every workgroup exits)
    uint64_t* spins;
    int64_t spin_stride;
    int64_t* sum_sh;
    int64_t* sum_bs;
    int64_t* flips;
    int32_t* nlin;
    int32_t* fields;
    int64_t field_stride;
    long long* prof;           // optional: 8 counters per workgroup (phase cycles of thread 0), or null
};

// Arguments of the large-instance kernel (anc_large.hip): NP(n) > 3072.
struct LargeArgs {
    const int8_t* Jf;          // J' in fragment-major order [row tile][K-block][mt][lane][16 B] (kernel slot count npk),
                               // columns of each K-pair in the exchange's nibble order
    const int32_t* rs;         // row sums of J' per permuted row (npk)
    const int32_t* bp;         // bias per permuted row (npk), or null
    int n;
    int kr;                    // rounds drawn per lane and step: NP(n)/256 <= npk/256
    int S;
    const int32_t* fourT;
    const int32_t* q;
    const int32_t* kconst;
    const int64_t* kcorr;      // counted correction (Q8.24 per step; read only when count_nlin)
    int count_nlin;            // 1: n_lin is reduced over the instance every step (counted correction or trace)
    uint64_t seed;
    const uint32_t* trials;    // R trial ids, or null: trial = trial0 + r
    uint32_t trial0;
    int R;
    int rpt;                   // replicas per work item (replica tile): 128 or 64
    int n_teams;               // replica tiles annealed concurrently
    uint8_t* xd;               // exchange: spin nibbles [team][parity][replica half][K-pair][n-tile][lane][16 B], then
                               // the n_lin counters
    int* status;
    long long timeout_ticks;
    uint64_t* spins;
    int64_t spin_stride;
    int64_t* sum_sh;
    int64_t* sum_bs;
    int64_t* flips;
    int32_t* nlin;             // optional n_lin trace (R x S)
    int32_t* fields;
    int64_t field_stride;
    long long* prof;
};
// Kernel slot count of the large path for NP(n) > 3072: 4096, 8192 or 16384 (0 above 16384).
int large_np_for(int np);
// Bytes of the spin exchange buffer of the large path (all teams, both parities) and its n_lin counters.
size_t large_xd_bytes(int npk);
int launch_permute_large(const int8_t* J, int64_t ldj, int n, const int32_t* bias, int npk, int8_t* Jf, int32_t* rs,
                         int32_t* bp, hipStream_t st);
int launch_large(const LargeArgs& a, int npk, bool wide, int cus, bool dry, hipStream_t st);

// Mean-field relaxation (anc_meanfield.hip).
struct MFArgs {
    const int8_t* J;     // B instances, instance b at J + b * bstride, row stride ldj bytes
    int jbits;           // 8: int8 entries (ldj a multiple of 16, >= 16 ceil(n/16)); 4 / 2: packed (mf_pack's layouts;
                         // ldj a multiple of 16, >= (16 jbits / 8) ceil(n/16))
    int64_t ldj;
    int64_t bstride;
    int n;
    int B;
    float q;
    const float* T;      // S temperatures (device)
    const float* coef;   // S correction coefficients (device)
    int counted;         // 1: beta_t = coef_t * sum (1 - m_t^2); 0: beta_t = coef_t
};
int launch_mf_forward(const MFArgs& a, const float* m_init_prev, const float* m_init, int t0, int t1, float* m_buf,
                      float* u_tape, float* beta_tape, hipStream_t st);
int launch_mf_bwd_step(const MFArgs& a, int t, const float* g_next, const float* m_next, const float* m_cur,
                       const float* m_prev, const float* u_t, const float* delta_next, const float* beta_next, float* delta_out,
                       float* g_out, double* dT, double* dcoef, hipStream_t st);
int launch_mf_gemv(const MFArgs& a, const float* x, float* y, hipStream_t st);
// Persistent team kernels (anc_mf_team.hip), n <= 2048 on >= 256 CUs: one launch per checkpoint segment.
struct MFTeamArgs {
    uint32_t* xch;             // exchange slots [slot][instance][2048] (encoded floats; zeroed by the launchers)
    double* part;              // adjoint instance-sum partials [segment step][instance][member][s1, s2] (uncounted)
    int* status;               // [0] abort flag of the launch, [2] first error code (kept until the host reads it)
    long long timeout_ticks;   // bound of every wait (wall-clock ticks)
    int stall_test;            // 1: every workgroup waits on a word nothing sets until the bound (tests of the error path)
};
bool mf_team_fits(int n, int cus);
size_t mf_team_slot_words(int B);
size_t mf_team_part_doubles(int B);   // partial-sum doubles per segment step
int launch_mf_team_forward(const MFArgs& a, const MFTeamArgs& x, const float* m_init_prev, const float* m_init, int t0, int t1,
                           float* m_buf, float* u_tape, float* beta_tape, hipStream_t st);
int launch_mf_team_backward(const MFArgs& a, const MFTeamArgs& x, int t0, int t1, const float* seg, const float* utape,
                            const float* btape, const float* cp, const float* g_in, const float* delta_in, const float* beta_in,
                            float* g_out, float* delta_out, double* dT, double* dcoef, hipStream_t st);
int launch_mf_pack(int bits, int B, int n, const int8_t* J, int64_t ldj, int64_t bstride, uint8_t* Jp, int64_t ldp, int64_t bstridep,
                   int* bad, hipStream_t st);

int launch_check_instance(const int8_t* J, int64_t ldj, int n, const int32_t* bias, unsigned long long* stats, hipStream_t st);
int launch_permute(const int8_t* J, int64_t ldj, int n, int npk, int8_t* Jp, hipStream_t st);
int launch_fill_zero(uint64_t* p, int64_t words, hipStream_t st);
// Up to 8 ranges of uint64 words zeroed by one kernel launch (graph-capture safe: no memset nodes).
struct FillRanges {
    uint64_t* p[8];
    int64_t n[8];
    int count;
    void add(void* q, int64_t words) {
        if (q && words > 0) { p[count] = static_cast<uint64_t*>(q); n[count] = words; ++count; }
    }
};
int launch_fill_zero_ranges(const FillRanges& r, hipStream_t st);
int launch_gen_ising(uint64_t seed, int n, int M, uint64_t zthr, int bmax, int8_t* J, int64_t ldj, int32_t* bias, hipStream_t st);
// Team kernel (NP = 2048: 32-workgroup XCD-local teams; NP = 3072: 64-workgroup teams): nt = 1 or 2 sixteen-replica
// tiles per item, wide = int64 thresholds, xflavor 2 = XCD-local (L2) exchange, 1 = device scope.
// Returns 0, -1 (launch error), -2 (not co-resident) or -3 (no team shape for NP).
// dry = only check (and cache) the launch configuration; call once outside stream capture before capturing.
int launch_team(const TeamArgs& a, int nt, bool wide, int xflavor, int cus, bool dry, hipStream_t st);
// Both shapes in one launch: 32-workgroup teams of the NP = 2048 items (a2) on XCDs 0 .. xcd2-1, and the NP = 3072
// items (a3) on the remaining XCDs (spin exchanges: one 32-workgroup 12-round team per XCD, 16 nt3-replica items; flip
// exchange: 64-workgroup teams on XCD pairs, xcd2 even, 16 nt-replica items); xcd2 = 8 (spin exchanges): every
// XCD's team runs both classes in turn. a1 (spin exchanges): 4-round items of n <= 1024 instances (J' laid out with
// 1024 slots), run in turn before the NP 2048 items on the same teams. xflavor 1: device scope for all. rot (xcd2 =
// 8): XCD x runs team (x + (rot & 7)) mod 8 of a2 and (x + (rot >> 3)) mod 8 of a3.
int launch_team_mixed(const TeamArgs& a1, const TeamArgs& a2, const TeamArgs& a3, int xcd2, int rot, int nt, int nt3, bool wide,
                      int xflavor, int cus, bool dry, hipStream_t st);
// Workgroups per team for a padded slot count NP(n) (32 for 2048; for 2304..3072 32 with the spin exchange, xmode 1,
// and 64 with the flip exchange), 0 if the team kernel has no shape; and the kernel's own padded slot count (2048 or
// 3072) that J' is laid out with.
int team_size_for(int np, int xmode);
int kernel_np_for(int np);

}  // namespace anc
