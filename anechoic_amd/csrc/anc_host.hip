// Host side of libanechoic_mi355x: the C ABI (include/anechoic.h), contexts, instance preparation and launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "anc_internal.hpp"
#include "anc_kernels.hpp"
#include "anc_team16.hpp"

namespace anc {

static thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}
void clear_error() { g_err[0] = 0; }


struct Instance {
    int n = 0, np = 0;      // active spins, NP(n)
    int npk = 0;            // the team kernel's slot count J' is laid out with (0: no team shape)
    const int8_t* J = nullptr;
    int64_t ldj = 0;
    const int32_t* bias = nullptr;
    int64_t hbound = 0;
    int8_t* Jp = nullptr;   // permuted copy owned by the context
    int8_t* Jp4 = nullptr;  // n <= 1024: a second copy laid out with 1024 slots (4-round team kernels)
    bool large = false;     // NP(n) > 3072: the streamed large-instance kernel (Jp in fragment-major order)
    int32_t* rs = nullptr;  // large: row sums of J' per permuted row
    int32_t* bp = nullptr;  // large: bias per permuted row
};

constexpr int kTeamsMax = 8;
constexpr int kTeam = 32;
constexpr int kRTMax = 32;
// uint64 words: the largest team exchange, a mixed launch with 8 teams of each class: 8 teams x 2 parities x (64 + 128
// + 192) k-groups (1024-slot, NP 2048 and NP 3072 items) x 2 N-tiles x 16 replicas x 2
constexpr int64_t kFragWords = int64_t(kTeamsMax) * 2 * ((1024 + 2048 + 3072) / 16) * 2 * 16 * 2;
constexpr int64_t kNlinWords = int64_t(kTeamsMax) * 2 * 3 * kTeam * kRTMax / 2;  // uint32 pairs (three classes)
// cost of a 4-round work-item step relative to an NP 2048 one (the mixed launch's split model)
constexpr double kRatio4 = 0.6;
constexpr int64_t kStatusWords = 4;
constexpr int64_t kProfWords = 8 * 4096;   // 8 counters per workgroup

// Exchange workspace of anc_run_batched_noalloc (team-kernel exchange chunks, then n_lin words: the layout of the
// context workspace without its status block), 3.19 MiB per device for the life of the process. Host code never uses
// the array's name as an address: anc_create resolves the device address once per device (hipGetSymbolAddress) and
// every use goes through that pointer. Every context on a device shares the one array, so no-alloc runs that overlap
// on one device race.
constexpr int64_t kScratchWords = kFragWords + kNlinWords;
static __device__ __align__(256) uint64_t scratch[kScratchWords];
constexpr int kMaxDevices = 64;
static uint64_t* g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

// A captured multi-instance run: the instantiated graph plus the device buffers it owns (work items, tables).
struct GraphRec {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<void*> mem;
    bool live = false;
};

}  // namespace anc

struct anc_ctx {
    int device = 0;
    int cus = 0;
    long long wall_khz = 100000;
    int xflavor = 3;   // 3: XCD-local exchange through L2 (plain stores, nt loads); 1: device scope (fallback)
    double timeout_s = 5.0;
    anc::Tables tb;
    int S = 0;
    int32_t *d_fourT = nullptr, *d_q = nullptr, *d_kconst = nullptr;
    int64_t* d_kcorr = nullptr;
    std::vector<anc::Instance> inst;
    uint64_t* ws = nullptr;          // exchange fragments | n_lin words | status
    unsigned long long* stats = nullptr;
    long long* prof = nullptr;      // phase counters (anc_set_option "profile")
    anc::TeamItem* d_items = nullptr;    // work-item table of the last anc_run_multi
    size_t items_bytes = 0;
    int multi_nt = 2;                    // 16-replica tiles per work item of the multi-instance runs (1 or 2)
    int multi_nt3 = 2;                   // of the NP 3072 work items with the spin exchange (1 or 2; also single runs)
    int small_kr4 = 1;                   // multi-instance runs, spin exchanges: n <= 1024 items on 4-round teams
    int last_kr4 = 0;                    // 1 if the last team launch ran 4-round teams
    int last_nt3 = 0;                    // 16-replica tiles of the NP 3072 items of the last team launch (0: none)
    double mixed_ratio = 2.0;            // cost of an NP 3072 work-item step relative to an NP 2048 one (mixed launches;
                                         // measured with 32-replica items of both classes)
    int mixed_xcd2 = 0;                  // forced XCD split of mixed launches (1..7, or 8 = both shapes in turn on every XCD;
                                         // 2, 4, 6 with exchange 0; 0 = model)
    int last_xcd2 = 0;                   // XCDs given to NP 2048 teams by the last anc_run_multi (8: all, 0: none)
    int exchange = 2;                    // team kernels: 2 = spin crumbs, 1 = spin nibbles, 0 = flip-matrix bytes
    int fallback = 0;                    // 1: an exchange time-out switches the context to the device-scope exchange
    int stall_test = 0;                  // 1: team launches time out on purpose (host-forced; tests of the error path)
    int last_wide = 0;                   // 1 if the last launch decided with int64 thresholds
    int batched_team = 0;                // anc_run_batched, n <= 2048: 16 or 32-member teams (0: the default rule)
    int batched_nt = 0;                  // 32-member teams: 1 or 2 sixteen-replica tiles per item (0: the default rule)
    int last_team = 0;                   // team size of the last team launch (16 or 32; 64 for the flip exchange's NP 3072)
    int large_tile = 0;                  // large kernel: replicas per work item, 128 or 64 (0: the default rule)
    int last_large_tile = 0;             // replicas per work item of the last large-kernel launch
    uint64_t* scratch = nullptr;         // device address of this device's no-alloc exchange workspace
    hipStream_t cap = nullptr;           // private stream for graph capture
    uint8_t* lxd = nullptr;              // spin exchange buffer of the large-instance kernel (allocated on first use)
    size_t lxd_bytes = 0;
    std::vector<anc::GraphRec> graphs;
};

using namespace anc;

static void free_tables(anc_ctx* c) {
    if (c->d_fourT) hipFree(c->d_fourT);
    if (c->d_q) hipFree(c->d_q);
    if (c->d_kconst) hipFree(c->d_kconst);
    if (c->d_kcorr) hipFree(c->d_kcorr);
    c->d_fourT = c->d_q = c->d_kconst = nullptr;
    c->d_kcorr = nullptr;
    c->S = 0;
}

static int set_device(anc_ctx* c) {
    ANC_HIP(hipSetDevice(c->device));
    return ANC_OK;
}

static void free_instance(anc::Instance& i) {
    if (i.Jp) hipFree(i.Jp);
    if (i.Jp4) hipFree(i.Jp4);
    i.Jp4 = nullptr;
    if (i.rs) hipFree(i.rs);
    if (i.bp) hipFree(i.bp);
    i.Jp = nullptr;
    i.rs = i.bp = nullptr;
}

static void free_graph(anc::GraphRec& g);
static void free_graphs(anc_ctx* c) {
    if (c->graphs.empty()) return;
    hipDeviceSynchronize();
    for (auto& g : c->graphs) free_graph(g);
    c->graphs.clear();
}

extern "C" {

const char* anc_last_error(void) { return anc::g_err; }
int anc_version(void) { return 1; }

anc_ctx* anc_create(int device) {
    clear_error();
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
        fail(ANC_ERR_HIP, "anc_create: no HIP device %d (count %d)", device, count);
        return nullptr;
    }
    std::unique_ptr<anc_ctx> c(new anc_ctx);
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) { fail(ANC_ERR_HIP, "hipSetDevice failed"); return nullptr; }
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) != hipSuccess) { fail(ANC_ERR_HIP, "hipGetDeviceProperties failed"); return nullptr; }
    c->cus = p.multiProcessorCount;
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0) c->wall_khz = khz;
    const size_t bytes = size_t(kFragWords + kNlinWords + kStatusWords) * 8;
    if (hipMalloc(&c->ws, bytes) != hipSuccess) { fail(ANC_ERR_HIP, "workspace allocation failed"); return nullptr; }
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (device >= kMaxDevices) {
            hipFree(c->ws);
            fail(ANC_ERR_HIP, "anc_create: device %d beyond %d", device, kMaxDevices);
            return nullptr;
        }
        if (!g_scratch_ptr[device]) {   // once per device, with that device current
            void* ptr = nullptr;
            const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !ptr) {
                hipFree(c->ws);
                fail(ANC_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
                return nullptr;
            }
            g_scratch_ptr[device] = static_cast<uint64_t*>(ptr);
        }
        c->scratch = g_scratch_ptr[device];
    }
    if (hipMemset(c->ws, 0, bytes) != hipSuccess) {   // the status block starts clear
        hipFree(c->ws);
        fail(ANC_ERR_HIP, "workspace clear failed");
        return nullptr;
    }
    if (hipMalloc(&c->stats, 8 * sizeof(unsigned long long)) != hipSuccess) {
        hipFree(c->ws);
        fail(ANC_ERR_HIP, "stats allocation failed");
        return nullptr;
    }
    return c.release();
}

void anc_destroy(anc_ctx* c) {
    if (!c) return;
    hipSetDevice(c->device);
    free_graphs(c);
    if (c->cap) hipStreamDestroy(c->cap);
    free_tables(c);
    for (auto& i : c->inst) free_instance(i);
    if (c->ws) hipFree(c->ws);
    if (c->lxd) hipFree(c->lxd);
    if (c->stats) hipFree(c->stats);
    if (c->prof) hipFree(c->prof);
    if (c->d_items) hipFree(c->d_items);
    delete c;
}

int anc_set_tables(anc_ctx* c, int S, const int32_t* fourT, const int32_t* q, const int64_t* kcorr, const int32_t* kconst) {
    clear_error();
    if (!c || !fourT || !q || !kcorr || !kconst) return fail(ANC_ERR_ARG, "anc_set_tables: null argument");
    if (S < 1 || S > ANC_MAX_STEPS) return fail(ANC_ERR_ARG, "anc_set_tables: S = %d outside 1..%d", S, ANC_MAX_STEPS);
    for (int t = 0; t < S; ++t)
        if (fourT[t] <= 0) return fail(ANC_ERR_RANGE, "anc_set_tables: fourT[%d] = %d must be positive", t, fourT[t]);
    if (int rc = set_device(c)) return rc;
    free_tables(c);
    c->tb.fourT.assign(fourT, fourT + S);
    c->tb.q.assign(q, q + S);
    c->tb.kcorr.assign(kcorr, kcorr + S);
    c->tb.kconst.assign(kconst, kconst + S);
    ANC_HIP(hipMalloc(&c->d_fourT, sizeof(int32_t) * S));
    ANC_HIP(hipMalloc(&c->d_q, sizeof(int32_t) * S));
    ANC_HIP(hipMalloc(&c->d_kconst, sizeof(int32_t) * S));
    ANC_HIP(hipMalloc(&c->d_kcorr, sizeof(int64_t) * S));
    ANC_HIP(hipMemcpy(c->d_fourT, fourT, sizeof(int32_t) * S, hipMemcpyHostToDevice));
    ANC_HIP(hipMemcpy(c->d_q, q, sizeof(int32_t) * S, hipMemcpyHostToDevice));
    ANC_HIP(hipMemcpy(c->d_kconst, kconst, sizeof(int32_t) * S, hipMemcpyHostToDevice));
    ANC_HIP(hipMemcpy(c->d_kcorr, kcorr, sizeof(int64_t) * S, hipMemcpyHostToDevice));
    c->S = S;
    return ANC_OK;
}

int anc_add_instance(anc_ctx* c, int n, const int8_t* J, int64_t ldj, const int32_t* bias, void* stream) {
    clear_error();
    if (!c || !J) return fail(ANC_ERR_ARG, "anc_add_instance: null argument");
    if (n < 1 || n > ANC_MAX_N) return fail(ANC_ERR_ARG, "anc_add_instance: n = %d outside 1..%d", n, ANC_MAX_N);
    if (ldj < n) return fail(ANC_ERR_ARG, "anc_add_instance: ldj = %lld < n = %d", (long long)ldj, n);
    if (int rc = set_device(c)) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    ANC_HIP(hipMemsetAsync(c->stats, 0, 8 * sizeof(unsigned long long), st));
    if (launch_check_instance(J, ldj, n, bias, c->stats, st)) return fail(ANC_ERR_HIP, "instance check launch failed");
    unsigned long long h[8];
    ANC_HIP(hipMemcpyAsync(h, c->stats, sizeof h, hipMemcpyDeviceToHost, st));
    ANC_HIP(hipStreamSynchronize(st));
    if (h[0]) return fail(ANC_ERR_INSTANCE, "J is not symmetric (%llu asymmetric entries)", h[0]);
    if (h[1]) return fail(ANC_ERR_INSTANCE, "J has %llu nonzero diagonal entries", h[1]);
    Instance in;
    in.n = n; in.np = padded_slots(n); in.J = J; in.ldj = ldj; in.bias = bias;
    in.hbound = int64_t(h[2]) + int64_t(h[3]);
    if (in.hbound >= (int64_t(1) << 30)) return fail(ANC_ERR_INSTANCE, "field bound %lld exceeds 2^30", (long long)in.hbound);
    in.npk = kernel_np_for(in.np);
    if (!in.npk) {
        in.npk = large_np_for(in.np);
        in.large = true;
    }
    if (!in.npk) return fail(ANC_ERR_UNSUPPORTED, "n = %d: no kernel shape", n);
    ANC_HIP(hipMalloc(&in.Jp, size_t(in.npk) * size_t(in.npk)));
    if (in.large) {
        if (hipMalloc(&in.rs, sizeof(int32_t) * size_t(in.npk)) != hipSuccess ||
            hipMalloc(&in.bp, sizeof(int32_t) * size_t(in.npk)) != hipSuccess) {
            free_instance(in);
            return fail(ANC_ERR_HIP, "instance buffer allocation failed");
        }
    }
    if (!in.large && n <= 1024 && hipMalloc(&in.Jp4, size_t(1024) * 1024) != hipSuccess) {
        free_instance(in);
        return fail(ANC_ERR_HIP, "instance buffer allocation failed");
    }
    int prc = in.large ? launch_permute_large(J, ldj, n, bias, in.npk, in.Jp, in.rs, in.bp, st)
                       : launch_permute(J, ldj, n, in.npk, in.Jp, st);
    if (!prc && in.Jp4) prc = launch_permute(J, ldj, n, 1024, in.Jp4, st);
    if (prc || hipStreamSynchronize(st) != hipSuccess) {
        free_instance(in);
        return fail(ANC_ERR_HIP, "permute launch failed");
    }
    c->inst.push_back(in);
    return int(c->inst.size()) - 1;
}

int anc_clear_instances(anc_ctx* c) {
    if (!c) return fail(ANC_ERR_ARG, "anc_clear_instances: null context");
    if (int rc = set_device(c)) return rc;
    free_graphs(c);   // captured runs reference the instances
    ANC_HIP(hipDeviceSynchronize());
    for (auto& i : c->inst) free_instance(i);
    c->inst.clear();
    return ANC_OK;
}

int anc_gen_ising(uint64_t seed, int n, int kbits, double pzero, int bmax, int8_t* J, int64_t ldj, int32_t* bias, void* stream) {
    clear_error();
    if (n < 1 || n > ANC_MAX_N || kbits < 2 || kbits > 8 || !J || ldj < n || bmax < 0 || !(pzero >= 0.0 && pzero <= 1.0))
        return fail(ANC_ERR_ARG, "anc_gen_ising: bad argument (1 <= n <= %d, 2 <= kbits <= 8, ldj >= n, 0 <= pzero <= 1)", ANC_MAX_N);
    const double lim = std::ldexp(pzero, 64);
    const uint64_t zthr = pzero > 0 ? (lim >= 18446744073709551615.0 ? ~0ull : uint64_t(lim)) : 0;
    if (launch_gen_ising(seed, n, (1 << (kbits - 1)) - 1, zthr, bmax, J, ldj, bias, static_cast<hipStream_t>(stream)))
        return fail(ANC_ERR_HIP, "anc_gen_ising launch failed");
    return ANC_OK;
}

int64_t anc_instance_bound(anc_ctx* c, int inst) {
    if (!c || inst < 0 || inst >= int(c->inst.size())) return fail(ANC_ERR_ARG, "anc_instance_bound: bad instance");
    return c->inst[size_t(inst)].hbound;
}

}  // extern "C"

// The accumulated outputs of R rows, as ranges for one fill launch (graph-capture safe).
static void out_ranges(const anc_out& o, int R, FillRanges& f) {
    f.add(o.spins, int64_t(R) * o.spin_stride);
    f.add(o.sum_sh, R);
    f.add(o.sum_bs, R);
    f.add(o.flips, R);
}

// ---------------------------------------------------------------------------------------------------------- launch plans
// Everything one persistent team launch needs (one team shape, or both shapes in one launch), enqueued eagerly or
// under stream capture. The device pointers it holds (tables, instances, work items, workspace, outputs) must stay
// valid while the plan is in use.
struct LaunchPlan {
    int mode = 0;            // 2: NP 2048 teams only (a2), 3: NP 3072 teams only (a3), 4: both (mixed launch), 5: NP 2048
                             // 16-member teams (a2)
    TeamArgs a1{}, a2{}, a3{};   // mixed: 4-round items of n <= 1024 instances, NP 2048 items, NP 3072 items
    int xcd2 = 0;            // mixed: XCDs given to the NP <= 2048 teams (8: every XCD runs all classes in turn)
    int rot = 0;             // mixed, xcd2 = 8: team rotation of the NP 2048 (bits 0-2) and NP 3072 (bits 3-5) items
    int nt = 2;              // 16-replica tiles of the NP 2048 work items (and of the NP 3072 ones with the flip exchange)
    int nt3 = 1;             // of the NP 3072 work items (1 with the spin exchange)
    bool wide = false;
    int xflavor = 3;
    int64_t frag_words = 0;  // exchange fragment words cleared before the launch
    uint64_t* xws = nullptr; // exchange workspace: fragments, then n_lin words at xws + kFragWords
    anc_out out{};
    int R = 0;               // output rows cleared before the launch
};

struct TablesRef {
    int S;
    const int32_t *fourT, *q, *kconst;
    const int64_t* kcorr;
};

static void common_args(anc_ctx* c, const TablesRef& t, TeamArgs& a, uint64_t seed, const uint32_t* trials, uint32_t trial0,
                        const anc_out& o) {
    a.S = t.S; a.fourT = t.fourT; a.q = t.q; a.kcorr = t.kcorr; a.kconst = t.kconst;
    a.seed = seed; a.trials = trials; a.trial0 = trial0;
    a.status = reinterpret_cast<int*>(c->ws + kFragWords + kNlinWords);
    a.timeout_ticks = (long long)(c->timeout_s * double(c->wall_khz) * 1000.0);
    a.spins = o.spins; a.spin_stride = o.spin_stride;
    a.sum_sh = o.sum_sh; a.sum_bs = o.sum_bs; a.flips = o.flips;
    a.nlin = o.nlin; a.fields = o.fields; a.field_stride = o.field_stride;
    a.prof = c->prof;
    a.xmode = c->exchange;
    a.stall_test = c->stall_test;
}

static TablesRef ctx_tables(anc_ctx* c) { return TablesRef{c->S, c->d_fourT, c->d_q, c->d_kconst, c->d_kcorr}; }

// The 16-member team kernel's conditions: NP 2048, the crumb exchange, and 12 x field bound < 2^31 (its odd crumb
// positions accumulate 4 J' N, whose partial sums must stay below 2^31).
static bool team16_ok(const anc_ctx* c, const Instance& in) {
    return kernel_np_for(in.np) == 2048 && c->exchange == 2 && in.hbound * 12 < (int64_t(1) << 31);
}

// R replicas of one instance: 16 nt replicas per work item, up to 8 teams (NP 2048) or 4 teams (NP 3072).
static int plan_single(anc_ctx* c, const Instance& in, uint64_t seed, const uint32_t* trials, uint32_t trial0, int R, int nt,
                       const anc_out& o, LaunchPlan& lp, bool team16 = false, uint64_t* xws = nullptr) {
    Plan p;
    if (int rc = plan(c->tb, in.n, in.hbound, p)) return rc;
    if (!xws) xws = c->ws;
    lp.xws = xws;
    if (team16) {   // 16-member teams (NP 2048, crumb exchange): 16-replica items, up to 16 teams (two per XCD)
        if (!team16_ok(c, in))
            return fail(ANC_ERR_UNSUPPORTED, "16-member teams need n <= 2048, the crumb exchange and 12 x field bound < 2^31");
        const int items = (R + 15) / 16;
        const int teams = std::min(std::min(16, 8 * (c->cus / 128)), items);
        if (teams < 1) return fail(ANC_ERR_UNSUPPORTED, "16-member teams need 128 CUs (have %d)", c->cus);
        lp.mode = 5;
        TeamArgs& a = lp.a2;
        common_args(c, ctx_tables(c), a, seed, trials, trial0, o);
        a.Jp = in.Jp; a.np = 2048; a.kr = 8; a.n = in.n; a.bias = in.bias; a.R = R;
        a.n_teams = teams; a.items = nullptr; a.n_items = 0;
        a.corr24 = p.corr24 ? 1 : 0;
        a.xfrag = xws;
        a.xnlin = reinterpret_cast<uint32_t*>(xws + kFragWords);
        lp.nt = lp.nt3 = 1; lp.wide = p.wide; lp.xflavor = c->xflavor; lp.out = o; lp.R = R;
        lp.frag_words = int64_t(teams) * dev::T16_FRAG_U64;
        if (lp.frag_words > kFragWords || int64_t(teams) * dev::T16_NLIN_U32 > 2 * kNlinWords)
            return fail(ANC_ERR_UNSUPPORTED, "exchange workspace too small");
        return ANC_OK;
    }
    const int knp = kernel_np_for(in.np), team = team_size_for(in.np, c->exchange);
    if (!team) return fail(ANC_ERR_UNSUPPORTED, "team kernel: no team shape for NP = %d", in.np);
    if (knp == 3072 && c->exchange) nt = std::min(nt, c->multi_nt3);   // 32 workgroups x 12 rounds
    const int rt = 16 * nt, items = (R + rt - 1) / rt;
    const int teams = std::min(std::min(kTeamsMax, c->cus / team), items);
    if (teams < 1) return fail(ANC_ERR_UNSUPPORTED, "team kernel: %d CUs cannot host a team of %d", c->cus, team);
    lp.mode = knp == 2048 ? 2 : 3;
    TeamArgs& a = knp == 2048 ? lp.a2 : lp.a3;
    common_args(c, ctx_tables(c), a, seed, trials, trial0, o);
    a.Jp = in.Jp; a.np = knp; a.kr = in.np / 256; a.n = in.n; a.bias = in.bias; a.R = R;
    a.n_teams = teams; a.items = nullptr; a.n_items = 0;
    a.xfrag = xws;
    a.xnlin = reinterpret_cast<uint32_t*>(xws + kFragWords);
    lp.nt = lp.nt3 = nt; lp.wide = p.wide; lp.xflavor = c->xflavor; lp.out = o; lp.R = R;
    lp.frag_words = int64_t(teams) * 2 * (knp / 16) * nt * 32;   // the kernel's slot count, not n's
    if (lp.frag_words > kFragWords) return fail(ANC_ERR_UNSUPPORTED, "exchange workspace too small");
    return ANC_OK;
}

// Work items of a multi-instance run: the 4-round items of n <= 1024 instances first (n1 of them), then the NP 2048
// items (n2), then the NP 2304..3072 items.
struct MultiItems {
    std::vector<TeamItem> items;
    int n1 = 0, n2 = 0, R = 0, nmax = 1, npmax = 0;
    int64_t hb = 0;
};

// Items of NP 2048 instances have 16 nt replicas, those of NP 3072 instances 16 multi_nt3 (spin exchange) or 16 nt.
// With the spin exchanges and option small_kr4, the items of n <= 1024 instances run on 4-round teams (J' laid out
// with 1024 slots; each lane still draws the 8 numbers per step of the NP 2048 layout).
static int build_items(anc_ctx* c, int n_groups, const int32_t* inst, const int32_t* count, int nt, MultiItems& m) {
    const int rt2 = 16 * nt, rt3 = c->exchange ? 16 * c->multi_nt3 : 16 * nt;
    std::vector<TeamItem> i1, i2, i3;
    for (int gi = 0; gi < n_groups; ++gi) {
        if (inst[gi] < 0 || inst[gi] >= int(c->inst.size())) return fail(ANC_ERR_STATE, "group %d: instance %d not registered", gi, inst[gi]);
        if (count[gi] < 0) return fail(ANC_ERR_ARG, "group %d: negative count", gi);
        const Instance& in = c->inst[size_t(inst[gi])];
        const int knp = kernel_np_for(in.np);
        if (!knp || in.large)
            return fail(ANC_ERR_UNSUPPORTED, "group %d: n = %d > 3072 runs through anc_run_batched only", gi, in.n);
        const int rt = knp == 2048 ? rt2 : rt3;
        const bool small = knp == 2048 && in.Jp4 && c->small_kr4 && c->exchange != 0;
        for (int q = 0; q < count[gi]; q += rt) {
            TeamItem it{};
            it.Jp = small ? in.Jp4 : in.Jp; it.bias = in.bias; it.n = in.n; it.r0 = m.R + q; it.count = std::min(rt, count[gi] - q);
            it.kr = small ? 4 : in.np / 256;
            (small ? i1 : knp == 2048 ? i2 : i3).push_back(it);
        }
        m.R += count[gi];
        m.nmax = std::max(m.nmax, in.n);
        m.npmax = std::max(m.npmax, in.np);
        m.hb = std::max(m.hb, in.hbound);
    }
    if (m.R < 1) return fail(ANC_ERR_ARG, "no replicas");
    m.n1 = int(i1.size());
    m.n2 = int(i2.size());
    m.items = std::move(i1);
    m.items.insert(m.items.end(), i2.begin(), i2.end());
    m.items.insert(m.items.end(), i3.begin(), i3.end());
    return ANC_OK;
}

static int check_multi_out(const anc_out* out, const MultiItems& m) {
    if (!out->spins || !out->sum_sh || !out->sum_bs || !out->flips) return fail(ANC_ERR_ARG, "null output array");
    if (out->spin_stride < m.npmax / 64) return fail(ANC_ERR_ARG, "spin_stride too small");
    if (out->fields && out->field_stride < m.nmax) return fail(ANC_ERR_ARG, "field_stride < n");
    return ANC_OK;
}

// Team counts for the item classes. NP 2048 items alone: as many 32-workgroup teams as fit (mode 2); NP 3072 items
// alone: likewise (flip exchange: 4 teams of 64; mode 3). Otherwise one mixed launch (mode 4): the teams of XCDs
// 0 .. xcd2-1 run the 4-round items and then the NP 2048 items, the remaining XCDs the NP 3072 items (spin exchange:
// one 32-workgroup team per XCD, xcd2 in 1..7; flip exchange: 64-workgroup teams on XCD pairs, xcd2 in {2, 4, 6}), or
// (spin exchanges, xcd2 = 8) every XCD's team runs all classes in turn; xcd2 minimises
// max(rounds_4 kRatio4 + rounds_2048, ratio rounds_3072) (xcd2 = 8: the sum) with `ratio` = the cost of an NP 3072 item
// step relative to an NP 2048 one, or is forced by `mixed_xcd2` (ignored without NP 3072 items: every XCD then runs
// the NP <= 2048 classes).
static int plan_items(anc_ctx* c, const TablesRef& tabs, const MultiItems& m, const TeamItem* d_items, uint64_t seed,
                      const uint32_t* trials, const anc_out& o, int nt, LaunchPlan& lp) {
    Plan p;
    if (int rc = plan(c->tb, m.nmax, m.hb, p)) return rc;
    const bool spin = c->exchange != 0;
    const int n1 = m.n1, n2 = m.n2, n3 = int(m.items.size()) - m.n1 - m.n2, rt2 = 16 * nt;
    const int nt3 = spin ? c->multi_nt3 : nt, rt3 = 16 * nt3, team3 = spin ? 32 : 64;
    int t1 = 0, t2 = 0, t3 = 0;
    if (n3 == 0 && n1 == 0) {
        lp.mode = 2;
        t2 = std::min(std::min(kTeamsMax, c->cus / 32), n2);
    } else if (n2 == 0 && n1 == 0) {
        lp.mode = 3;
        t3 = std::min(std::min(spin ? kTeamsMax : 4, c->cus / team3), n3);
    } else {
        if (c->cus < 256) return fail(ANC_ERR_UNSUPPORTED, "mixed team launch needs 256 CUs (have %d)", c->cus);
        lp.mode = 4;
        double best = 1e300;
        for (int x2 = spin ? 1 : 2; x2 <= (spin ? 8 : 6); x2 += spin ? 1 : 2) {
            if (n3 == 0 ? x2 != 8 : c->mixed_xcd2 && x2 != c->mixed_xcd2) continue;
            const int a1 = std::min(x2, n1), a2 = std::min(x2, n2);
            const int b = std::min(x2 == 8 ? 8 : spin ? 8 - x2 : (8 - x2) / 2, n3);
            const double r12 = (a1 ? kRatio4 * double((n1 + a1 - 1) / a1) : 0.0) + (a2 ? double((n2 + a2 - 1) / a2) : 0.0);
            const double r3 = b ? c->mixed_ratio * double((n3 + b - 1) / b) : 0.0;
            const double est = x2 == 8 ? r12 + r3 : std::max(r12, r3);
            if (est < best) { best = est; lp.xcd2 = x2; t1 = a1; t2 = a2; t3 = b; }
        }
        if (!lp.xcd2) return fail(ANC_ERR_ARG, "mixed_xcd2 = %d is not a split of this exchange mode", c->mixed_xcd2);
    }
    if ((n1 && t1 < 1) || (n2 && t2 < 1) || (n3 && t3 < 1)) return fail(ANC_ERR_UNSUPPORTED, "team kernel: not enough CUs (%d)", c->cus);
    lp.rot = 0;
    if (lp.mode == 4 && lp.xcd2 == 8) {
        // items go round-robin from team 0 within each class: rotate the later classes' team numbering over the XCDs so
        // that their teams with an extra item land where the classes before them left the least work
        auto items = [](int n, int t, int j) { return j < t ? (n - j + t - 1) / t : 0; };
        double load[8];
        for (int x = 0; x < 8; ++x) load[x] = kRatio4 * items(n1, t1, x);
        double best = 1e300;
        int r2b = 0, r3b = 0;
        for (int r2 = 0; r2 < 8; ++r2)
            for (int r3 = 0; r3 < 8; ++r3) {
                double mx = 0.0;
                for (int x = 0; x < 8; ++x)
                    mx = std::max(mx, load[x] + items(n2, t2, (x + r2) & 7) + c->mixed_ratio * items(n3, t3, (x + r3) & 7));
                if (mx < best - 1e-9) { best = mx; r2b = r2; r3b = r3; }
            }
        lp.rot = r2b | (r3b << 3);
    }
    common_args(c, tabs, lp.a1, seed, trials, 0, o);
    common_args(c, tabs, lp.a2, seed, trials, 0, o);
    common_args(c, tabs, lp.a3, seed, trials, 0, o);
    lp.a1.np = 1024; lp.a1.kr = 4; lp.a1.R = m.R; lp.a1.n_teams = t1; lp.a1.items = d_items; lp.a1.n_items = n1;
    lp.a2.np = 2048; lp.a2.kr = 8; lp.a2.R = m.R; lp.a2.n_teams = t2; lp.a2.items = d_items + n1; lp.a2.n_items = n2;
    lp.a3.np = 3072; lp.a3.kr = 12; lp.a3.R = m.R; lp.a3.n_teams = t3; lp.a3.items = d_items + n1 + n2; lp.a3.n_items = n3;
    const int64_t f2 = int64_t(t2) * 2 * (2048 / 16) * nt * 32, f3 = int64_t(t3) * 2 * (3072 / 16) * nt3 * 32;
    const int64_t f1 = int64_t(t1) * 2 * (1024 / 16) * nt * 32;
    lp.xws = c->ws;
    lp.a2.xfrag = c->ws;
    lp.a3.xfrag = c->ws + f2;
    lp.a1.xfrag = c->ws + f2 + f3;
    lp.a2.xnlin = reinterpret_cast<uint32_t*>(c->ws + kFragWords);
    lp.a3.xnlin = lp.a2.xnlin + size_t(t2) * 2 * 32 * rt2;
    lp.a1.xnlin = lp.a3.xnlin + size_t(t3) * 2 * team3 * rt3;
    lp.frag_words = f1 + f2 + f3;
    if (lp.frag_words > kFragWords ||
        int64_t(t2) * 2 * 32 * rt2 + int64_t(t3) * 2 * team3 * rt3 + int64_t(t1) * 2 * 32 * rt2 > 2 * kNlinWords)
        return fail(ANC_ERR_UNSUPPORTED, "exchange workspace too small");
    lp.nt = nt; lp.nt3 = nt3; lp.wide = p.wide; lp.xflavor = c->xflavor; lp.out = o; lp.R = m.R;
    return ANC_OK;
}

// Clears the exchange workspace, the per-launch status word (the abort flag; the error word is kept until
// anc_sync_status reports it) and the output rows (fill kernels: no memset nodes under capture), then launches.
// dry = only validate and cache the launch configuration.
static int enqueue(anc_ctx* c, const LaunchPlan& lp, bool dry, hipStream_t st) {
    if (!dry) {
        FillRanges f{};
        f.add(lp.xws, lp.frag_words);
        f.add(lp.xws + kFragWords, kNlinWords);
        f.add(c->ws + kFragWords + kNlinWords, 1);   // the abort flag (the error word is kept)
        out_ranges(lp.out, lp.R, f);
        if (launch_fill_zero_ranges(f, st)) return fail(ANC_ERR_HIP, "fill launch failed");
    }
    int rc;
    if (!dry) {
        c->last_wide = lp.wide ? 1 : 0;
        c->last_team = lp.mode == 5 ? 16 : (lp.mode == 3 && c->exchange == 0) ? 64 : 32;
        c->last_nt3 = (lp.mode == 3 || lp.mode == 4) && lp.a3.n_teams > 0 ? lp.nt3 : 0;
        c->last_kr4 = lp.mode == 4 && lp.a1.n_teams > 0 ? 1 : 0;
    }
    if (lp.mode == 5) rc = launch_team16(lp.a2, lp.wide, lp.xflavor, c->cus, dry, st);
    else if (lp.mode == 4) rc = launch_team_mixed(lp.a1, lp.a2, lp.a3, lp.xcd2, lp.rot, lp.nt, lp.nt3, lp.wide, lp.xflavor, c->cus, dry, st);
    else rc = launch_team(lp.mode == 2 ? lp.a2 : lp.a3, lp.mode == 2 ? lp.nt : lp.nt3, lp.wide, lp.xflavor, c->cus, dry, st);
    if (rc == -2) return fail(ANC_ERR_UNSUPPORTED, "team kernel: the teams are not co-resident on %d CUs", c->cus);
    if (rc) return fail(ANC_ERR_HIP, "team kernel launch failed: %s", hipGetErrorString(hipGetLastError()));
    return ANC_OK;
}

static int check_run(anc_ctx* c, int inst, const anc_out* out, int R) {
    if (!c || !out) return fail(ANC_ERR_ARG, "null context or outputs");
    if (c->S < 1) return fail(ANC_ERR_STATE, "no tables set");
    if (inst < 0 || inst >= int(c->inst.size())) return fail(ANC_ERR_STATE, "instance %d not registered", inst);
    if (!out->spins || !out->sum_sh || !out->sum_bs || !out->flips) return fail(ANC_ERR_ARG, "null output array");
    if (R < 1) return fail(ANC_ERR_ARG, "R = %d must be positive", R);
    const Instance& in = c->inst[size_t(inst)];
    if (out->spin_stride < in.np / 64) return fail(ANC_ERR_ARG, "spin_stride %lld < %d words", (long long)out->spin_stride, in.np / 64);
    if (out->fields && out->field_stride < in.n) return fail(ANC_ERR_ARG, "field_stride < n");
    return set_device(c);
}

static void free_graph(anc::GraphRec& g) {
    if (g.exec) hipGraphExecDestroy(g.exec);
    if (g.graph) hipGraphDestroy(g.graph);
    for (void* p : g.mem)
        if (p) hipFree(p);
    g = anc::GraphRec{};
}

// R replicas of a large instance (NP(n) > 3072): 128 or 64 replicas per work item (option large_tile; 0 = 64 when
// 128-replica items would leave teams idle), 256 / KR items in flight.
static int run_large(anc_ctx* c, const Instance& in, uint64_t seed, const uint32_t* trials, uint32_t trial0, int R,
                     const anc_out& o, hipStream_t st) {
    bool counted = false;
    for (int t = 0; t < c->S; ++t) counted = counted || c->tb.kcorr[size_t(t)] != 0;
    Plan p;
    if (int rc = plan(c->tb, in.n, in.hbound, p)) return rc;
    const size_t xb = large_xd_bytes(in.npk);
    if (xb > c->lxd_bytes) {
        ANC_HIP(hipStreamSynchronize(st));
        if (c->lxd) ANC_HIP(hipFree(c->lxd));
        c->lxd = nullptr;
        c->lxd_bytes = 0;
        ANC_HIP(hipMalloc(&c->lxd, xb));
        c->lxd_bytes = xb;
    }
    LargeArgs a{};
    a.Jf = in.Jp; a.rs = in.rs; a.bp = in.bp; a.n = in.n; a.kr = in.np / 256;
    a.S = c->S; a.fourT = c->d_fourT; a.q = c->d_q; a.kconst = c->d_kconst; a.kcorr = c->d_kcorr;
    a.count_nlin = (counted || o.nlin) ? 1 : 0;
    a.seed = seed; a.trials = trials; a.trial0 = trial0; a.R = R;
    const int kr_k = in.npk / 256, teams = 256 / kr_k;
    a.rpt = c->large_tile ? c->large_tile : ((R + 127) / 128 < teams ? 64 : 128);
    const int items = (R + a.rpt - 1) / a.rpt;
    a.n_teams = std::min(teams, items);
    c->last_large_tile = a.rpt;
    a.xd = c->lxd;
    a.status = reinterpret_cast<int*>(c->ws + kFragWords + kNlinWords);
    a.timeout_ticks = (long long)(c->timeout_s * double(c->wall_khz) * 1000.0);
    a.spins = o.spins; a.spin_stride = o.spin_stride;
    a.sum_sh = o.sum_sh; a.sum_bs = o.sum_bs; a.flips = o.flips;
    a.nlin = o.nlin; a.fields = o.fields; a.field_stride = o.field_stride;
    a.prof = c->prof;
    FillRanges f{};
    f.add(c->lxd, int64_t(xb / 8));
    f.add(c->ws + kFragWords, kNlinWords + 1);   // n_lin words and the abort flag (the error word is kept)
    out_ranges(o, R, f);
    if (launch_fill_zero_ranges(f, st)) return fail(ANC_ERR_HIP, "fill launch failed");
    c->last_wide = p.wide ? 1 : 0;
    const int rc = launch_large(a, in.npk, p.wide, c->cus, false, st);
    if (rc == -2) return fail(ANC_ERR_UNSUPPORTED, "large kernel: 256 co-resident workgroups needed (%d CUs)", c->cus);
    if (rc) return fail(ANC_ERR_HIP, "large kernel launch failed: %s", hipGetErrorString(hipGetLastError()));
    return ANC_OK;
}

extern "C" {

int anc_run_plain(anc_ctx* c, int inst, uint64_t seed, uint32_t trial, const anc_out* out, void* stream) {
    clear_error();
    if (int rc = check_run(c, inst, out, 1)) return rc;
    if (c->inst[size_t(inst)].large)
        return run_large(c, c->inst[size_t(inst)], seed, nullptr, trial, 1, *out, static_cast<hipStream_t>(stream));
    LaunchPlan lp;
    if (int rc = plan_single(c, c->inst[size_t(inst)], seed, nullptr, trial, 1, 1, *out, lp)) return rc;
    return enqueue(c, lp, false, static_cast<hipStream_t>(stream));
}

int anc_run_batched(anc_ctx* c, int inst, uint64_t seed, const uint32_t* trials, int R, const anc_out* out, void* stream) {
    clear_error();
    if (int rc = check_run(c, inst, out, R)) return rc;
    if (c->inst[size_t(inst)].large)
        return run_large(c, c->inst[size_t(inst)], seed, trials, 0, R, *out, static_cast<hipStream_t>(stream));
    const Instance& in = c->inst[size_t(inst)];
    // default shape rule: 16-member teams for n <= 2048 with the crumb exchange (options batched_team / batched_nt, set
    // by the autotuner, override it)
    const bool team16 = c->batched_team == 16 || (c->batched_team == 0 && team16_ok(c, in));
    const int nt = c->batched_nt ? c->batched_nt : (R <= kTeamsMax * 16 ? 1 : 2);
    LaunchPlan lp;
    if (int rc = plan_single(c, in, seed, trials, 0, R, nt, *out, lp, team16)) return rc;
    return enqueue(c, lp, false, static_cast<hipStream_t>(stream));
}

int anc_run_batched_noalloc(anc_ctx* c, int inst, uint64_t seed, const uint32_t* trials, int R, const anc_out* out, void* stream) {
    clear_error();
    if (int rc = check_run(c, inst, out, R)) return rc;
    const Instance& in = c->inst[size_t(inst)];
    if (in.large)
        return fail(ANC_ERR_UNSUPPORTED, "anc_run_batched_noalloc: n = %d > 3072 is not supported (its exchange needs %zu "
                    "bytes beyond the %lld-byte static workspace); use anc_run_batched", in.n, large_xd_bytes(in.npk),
                    (long long)(kScratchWords * 8));
    if (!c->scratch) return fail(ANC_ERR_STATE, "anc_run_batched_noalloc: no static workspace on device %d", c->device);
    const bool team16 = c->batched_team == 16 || (c->batched_team == 0 && team16_ok(c, in));
    const int nt = c->batched_nt ? c->batched_nt : (R <= kTeamsMax * 16 ? 1 : 2);
    LaunchPlan lp;
    if (int rc = plan_single(c, in, seed, trials, 0, R, nt, *out, lp, team16, c->scratch)) return rc;
    // plan_single refuses exchange needs beyond the workspace (kFragWords fragment words, kNlinWords n_lin words)
    return enqueue(c, lp, false, static_cast<hipStream_t>(stream));
}

int anc_run_multi(anc_ctx* c, int n_groups, const int32_t* inst, const int32_t* count, uint64_t seed, const uint32_t* trials,
                  const anc_out* out, void* stream) {
    clear_error();
    if (!c || !inst || !count || !out || n_groups < 1) return fail(ANC_ERR_ARG, "anc_run_multi: bad argument");
    if (c->S < 1) return fail(ANC_ERR_STATE, "no tables set");
    MultiItems m;
    if (int rc = build_items(c, n_groups, inst, count, c->multi_nt, m)) return rc;
    if (int rc = check_multi_out(out, m)) return rc;
    if (int rc = set_device(c)) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const size_t bytes = sizeof(TeamItem) * m.items.size();
    // the previous launch on this stream may still read the item table
    ANC_HIP(hipStreamSynchronize(st));
    if (bytes > c->items_bytes) {
        if (c->d_items) ANC_HIP(hipFree(c->d_items));
        c->d_items = nullptr;
        c->items_bytes = 0;
        ANC_HIP(hipMalloc(&c->d_items, bytes));
        c->items_bytes = bytes;
    }
    ANC_HIP(hipMemcpy(c->d_items, m.items.data(), bytes, hipMemcpyHostToDevice));
    LaunchPlan lp;
    if (int rc = plan_items(c, ctx_tables(c), m, c->d_items, seed, trials, *out, c->multi_nt, lp)) return rc;
    c->last_xcd2 = lp.mode == 4 ? lp.xcd2 : lp.mode == 2 ? 8 : 0;
    return enqueue(c, lp, false, st);
}

int anc_graph_multi(anc_ctx* c, int n_groups, const int32_t* inst, const int32_t* count, uint64_t seed, const uint32_t* trials,
                    const anc_out* out) {
    clear_error();
    if (!c || !inst || !count || !out || n_groups < 1) return fail(ANC_ERR_ARG, "anc_graph_multi: bad argument");
    if (c->S < 1) return fail(ANC_ERR_STATE, "no tables set");
    MultiItems m;
    if (int rc = build_items(c, n_groups, inst, count, c->multi_nt, m)) return rc;
    if (int rc = check_multi_out(out, m)) return rc;
    if (int rc = set_device(c)) return rc;
    GraphRec g;
    // the graph owns its work-item table and a copy of the current tables
    const size_t ib = sizeof(TeamItem) * m.items.size(), S = size_t(c->S);
    TablesRef tabs{c->S, nullptr, nullptr, nullptr, nullptr};
    void* mem[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    const size_t sizes[5] = {ib, 4 * S, 4 * S, 4 * S, 8 * S};
    for (int i = 0; i < 5; ++i)
        if (hipMalloc(&mem[i], sizes[i]) != hipSuccess) {
            for (void* p : mem) if (p) hipFree(p);
            return fail(ANC_ERR_HIP, "graph buffer allocation failed");
        }
    g.mem.assign(mem, mem + 5);
    const void* src[5] = {m.items.data(), c->tb.fourT.data(), c->tb.q.data(), c->tb.kconst.data(), c->tb.kcorr.data()};
    for (int i = 0; i < 5; ++i)
        if (hipMemcpy(mem[i], src[i], sizes[i], hipMemcpyHostToDevice) != hipSuccess) {
            free_graph(g);
            return fail(ANC_ERR_HIP, "graph buffer upload failed");
        }
    tabs.fourT = static_cast<const int32_t*>(mem[1]);
    tabs.q = static_cast<const int32_t*>(mem[2]);
    tabs.kconst = static_cast<const int32_t*>(mem[3]);
    tabs.kcorr = static_cast<const int64_t*>(mem[4]);
    LaunchPlan lp;
    int rc = plan_items(c, tabs, m, static_cast<const TeamItem*>(mem[0]), seed, trials, *out, c->multi_nt, lp);
    if (!rc) rc = enqueue(c, lp, true, nullptr);
    if (!rc && !c->cap && hipStreamCreateWithFlags(&c->cap, hipStreamNonBlocking) != hipSuccess)
        rc = fail(ANC_ERR_HIP, "capture stream creation failed");
    if (!rc && hipStreamBeginCapture(c->cap, hipStreamCaptureModeThreadLocal) != hipSuccess)
        rc = fail(ANC_ERR_HIP, "hipStreamBeginCapture failed");
    if (!rc) {
        const int lrc = enqueue(c, lp, false, c->cap);
        const hipError_t e = hipStreamEndCapture(c->cap, &g.graph);
        rc = lrc ? lrc : e != hipSuccess ? fail(ANC_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e)) : ANC_OK;
    }
    if (!rc && hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0) != hipSuccess)
        rc = fail(ANC_ERR_HIP, "hipGraphInstantiate failed");
    if (rc) {
        free_graph(g);
        return rc;
    }
    g.live = true;
    for (size_t i = 0; i < c->graphs.size(); ++i)
        if (!c->graphs[i].live) {
            c->graphs[i] = g;
            return int(i);
        }
    c->graphs.push_back(g);
    return int(c->graphs.size()) - 1;
}

int anc_graph_launch(anc_ctx* c, int graph, void* stream) {
    clear_error();
    if (!c || graph < 0 || graph >= int(c->graphs.size()) || !c->graphs[size_t(graph)].live)
        return fail(ANC_ERR_STATE, "graph %d does not exist", graph);
    if (int rc = set_device(c)) return rc;
    ANC_HIP(hipGraphLaunch(c->graphs[size_t(graph)].exec, static_cast<hipStream_t>(stream)));
    return ANC_OK;
}

int anc_graph_destroy(anc_ctx* c, int graph) {
    clear_error();
    if (!c || graph < 0 || graph >= int(c->graphs.size()) || !c->graphs[size_t(graph)].live)
        return fail(ANC_ERR_STATE, "graph %d does not exist", graph);
    if (int rc = set_device(c)) return rc;
    ANC_HIP(hipDeviceSynchronize());
    free_graph(c->graphs[size_t(graph)]);
    return ANC_OK;
}

int anc_sync_status(anc_ctx* c, void* stream) {
    clear_error();
    if (!c) return fail(ANC_ERR_ARG, "null context");
    if (int rc = set_device(c)) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    ANC_HIP(hipStreamSynchronize(st));
    // the error word holds the first error of any launch since the last report (a later launch does not clear it)
    uint64_t* err_word = c->ws + kFragWords + kNlinWords + 1;
    int s[2] = {0, 0};
    ANC_HIP(hipMemcpy(s, err_word, sizeof s, hipMemcpyDeviceToHost));
    if (!s[0]) return ANC_OK;
    ANC_HIP(hipMemset(err_word, 0, sizeof(uint64_t)));
    if (s[0] == 1) {
        const int was = c->xflavor;
        if (c->fallback) c->xflavor = 1;
        return fail(ANC_ERR_HIP, "team exchange timed out (flavour %d)%s", was,
                    (c->fallback && was != 1) ? "; the context now uses the device-scope exchange (option fallback)" : "");
    }
    return fail(ANC_ERR_HIP, "device status %d", s[0]);
}

double anc_get_option(anc_ctx* c, const char* key) {
    clear_error();
    if (!c || !key) return fail(ANC_ERR_ARG, "null argument");
    if (!std::strcmp(key, "xflavor")) return c->xflavor;
    if (!std::strcmp(key, "timeout_s")) return c->timeout_s;
    if (!std::strcmp(key, "multi_nt")) return c->multi_nt;
    if (!std::strcmp(key, "multi_nt3")) return c->multi_nt3;
    if (!std::strcmp(key, "last_nt3")) return c->last_nt3;
    if (!std::strcmp(key, "small_kr4")) return c->small_kr4;
    if (!std::strcmp(key, "last_kr4")) return c->last_kr4;
    if (!std::strcmp(key, "mixed_ratio")) return c->mixed_ratio;
    if (!std::strcmp(key, "mixed_xcd2")) return c->mixed_xcd2;
    if (!std::strcmp(key, "last_xcd2")) return c->last_xcd2;
    if (!std::strcmp(key, "fallback")) return c->fallback;
    if (!std::strcmp(key, "stall_test")) return c->stall_test;
    if (!std::strcmp(key, "last_wide")) return c->last_wide;
    if (!std::strcmp(key, "batched_team")) return c->batched_team;
    if (!std::strcmp(key, "batched_nt")) return c->batched_nt;
    if (!std::strcmp(key, "last_team")) return c->last_team;
    if (!std::strcmp(key, "large_tile")) return c->large_tile;
    if (!std::strcmp(key, "last_large_tile")) return c->last_large_tile;
    if (!std::strcmp(key, "exchange")) return c->exchange;
    if (!std::strcmp(key, "profile")) return c->prof ? 1 : 0;
    return fail(ANC_ERR_ARG, "unknown option %s", key);
}

int anc_get_profile(anc_ctx* c, long long* out, int count) {
    if (!c || !out) return fail(ANC_ERR_ARG, "null argument");
    if (!c->prof) return fail(ANC_ERR_STATE, "profiling is off (anc_set_option \"profile\" 1)");
    if (int rc = set_device(c)) return rc;
    const int n = std::min<int>(count, int(kProfWords));
    ANC_HIP(hipDeviceSynchronize());
    ANC_HIP(hipMemcpy(out, c->prof, sizeof(long long) * size_t(n), hipMemcpyDeviceToHost));
    return n;
}

int anc_set_option(anc_ctx* c, const char* key, double value) {
    if (!c || !key) return fail(ANC_ERR_ARG, "null argument");
    if (!std::strcmp(key, "xflavor")) {
        if (value != 1 && value != 2 && value != 3)
            return fail(ANC_ERR_ARG, "xflavor must be 1 (device scope), 2 or 3 (XCD-local through L2)");
        c->xflavor = int(value);
        return ANC_OK;
    }
    if (!std::strcmp(key, "timeout_s")) { c->timeout_s = value > 0 ? value : 5.0; return ANC_OK; }
    if (!std::strcmp(key, "multi_nt")) {
        if (value != 1 && value != 2) return fail(ANC_ERR_ARG, "multi_nt must be 1 or 2 (16 or 32 replicas per work item)");
        c->multi_nt = int(value);
        return ANC_OK;
    }
    if (!std::strcmp(key, "small_kr4")) {
        if (value != 0 && value != 1) return fail(ANC_ERR_ARG, "small_kr4 must be 0 or 1");
        c->small_kr4 = int(value);
        return ANC_OK;
    }
    if (!std::strcmp(key, "multi_nt3")) {
        if (value != 1 && value != 2) return fail(ANC_ERR_ARG, "multi_nt3 must be 1 or 2 (16 or 32 replicas per NP 3072 work item)");
        c->multi_nt3 = int(value);
        return ANC_OK;
    }
    if (!std::strcmp(key, "mixed_ratio")) {
        if (!(value > 0.05 && value < 20)) return fail(ANC_ERR_ARG, "mixed_ratio must be in (0.05, 20)");
        c->mixed_ratio = value;
        return ANC_OK;
    }
    if (!std::strcmp(key, "mixed_xcd2")) {
        if (!(value >= 0 && value <= 8 && value == double(int(value)))) return fail(ANC_ERR_ARG, "mixed_xcd2 must be 0..8");
        c->mixed_xcd2 = int(value);
        return ANC_OK;
    }
    if (!std::strcmp(key, "exchange")) {
        if (value < 0 || value > 2) return fail(ANC_ERR_ARG, "exchange must be 2 (spin crumbs), 1 (spin nibbles) or 0 (flip matrix)");
        c->exchange = int(value);
        return ANC_OK;
    }
    if (!std::strcmp(key, "fallback")) {
        if (value != 0 && value != 1) return fail(ANC_ERR_ARG, "fallback must be 0 (report time-outs only) or 1");
        c->fallback = int(value);
        return ANC_OK;
    }
    if (!std::strcmp(key, "batched_team")) {
        if (value != 0 && value != 16 && value != 32) return fail(ANC_ERR_ARG, "batched_team must be 0 (default rule), 16 or 32");
        c->batched_team = int(value);
        return ANC_OK;
    }
    if (!std::strcmp(key, "batched_nt")) {
        if (value != 0 && value != 1 && value != 2) return fail(ANC_ERR_ARG, "batched_nt must be 0 (default rule), 1 or 2");
        c->batched_nt = int(value);
        return ANC_OK;
    }
    if (!std::strcmp(key, "large_tile")) {
        if (value != 0 && value != 64 && value != 128) return fail(ANC_ERR_ARG, "large_tile must be 0 (default rule), 64 or 128");
        c->large_tile = int(value);
        return ANC_OK;
    }
    if (!std::strcmp(key, "stall_test")) {
        if (value != 0 && value != 1) return fail(ANC_ERR_ARG, "stall_test must be 0 or 1");
        c->stall_test = int(value);
        return ANC_OK;
    }
    if (!std::strcmp(key, "profile")) {
        for (const auto& g : c->graphs)   // captured runs hold the counter buffer's address
            if (g.live) return fail(ANC_ERR_STATE, "profile cannot change while captured runs exist (anc_graph_destroy them first)");
        if (int rc = set_device(c)) return rc;
        if (value != 0 && !c->prof) {
            ANC_HIP(hipMalloc(&c->prof, kProfWords * sizeof(long long)));
            ANC_HIP(hipMemset(c->prof, 0, kProfWords * sizeof(long long)));
        } else if (value == 0 && c->prof) {
            ANC_HIP(hipFree(c->prof));
            c->prof = nullptr;
        }
        return ANC_OK;
    }
    return fail(ANC_ERR_ARG, "unknown option %s", key);
}

}  // extern "C"
