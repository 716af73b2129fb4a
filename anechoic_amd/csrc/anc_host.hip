// Host side of libanechoic_mi355x: the C ABI (include/anechoic.h), contexts, instance preparation and launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "anc_internal.hpp"
#include "anc_kernels.hpp"

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
};

constexpr int kTeamsMax = 8;
constexpr int kTeam = 32;
constexpr int kRTMax = 32;
constexpr int64_t kFragWords = int64_t(kTeamsMax) * 2 * kTeam * 2 * 64 * 2;   // uint64 words
constexpr int64_t kNlinWords = int64_t(kTeamsMax) * 2 * kTeam * kRTMax / 2;  // uint32 pairs
constexpr int64_t kStatusWords = 4;
constexpr int64_t kProfWords = 8 * 4096;   // 8 counters per workgroup

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
    free_tables(c);
    for (auto& i : c->inst)
        if (i.Jp) hipFree(i.Jp);
    if (c->ws) hipFree(c->ws);
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
    if (!in.npk) return fail(ANC_ERR_UNSUPPORTED, "n = %d: the team kernels cover n <= 3072", n);
    ANC_HIP(hipMalloc(&in.Jp, size_t(in.npk) * size_t(in.npk)));
    if (launch_permute(J, ldj, n, in.npk, in.Jp, st)) {
        hipFree(in.Jp);
        return fail(ANC_ERR_HIP, "permute launch failed");
    }
    ANC_HIP(hipStreamSynchronize(st));
    c->inst.push_back(in);
    return int(c->inst.size()) - 1;
}

int anc_clear_instances(anc_ctx* c) {
    if (!c) return fail(ANC_ERR_ARG, "anc_clear_instances: null context");
    if (int rc = set_device(c)) return rc;
    for (auto& i : c->inst)
        if (i.Jp) hipFree(i.Jp);
    c->inst.clear();
    return ANC_OK;
}

int64_t anc_instance_bound(anc_ctx* c, int inst) {
    if (!c || inst < 0 || inst >= int(c->inst.size())) return fail(ANC_ERR_ARG, "anc_instance_bound: bad instance");
    return c->inst[size_t(inst)].hbound;
}

}  // extern "C"

// Zeroes the accumulated outputs of R rows (fill kernels: graph-capture safe).
static int zero_out(const anc_out& o, int R, hipStream_t st) {
    return launch_fill_zero(o.spins, int64_t(R) * o.spin_stride, st) ||
           launch_fill_zero(reinterpret_cast<uint64_t*>(o.sum_sh), R, st) ||
           launch_fill_zero(reinterpret_cast<uint64_t*>(o.sum_bs), R, st) ||
           launch_fill_zero(reinterpret_cast<uint64_t*>(o.flips), R, st);
}

// One persistent team launch: R replicas of one instance (items == null) or a table of work items.
static int run_team(anc_ctx* c, const Instance& in, uint64_t seed, const uint32_t* trials, uint32_t trial0, int R, int nt,
                    const TeamItem* d_items, int n_items, int nmax, int64_t hbound, int np, const anc_out& o, hipStream_t st,
                    bool zero_outputs = true) {
    Plan p;
    if (int rc = plan(c->tb, nmax, hbound, p)) return rc;
    const int team = team_size_for(np);
    if (!team) return fail(ANC_ERR_UNSUPPORTED, "team kernel: no team shape for NP = %d (n <= 2048 or 2305..3072)", np);
    const int rt = 16 * nt;
    const int items = d_items ? n_items : (R + rt - 1) / rt;
    const int teams_max = std::min(kTeamsMax, c->cus / team);
    TeamArgs a{};
    a.Jp = in.Jp; a.np = kernel_np_for(np); a.kr = np / 256; a.n = in.n; a.bias = in.bias;
    a.S = c->S; a.fourT = c->d_fourT; a.q = c->d_q; a.kcorr = c->d_kcorr; a.kconst = c->d_kconst;
    a.seed = seed; a.trials = trials; a.trial0 = trial0; a.R = R;
    a.n_teams = std::min(teams_max, items);
    if (a.n_teams < 1) return fail(ANC_ERR_UNSUPPORTED, "team kernel: %d CUs cannot host a team of %d", c->cus, team);
    a.items = d_items; a.n_items = n_items;
    a.xfrag = c->ws;
    a.xnlin = reinterpret_cast<uint32_t*>(c->ws + kFragWords);
    a.status = reinterpret_cast<int*>(c->ws + kFragWords + kNlinWords);
    a.timeout_ticks = (long long)(c->timeout_s * double(c->wall_khz) * 1000.0);
    a.spins = o.spins; a.spin_stride = o.spin_stride;
    a.sum_sh = o.sum_sh; a.sum_bs = o.sum_bs; a.flips = o.flips;
    a.nlin = o.nlin; a.fields = o.fields; a.field_stride = o.field_stride;
    a.prof = c->prof;
    // zero the exchange workspace, the status words and the accumulated outputs (fill kernels: graph-capture safe)
    const int64_t frag_words = int64_t(a.n_teams) * 2 * (a.np / 16) * nt * 32;   // the kernel's slot count, not n's
    if (frag_words > kFragWords) return fail(ANC_ERR_UNSUPPORTED, "exchange workspace too small");
    if (launch_fill_zero(c->ws, frag_words, st) || launch_fill_zero(c->ws + kFragWords, kNlinWords + kStatusWords, st))
        return fail(ANC_ERR_HIP, "fill launch failed");
    if (zero_outputs && zero_out(o, R, st)) return fail(ANC_ERR_HIP, "fill launch failed");
    const int rc = launch_team(a, nt, p.wide, c->xflavor, c->cus, st);
    if (rc == -2) return fail(ANC_ERR_UNSUPPORTED, "team kernel: %d teams of %d workgroups are not co-resident on %d CUs", a.n_teams, kTeam, c->cus);
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

extern "C" {

int anc_run_plain(anc_ctx* c, int inst, uint64_t seed, uint32_t trial, const anc_out* out, void* stream) {
    clear_error();
    if (int rc = check_run(c, inst, out, 1)) return rc;
    const Instance& in = c->inst[size_t(inst)];
    return run_team(c, in, seed, nullptr, trial, 1, 1, nullptr, 0, in.n, in.hbound, in.np, *out, static_cast<hipStream_t>(stream));
}

int anc_run_batched(anc_ctx* c, int inst, uint64_t seed, const uint32_t* trials, int R, const anc_out* out, void* stream) {
    clear_error();
    if (int rc = check_run(c, inst, out, R)) return rc;
    const Instance& in = c->inst[size_t(inst)];
    const int nt = R <= kTeamsMax * 16 ? 1 : 2;
    return run_team(c, in, seed, trials, 0, R, nt, nullptr, 0, in.n, in.hbound, in.np, *out, static_cast<hipStream_t>(stream));
}

int anc_run_multi(anc_ctx* c, int n_groups, const int32_t* inst, const int32_t* count, uint64_t seed, const uint32_t* trials,
                  const anc_out* out, void* stream) {
    clear_error();
    if (!c || !inst || !count || !out || n_groups < 1) return fail(ANC_ERR_ARG, "anc_run_multi: bad argument");
    if (c->S < 1) return fail(ANC_ERR_STATE, "no tables set");
    const int nt = 2, rt = 16 * nt;
    // work items per padded-slot class (one persistent launch per class: 32-workgroup teams for NP = 2048,
    // 64-workgroup teams for NP = 3072)
    struct Cls { int np = 0, nmax = 1; int64_t hb = 0; int first_inst = 0; std::vector<TeamItem> items; };
    std::vector<Cls> cls;
    int R = 0, npmax = 0, nmax_all = 1;
    for (int gi = 0; gi < n_groups; ++gi) {
        if (inst[gi] < 0 || inst[gi] >= int(c->inst.size())) return fail(ANC_ERR_STATE, "group %d: instance %d not registered", gi, inst[gi]);
        if (count[gi] < 0) return fail(ANC_ERR_ARG, "group %d: negative count", gi);
        const Instance& in = c->inst[size_t(inst[gi])];
        if (!team_size_for(in.np)) return fail(ANC_ERR_UNSUPPORTED, "group %d: NP = %d has no team shape", gi, in.np);
        Cls* k = nullptr;
        for (auto& x : cls) if (x.np == in.np) k = &x;
        if (!k) { cls.emplace_back(); k = &cls.back(); k->np = in.np; k->first_inst = inst[gi]; }
        k->nmax = std::max(k->nmax, in.n);
        k->hb = std::max(k->hb, in.hbound);
        for (int q = 0; q < count[gi]; q += rt) {
            TeamItem it{};
            it.Jp = in.Jp; it.bias = in.bias; it.n = in.n; it.r0 = R + q; it.count = std::min(rt, count[gi] - q);
            k->items.push_back(it);
        }
        R += count[gi];
        npmax = std::max(npmax, in.np);
        nmax_all = std::max(nmax_all, in.n);
    }
    if (R < 1) return fail(ANC_ERR_ARG, "anc_run_multi: no replicas");
    if (!out->spins || !out->sum_sh || !out->sum_bs || !out->flips) return fail(ANC_ERR_ARG, "null output array");
    if (out->spin_stride < npmax / 64) return fail(ANC_ERR_ARG, "spin_stride too small");
    if (out->fields && out->field_stride < nmax_all) return fail(ANC_ERR_ARG, "field_stride < n");
    if (int rc = set_device(c)) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    std::vector<TeamItem> all;
    for (auto& k : cls) all.insert(all.end(), k.items.begin(), k.items.end());
    const size_t bytes = sizeof(TeamItem) * all.size();
    if (bytes > c->items_bytes) {
        if (c->d_items) ANC_HIP(hipFree(c->d_items));
        c->d_items = nullptr;
        c->items_bytes = 0;
        ANC_HIP(hipMalloc(&c->d_items, bytes));
        c->items_bytes = bytes;
    }
    ANC_HIP(hipMemcpy(c->d_items, all.data(), bytes, hipMemcpyHostToDevice));
    if (zero_out(*out, R, st)) return fail(ANC_ERR_HIP, "fill launch failed");
    size_t first = 0;
    for (auto& k : cls) {
        if (int rc = run_team(c, c->inst[size_t(k.first_inst)], seed, trials, 0, R, nt, c->d_items + first, int(k.items.size()), k.nmax,
                              k.hb, k.np, *out, st, false))
            return rc;
        first += k.items.size();
    }
    return ANC_OK;
}

int anc_sync_status(anc_ctx* c, void* stream) {
    clear_error();
    if (!c) return fail(ANC_ERR_ARG, "null context");
    if (int rc = set_device(c)) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    ANC_HIP(hipStreamSynchronize(st));
    int s[2] = {0, 0};
    ANC_HIP(hipMemcpy(s, c->ws + kFragWords + kNlinWords, sizeof s, hipMemcpyDeviceToHost));
    if (s[0] == 1) {
        const int was = c->xflavor;
        c->xflavor = 1;
        return fail(ANC_ERR_HIP, "team exchange timed out (flavour %d)%s", was,
                    was != 1 ? "; the context now uses the device-scope exchange" : "");
    }
    if (s[0]) return fail(ANC_ERR_HIP, "device status %d", s[0]);
    return ANC_OK;
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
    if (!std::strcmp(key, "profile")) {
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
