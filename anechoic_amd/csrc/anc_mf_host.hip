// Host side of the mean-field relaxation: forward with checkpoints every `ckpt` steps, adjoint with recomputation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "anc_internal.hpp"
#include "anc_kernels.hpp"

using namespace anc;

namespace {

int check_layout(int n, int jbits, int64_t ldj, int64_t bstride) {
    if (jbits != 8 && jbits != 4 && jbits != 2) return fail(ANC_ERR_ARG, "mean-field: jbits must be 8 (int8 J), 4 or 2 (packed J)");
    const int64_t row = 2 * int64_t(jbits) * ((n + 15) / 16);
    if (ldj % 16 || ldj < row) return fail(ANC_ERR_ARG, "mean-field: ldj must be a multiple of 16 covering the row");
    if (bstride < int64_t(n) * ldj) return fail(ANC_ERR_ARG, "mean-field: instance stride smaller than n * ldj");
    // rows are read through 32-bit buffer offsets
    if (int64_t(n) * ldj >= (int64_t(1) << 31)) return fail(ANC_ERR_ARG, "mean-field: instance beyond 2 GiB");
    return ANC_OK;
}

int check_args(int B, int n, const int8_t* J, int jbits, int64_t ldj, int64_t bstride, int S, const float* T, const float* coef) {
    // B is the grid's second dimension
    if (B < 1 || B > 65535 || n < 1 || n > ANC_MAX_N || !J || !T || !coef) return fail(ANC_ERR_ARG, "mean-field: bad B, n or null pointer");
    if (int rc = check_layout(n, jbits, ldj, bstride)) return rc;
    if (S < 1 || S > ANC_MAX_STEPS) return fail(ANC_ERR_ARG, "mean-field: S out of range");
    return ANC_OK;
}

MFArgs make_args(int B, int n, const int8_t* J, int jbits, int64_t ldj, int64_t bstride, float q, const float* T, const float* coef,
                 int counted) {
    MFArgs a{};
    a.J = J; a.jbits = jbits; a.ldj = ldj; a.bstride = bstride; a.n = n; a.B = B; a.q = q; a.T = T; a.coef = coef;
    a.counted = counted ? 1 : 0;
    return a;
}

int pack(int bits, int B, int n, const int8_t* J, int64_t ldj, int64_t bstride, uint8_t* Jp, int64_t ldp, int64_t bstridep,
         int* bad, void* stream) {
    clear_error();
    if (B < 1 || B > 65535 || n < 1 || n > ANC_MAX_N || !J || !Jp || !bad) return fail(ANC_ERR_ARG, "anc_mf_pack: bad argument");
    if (int rc = check_layout(n, 8, ldj, bstride)) return rc;
    if (int rc = check_layout(n, bits, ldp, bstridep)) return rc;
    if (launch_mf_pack(bits, B, n, J, ldj, bstride, Jp, ldp, bstridep, bad, static_cast<hipStream_t>(stream)))
        return fail(ANC_ERR_HIP, "anc_mf_pack launch failed");
    return ANC_OK;
}

// Process-wide settings of the mean-field path (anc_mf_set_option): the persistent team kernels where they apply
// (n <= 2048 on >= 256 CUs; 1, the default) or the one-launch-per-step kernels (0), and the bound of the team
// kernels' waits.
int g_team = 1;
double g_timeout_s = 5.0;
int g_stall_test = 0;
int g_last_team = 0;   // whether the last forward or backward ran the team kernels

struct DevInfo {
    int cus = 0;
    long long wall_khz = 100000;
};

int dev_info(DevInfo& di) {
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess || d < 0 || d >= 64) return fail(ANC_ERR_HIP, "mean-field: no current device");
    static DevInfo cache[64];
    static bool have[64] = {};
    if (!have[d]) {
        int cus = 0, khz = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d) != hipSuccess) return fail(ANC_ERR_HIP, "mean-field: device query failed");
        cache[d].cus = cus;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, d) == hipSuccess && khz > 0) cache[d].wall_khz = khz;
        have[d] = true;
    }
    di = cache[d];
    return ANC_OK;
}

// Workspace: the per-step buffers (floats), then, for n <= 2048, a 256-byte status block, ckpt exchange slots of the
// team kernels (B x 2048 words each) and the adjoint's instance-sum partials of ckpt steps (B x 64 doubles each).
size_t base_bytes(int B, int n, int ckpt) {
    const size_t vec = size_t(B) * size_t(n);
    // segment m (ckpt + 1 vectors), u tape (ckpt), 2 g, 2 delta, beta tape (ckpt x B) + beta hold (B)
    const size_t b = sizeof(float) * (vec * (size_t(ckpt) + 1 + size_t(ckpt) + 4) + size_t(B) * (size_t(ckpt) + 1));
    return (b + 255) / 256 * 256;
}
bool team_area(int n) { return n <= 2048; }

MFTeamArgs team_args(float* work, int B, int n, int ckpt, const DevInfo& di) {
    MFTeamArgs x{};
    uint8_t* p = reinterpret_cast<uint8_t*>(work) + base_bytes(B, n, ckpt);
    x.status = reinterpret_cast<int*>(p);
    x.xch = reinterpret_cast<uint32_t*>(p + 256);
    x.part = reinterpret_cast<double*>(p + 256 + sizeof(uint32_t) * size_t(ckpt) * mf_team_slot_words(B));
    x.timeout_ticks = (long long)(g_timeout_s * double(di.wall_khz) * 1000.0);
    x.stall_test = g_stall_test;
    return x;
}

bool use_team(int n, const DevInfo& di) { return g_team && team_area(n) && mf_team_fits(n, di.cus); }

}  // namespace

extern "C" {

size_t anc_mf_workspace(int B, int n, int ckpt) {
    size_t b = base_bytes(B, n, ckpt);
    if (team_area(n))
        b += 256 + sizeof(uint32_t) * size_t(ckpt) * mf_team_slot_words(B) + sizeof(double) * size_t(ckpt) * mf_team_part_doubles(B);
    return b;
}

int anc_mf_set_option(const char* key, double value) {
    clear_error();
    if (!key) return fail(ANC_ERR_ARG, "null option key");
    if (!std::strcmp(key, "team")) {
        g_team = value != 0.0 ? 1 : 0;
        return ANC_OK;
    }
    if (!std::strcmp(key, "timeout_s")) {
        if (!(value > 0.0)) return fail(ANC_ERR_ARG, "timeout_s must be positive");
        g_timeout_s = value;
        return ANC_OK;
    }
    if (!std::strcmp(key, "stall_test")) {
        if (value != 0.0 && value != 1.0) return fail(ANC_ERR_ARG, "stall_test must be 0 or 1");
        g_stall_test = int(value);
        return ANC_OK;
    }
    return fail(ANC_ERR_ARG, "unknown mean-field option '%s'", key);
}

double anc_mf_get_option(const char* key) {
    if (key && !std::strcmp(key, "team")) return g_team;
    if (key && !std::strcmp(key, "timeout_s")) return g_timeout_s;
    if (key && !std::strcmp(key, "stall_test")) return g_stall_test;
    if (key && !std::strcmp(key, "last_team")) return g_last_team;
    return -1.0;
}

int anc_mf_status(float* work, int B, int n, int ckpt, void* stream) {
    clear_error();
    if (!work || B < 1 || n < 1 || ckpt < 1) return fail(ANC_ERR_ARG, "anc_mf_status: bad argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    ANC_HIP(hipStreamSynchronize(st));
    if (!team_area(n)) return ANC_OK;
    int* status = reinterpret_cast<int*>(reinterpret_cast<uint8_t*>(work) + base_bytes(B, n, ckpt));
    int code = 0;
    ANC_HIP(hipMemcpy(&code, status + 2, sizeof code, hipMemcpyDeviceToHost));
    if (!code) return ANC_OK;
    ANC_HIP(hipMemset(status + 2, 0, sizeof(int)));
    return fail(ANC_ERR_HIP, "mean-field team exchange timed out (error %d)", code);
}

int anc_mf_forward(int B, int n, const int8_t* J, int jbits, int64_t ldj, int64_t bstride, float q, int S, const float* T, const float* coef,
                   int counted, const float* m0, int ckpt, float* ckpt_buf, float* m_final, float* work, void* stream) {
    clear_error();
    if (int rc = check_args(B, n, J, jbits, ldj, bstride, S, T, coef)) return rc;
    if (!m0 || !ckpt_buf || !m_final || !work || ckpt < 1) return fail(ANC_ERR_ARG, "mean-field forward: null buffer or ckpt < 1");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const MFArgs a = make_args(B, n, J, jbits, ldj, bstride, q, T, coef, counted);
    const size_t vec = size_t(B) * size_t(n);
    float* seg = work;
    const int nseg = (S + ckpt - 1) / ckpt;
    DevInfo di;
    if (int rc = dev_info(di)) return rc;
    const bool team = use_team(n, di);
    g_last_team = team ? 1 : 0;
    const MFTeamArgs x = team_area(n) ? team_args(work, B, n, ckpt, di) : MFTeamArgs{};
    // a forward pass starts a run: its status block starts clear, whichever kernels run (anc_mf_status reports the
    // first error since)
    if (team_area(n)) {
        FillRanges f{};
        f.add(x.status, 32);
        if (launch_fill_zero_ranges(f, st)) return fail(ANC_ERR_HIP, "fill launch failed");
    }
    // checkpoint k holds (m_{t0-1}, m_{t0}), t0 = k ckpt; m_{-1} = m_0 (unused: beta_0 = 0)
    ANC_HIP(hipMemcpyAsync(ckpt_buf, m0, vec * sizeof(float), hipMemcpyDeviceToDevice, st));
    ANC_HIP(hipMemcpyAsync(ckpt_buf + vec, m0, vec * sizeof(float), hipMemcpyDeviceToDevice, st));
    for (int k = 0; k < nseg; ++k) {
        const int t0 = k * ckpt, t1 = std::min(S, t0 + ckpt);
        const float* cp = ckpt_buf + size_t(2 * k) * vec;
        const int lrc = team ? launch_mf_team_forward(a, x, cp, cp + vec, t0, t1, seg, nullptr, nullptr, st)
                             : launch_mf_forward(a, cp, cp + vec, t0, t1, seg, nullptr, nullptr, st);
        if (lrc) return fail(ANC_ERR_HIP, "mf forward launch failed");
        const int len = t1 - t0;
        if (k + 1 < nseg) {
            float* nx = ckpt_buf + size_t(2 * (k + 1)) * vec;
            ANC_HIP(hipMemcpyAsync(nx, seg + size_t(len - 1) * vec, vec * sizeof(float), hipMemcpyDeviceToDevice, st));
            ANC_HIP(hipMemcpyAsync(nx + vec, seg + size_t(len) * vec, vec * sizeof(float), hipMemcpyDeviceToDevice, st));
        } else {
            ANC_HIP(hipMemcpyAsync(m_final, seg + size_t(len) * vec, vec * sizeof(float), hipMemcpyDeviceToDevice, st));
        }
    }
    return ANC_OK;
}

int anc_mf_backward(int B, int n, const int8_t* J, int jbits, int64_t ldj, int64_t bstride, float q, int S, const float* T, const float* coef,
                    int counted, int ckpt, const float* ckpt_buf, const float* gS, double* dT, double* dcoef, float* g0, float* work,
                    void* stream) {
    clear_error();
    if (int rc = check_args(B, n, J, jbits, ldj, bstride, S, T, coef)) return rc;
    if (!ckpt_buf || !gS || !dT || !dcoef || !work || ckpt < 1) return fail(ANC_ERR_ARG, "mean-field backward: null buffer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const MFArgs a = make_args(B, n, J, jbits, ldj, bstride, q, T, coef, counted);
    const size_t vec = size_t(B) * size_t(n);
    float* seg = work;                                   // ckpt + 1 vectors
    float* utape = seg + (size_t(ckpt) + 1) * vec;       // ckpt vectors
    float* gbuf = utape + size_t(ckpt) * vec;            // 2 vectors
    float* dbuf = gbuf + 2 * vec;                        // 2 vectors
    float* btape = dbuf + 2 * vec;                       // ckpt x B
    float* bhold = btape + size_t(ckpt) * B;             // B
    const int nseg = (S + ckpt - 1) / ckpt;
    const float* g_cur = gS;
    const float* delta_next = nullptr;
    const float* beta_next = nullptr;
    int gsel = 0, dsel = 0;
    DevInfo di;
    if (int rc = dev_info(di)) return rc;
    const bool team = use_team(n, di);
    g_last_team = team ? 1 : 0;
    const MFTeamArgs x = team ? team_args(work, B, n, ckpt, di) : MFTeamArgs{};
    for (int k = nseg - 1; k >= 0; --k) {
        const int t0 = k * ckpt, t1 = std::min(S, t0 + ckpt);
        const float* cp = ckpt_buf + size_t(2 * k) * vec;
        if (k + 1 < nseg) {   // beta_{t1} lives in the tape about to be overwritten
            ANC_HIP(hipMemcpyAsync(bhold, btape, sizeof(float) * B, hipMemcpyDeviceToDevice, st));
            beta_next = bhold;
        }
        if (team) {
            // the segment's tapes, then its adjoint steps in one launch
            if (launch_mf_team_forward(a, x, cp, cp + vec, t0, t1, seg, utape, btape, st)) return fail(ANC_ERR_HIP, "mf recompute failed");
            float* g_out = gbuf + size_t(gsel) * vec;
            float* d_out = dbuf + size_t(dsel) * vec;
            if (launch_mf_team_backward(a, x, t0, t1, seg, utape, btape, cp, g_cur, delta_next, delta_next ? beta_next : nullptr, g_out,
                                        d_out, dT, dcoef, st))
                return fail(ANC_ERR_HIP, "mf backward launch failed");
            g_cur = g_out;
            delta_next = d_out;
            gsel ^= 1;
            dsel ^= 1;
            continue;
        }
        if (launch_mf_forward(a, cp, cp + vec, t0, t1, seg, utape, btape, st)) return fail(ANC_ERR_HIP, "mf recompute failed");
        for (int t = t1 - 1; t >= t0; --t) {
            const int kk = t - t0;
            const float* m_prev = kk > 0 ? seg + size_t(kk - 1) * vec : cp;
            const float* bnext = (t + 1 < t1) ? btape + size_t(kk + 1) * B : beta_next;
            float* g_out = gbuf + size_t(gsel) * vec;
            float* d_out = dbuf + size_t(dsel) * vec;
            if (launch_mf_bwd_step(a, t, g_cur, seg + size_t(kk + 1) * vec, seg + size_t(kk) * vec, m_prev, utape + size_t(kk) * vec,
                                   delta_next, delta_next ? bnext : nullptr, d_out, g_out, dT, dcoef, st))
                return fail(ANC_ERR_HIP, "mf backward launch failed");
            g_cur = g_out;
            delta_next = d_out;
            gsel ^= 1;
            dsel ^= 1;
        }
    }
    if (g0) ANC_HIP(hipMemcpyAsync(g0, g_cur, vec * sizeof(float), hipMemcpyDeviceToDevice, st));
    return ANC_OK;
}

int anc_mf_gemv(int B, int n, const int8_t* J, int jbits, int64_t ldj, int64_t bstride, const float* x, float* y, void* stream) {
    clear_error();
    if (B < 1 || B > 65535 || n < 1 || n > ANC_MAX_N || !J || !x || !y) return fail(ANC_ERR_ARG, "anc_mf_gemv: bad argument");
    if (int rc = check_layout(n, jbits, ldj, bstride)) return rc;
    float dummy = 0.f;
    const MFArgs a = make_args(B, n, J, jbits, ldj, bstride, 0.f, &dummy, &dummy, 0);
    if (launch_mf_gemv(a, x, y, static_cast<hipStream_t>(stream))) return fail(ANC_ERR_HIP, "anc_mf_gemv launch failed");
    return ANC_OK;
}

int anc_mf_pack_j4(int B, int n, const int8_t* J, int64_t ldj, int64_t bstride, uint8_t* J4, int64_t ldj4, int64_t bstride4,
                   int* bad, void* stream) {
    return pack(4, B, n, J, ldj, bstride, J4, ldj4, bstride4, bad, stream);
}

int anc_mf_pack_j2(int B, int n, const int8_t* J, int64_t ldj, int64_t bstride, uint8_t* J2, int64_t ldj2, int64_t bstride2,
                   int* bad, void* stream) {
    return pack(2, B, n, J, ldj, bstride, J2, ldj2, bstride2, bad, stream);
}

}  // extern "C"
