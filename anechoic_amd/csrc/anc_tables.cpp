// Host-side schedule -> fixed-point table builder and exact-range planner of the engine.
//
// Tables (reference: fpga/v80_sca/src/sca_ref.hpp:57-68 make_tables; fpga/v80_sca/src/mb_common.hpp:132-156 mb::tables;
// gpu/onsager_v2_20261007/src/tables.hpp:14-40 build_tables):
//   fourT[t]  = llround(4 T_t 2^16)                     (Q16.16 of 4T)
//   q[t]      = llround(q 2^16)
//   kcorr[t]  = t == 0 ? 0 : llround(lambda_t / (2 T_{t-1}) 2^24)      (Q8.24, counted Onsager correction)
//   kconst[t] = t == 0 ? 0 : llround(-jv 2^16)                          (TEC constant)
//   temperature-scaled mode (kappa > 0): kconst[t] = llround(kappa scale T_{t-1} rho_t 2^16), kcorr[t] = 0.
//   lambda_t = lambda rho_t scale, rho_t = 1 for t < 0.7 S else (S-1-t)/max(1, S-1-0.7 S) when the ramp is on.
// The double-precision expressions are evaluated in the same order as the reference so that every rounding matches.
#include "anc_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

namespace anc {

int temperatures(const anc_sched& c, std::vector<double>& T) {
    const int S = c.S;
    if (S < 1 || S > ANC_MAX_STEPS) return fail(ANC_ERR_ARG, "schedule: S must be in 1..%d (got %d)", ANC_MAX_STEPS, S);
    T.resize(size_t(S));
    const int den = std::max(1, S - 1);
    switch (c.shape) {
        case ANC_SHAPE_GEOMETRIC:
            if (!(c.t0 > 0) || !(c.t1 > 0)) return fail(ANC_ERR_ARG, "schedule: geometric needs t0 > 0 and t1 > 0");
            for (int t = 0; t < S; ++t) T[size_t(t)] = c.t0 * std::pow(c.t1 / c.t0, double(t) / den);
            break;
        case ANC_SHAPE_LINEAR:
            for (int t = 0; t < S; ++t) T[size_t(t)] = c.t0 + (c.t1 - c.t0) * (double(t) / den);
            break;
        case ANC_SHAPE_EXPLICIT:
            if (!c.T) return fail(ANC_ERR_ARG, "schedule: explicit shape needs a T array");
            for (int t = 0; t < S; ++t) T[size_t(t)] = c.T[t];
            break;
        default:
            return fail(ANC_ERR_ARG, "schedule: unknown shape %d", c.shape);
    }
    for (int t = 0; t < S; ++t)
        if (!(T[size_t(t)] > 0) || 4.0 * T[size_t(t)] * 65536.0 >= 2147483647.0)
            return fail(ANC_ERR_RANGE, "schedule: T[%d] = %g outside (0, 8192)", t, T[size_t(t)]);
    return ANC_OK;
}

int make_tables(const anc_sched& c, int32_t* fourT, int32_t* q, int64_t* kcorr, int32_t* kconst) {
    std::vector<double> T;
    if (int rc = temperatures(c, T)) return rc;
    const int S = c.S;
    if (std::fabs(c.q) * 65536.0 >= 2147483647.0 || std::fabs(c.jv) * 65536.0 >= 2147483647.0)
        return fail(ANC_ERR_RANGE, "schedule: q or jv outside the Q16.16 int32 range");
    std::vector<double> L(size_t(S), c.lam);
    if (c.ramp)
        for (int t = 0; t < S; ++t)
            if (t >= 0.7 * S) L[size_t(t)] = c.lam * (S - 1 - t) / std::max(1.0, S - 1 - 0.7 * S);
    if (c.scale != 1.0)
        for (double& l : L) l *= c.scale;
    for (int t = 0; t < S; ++t) {
        fourT[t] = int32_t(std::llround(4.0 * T[size_t(t)] * 65536.0));
        q[t] = int32_t(std::llround(c.q * 65536.0));
        kcorr[t] = t == 0 ? 0 : int64_t(std::llround(L[size_t(t)] / (2.0 * T[size_t(t) - 1]) * 16777216.0));
        kconst[t] = t == 0 ? 0 : int32_t(std::llround(-c.jv * 65536.0));
    }
    if (c.kappa > 0)
        for (int t = 1; t < S; ++t) {
            const double rf = (c.ramp && t >= 0.7 * S) ? (S - 1 - t) / std::max(1.0, S - 1 - 0.7 * S) : 1.0;
            const double v = c.kappa * c.scale * T[size_t(t) - 1] * rf * 65536.0;
            if (std::fabs(v) >= 2147483647.0) return fail(ANC_ERR_RANGE, "schedule: kappa T at step %d outside int32", t);
            kconst[t] = int32_t(std::llround(v));
            kcorr[t] = 0;
        }
    return ANC_OK;
}

// Exact-range plan. The device decides with v = s h (int32) against per-step thresholds derived from
// z = v 2^16 + c, c = q -/+ corr, corr = ((n_lin_prev kcorr) >> 8) + kconst (reference: sca_ref.hpp:119, :127-130).
// The fast path keeps c and the band/threshold arithmetic in int32; it needs, for every step and every reachable
// n_lin in [0, nmax]:  |q| + |corr| + fourT / 2 + 2 < 2^31  and  fourT > 0.
// (The reference GPU kernel proves the analogous bound with gpu/onsager_v2_20261007/src/sca_gpu.cu:519-528 and
// gpu/multibit_bias_20261008/src/sca_gpu_bias.cu:543-557; this engine never forms v 2^16, so |h| only has to fit int32.)
int plan(const Tables& tb, int nmax, int64_t hbound, Plan& p) {
    p = Plan{};
    const int64_t lim = (int64_t(1) << 31) - 1;
    if (hbound < 0 || hbound > lim / 2) return fail(ANC_ERR_RANGE, "field bound %lld does not fit int32", (long long)hbound);
    p.wide = false;
    p.corr24 = true;
    for (size_t t = 0; t < tb.fourT.size(); ++t) {
        const int64_t kh = tb.kcorr[t] >> 8;   // floor, as the kernels' split n (kc >> 8) + ((n (kc & 255)) >> 8)
        if (kh >= (int64_t(1) << 23) || kh < -(int64_t(1) << 23) || int64_t(nmax) * std::llabs(kh) >= (int64_t(1) << 31))
            p.corr24 = false;
        if (tb.fourT[t] <= 0) return fail(ANC_ERR_RANGE, "fourT[%zu] must be positive", t);
        const int64_t kc = tb.kcorr[t];
        if (kc > (int64_t(1) << 40) || kc < -(int64_t(1) << 40)) return fail(ANC_ERR_RANGE, "kcorr[%zu] out of range", t);
        const int64_t c0 = int64_t(tb.kconst[t]);
        const int64_t c1 = ((int64_t(nmax) * kc) >> 8) + int64_t(tb.kconst[t]);
        const int64_t corr_max = std::max(std::llabs(c0), std::llabs(c1));
        const int64_t m = std::llabs(int64_t(tb.q[t])) + corr_max + int64_t(tb.fourT[t]) / 2 + 2;
        if (m >= lim) p.wide = true;
        if (m >= (int64_t(1) << 46)) return fail(ANC_ERR_RANGE, "tables at step %zu exceed the int64 decision range", t);
    }
    if (p.wide) p.corr24 = false;
    p.ok = true;
    return ANC_OK;
}

}  // namespace anc

extern "C" int anc_make_tables(const anc_sched* s, int32_t* fourT, int32_t* q, int64_t* kcorr, int32_t* kconst) {
    if (!s || !fourT || !q || !kcorr || !kconst) return anc::fail(ANC_ERR_ARG, "anc_make_tables: null argument");
    return anc::make_tables(*s, fourT, q, kcorr, kconst);
}

extern "C" int anc_np(int n) { return anc::padded_slots(n); }
