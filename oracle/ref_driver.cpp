// ORACLE -- test infrastructure only.
// A thin C-callable driver around the reference's own golden-model headers, compiled from the reference sources where
// they lie (see oracle/Makefile; the build needs the reference sources, so it happens in the development container).
// It writes golden vectors (tests/golden/make_golden.py), cross-checks the C restatement (oracle/golden.c) and, as
// the built oracle/_ref/libref.so that travels with the working tree, is the host-CPU baseline that bench.py times
// and compares with the GPU. It is never part of the product library.
//
// Entry points wrap: sca::run_trial_bias (fpga/v80_sca/src/sca_ref_bias.hpp:25), sca::run_trial (sca_ref.hpp:101),
// mb::tables (fpga/v80_sca/src/mb_common.hpp:141), v80host::build_tables (in ref_tables.cpp),
// sca::make_tables (sca_ref.hpp:57), mb::random_k (mb_common.hpp:81), mb3::random_k (mb_common3.hpp:88),
// mb::fingerprint (mb_common.hpp:183), mb::j2_scale (mb_common.hpp:123), mb3::j2_scale.
#include "sca_ref_bias.hpp"
#include "mb_common.hpp"
#include "mb_common3.hpp"

#include <cstring>
#include <exception>

namespace {
void put_tables(const sca::Tables& tb, int32_t* fourT, int32_t* q, int64_t* kcorr, int32_t* kconst) {
    for (size_t t = 0; t < tb.fourT.size(); ++t) {
        fourT[t] = tb.fourT[t]; q[t] = tb.q[t]; kcorr[t] = tb.kcorr[t]; kconst[t] = tb.kconst[t];
    }
}
sca::Tables get_tables(int S, const int32_t* fourT, const int32_t* q, const int64_t* kcorr, const int32_t* kconst) {
    sca::Tables tb;
    tb.fourT.assign(fourT, fourT + S); tb.q.assign(q, q + S); tb.kcorr.assign(kcorr, kcorr + S); tb.kconst.assign(kconst, kconst + S);
    return tb;
}
}  // namespace

extern "C" {

int ref_np() { return sca::NP; }
int ref_lanes() { return sca::LANES; }

// mb::tables(Sched): geometric schedule with scale.
int ref_mb_tables(int S, double t0, double t1, double q, double lam, double jv, double kappa, double scale, int ramp,
                  int32_t* fourT, int32_t* qo, int64_t* kcorr, int32_t* kconst) {
    mb::Sched c; c.S = S; c.t0 = t0; c.t1 = t1; c.q = q; c.lam = lam; c.jv = jv; c.kappa = kappa; c.scale = scale; c.ramp = ramp != 0;
    put_tables(mb::tables(c), fourT, qo, kcorr, kconst);
    return 0;
}

// sca::make_tables on explicit per-step T, q, lambda vectors.
int ref_make_tables(int S, const double* T, const double* q, const double* lam, double jv,
                    int32_t* fourT, int32_t* qo, int64_t* kcorr, int32_t* kconst) {
    put_tables(sca::make_tables(std::vector<double>(T, T + S), std::vector<double>(q, q + S), std::vector<double>(lam, lam + S), jv),
               fourT, qo, kcorr, kconst);
    return 0;
}

// sca::run_trial_bias. dense: NP x NP int8 (reference layout). bias: NP entries or NULL. spins: NP/64 words.
int ref_run_trial_bias(int n, const int8_t* dense, const int32_t* bias, int S, const int32_t* fourT, const int32_t* q,
                       const int64_t* kcorr, const int32_t* kconst, uint64_t seed, uint32_t trial, uint64_t* spins,
                       int64_t* sum_sh, int64_t* sum_bs, int64_t* flips, int32_t* nlin) {
    try {
        const sca::ResultBias r = sca::run_trial_bias(n, dense, bias, get_tables(S, fourT, q, kcorr, kconst), seed, trial, nlin != nullptr);
        std::memcpy(spins, r.spins.data(), sizeof(uint64_t) * sca::WORDS);
        *sum_sh = r.sum_sh; *sum_bs = r.sum_bs; *flips = r.flips;
        if (nlin) std::memcpy(nlin, r.n_lin.data(), sizeof(int32_t) * size_t(S));
        return 0;
    } catch (const std::exception&) { return -1; }
}

// sca::run_trial on the dense path (N = 2000, b = 0). sumw is the instance's sum_{i<j} -J_ij.
int ref_run_trial(const int8_t* dense, int64_t sumw, int S, const int32_t* fourT, const int32_t* q, const int64_t* kcorr,
                  const int32_t* kconst, uint64_t seed, uint32_t trial, uint64_t* spins, int64_t* cut, int64_t* flips,
                  int32_t* nlin) {
    sca::Graph g; g.sumw = sumw;
    const sca::Result r = sca::run_trial(g, get_tables(S, fourT, q, kcorr, kconst), seed, trial, nlin != nullptr, dense);
    std::memcpy(spins, r.spins.data(), sizeof(uint64_t) * sca::WORDS);
    *cut = r.cut; *flips = r.flips;
    if (nlin) std::memcpy(nlin, r.n_lin.data(), sizeof(int32_t) * size_t(S));
    return 0;
}

// mb::random_k (N = 2000): writes the NP x NP dense matrix; returns maxabs.
int ref_mb_random_k(uint64_t seed, double density, int maxabs, int8_t* dense, int64_t* sumw, uint64_t* fingerprint, double* j2) {
    try {
        const mb::Matrix m = mb::random_k(seed, density, maxabs);
        std::memcpy(dense, m.J.data(), m.J.size());
        *sumw = m.sumw; *fingerprint = mb::fingerprint(m); *j2 = mb::j2_scale(m);
        return m.maxabs;
    } catch (const std::exception&) { return -1; }
}

// mb3::random_k(seed, density, maxabs, n, bmax): dense matrix plus NP biases.
int ref_mb3_random_k(uint64_t seed, double density, int maxabs, int n, int bmax, int8_t* dense, int32_t* bias,
                     int64_t* sumw, uint64_t* fingerprint, double* j2) {
    try {
        const mb3::Matrix m = mb3::random_k(seed, density, maxabs, n, bmax);
        std::memcpy(dense, m.J.data(), m.J.size());
        std::memcpy(bias, m.bias.data(), sizeof(int32_t) * m.bias.size());
        *sumw = m.sumw; *fingerprint = mb3::fingerprint(m); *j2 = mb3::j2_scale(m);
        return m.maxabs;
    } catch (const std::exception&) { return -1; }
}

}  // extern "C"
