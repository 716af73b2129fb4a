/*
 * ORACLE -- test infrastructure only.
 *
 * A plain-C restatement of the reference's integer SCA golden model, used as the checker for the HIP engine.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only to check or to time
 * the CPU baseline; the product library (anechoic_amd) never links or calls it.
 *
 * Followed reference code (paths relative to the reference repository):
 *   RNG (splitmix64 seeding, xoshiro128**)      fpga/v80_sca/src/sca_ref.hpp:71-91
 *   tables (make_tables, geom, ramp, kappa)      fpga/v80_sca/src/sca_ref.hpp:57-68, fpga/v80_sca/src/mb_common.hpp:132-156,
 *                                               gpu/onsager_v2_20261007/src/tables.hpp:14-40
 *   one anneal with bias and runtime n           fpga/v80_sca/src/sca_ref_bias.hpp:25-75 (dense path of sca_ref.hpp:101-155)
 *   instance generators                          mb_common.hpp:81-96 (random_k), mb_common3.hpp:88-109 (random_k with n, bias),
 *                                               asic/gf22_engine_v64_mb_20261008/sim/src/gen_vectors_mbb.cpp:18-34, :128-147,
 *                                               gpu/multibit_bias_20261008/src/gen_ising.cpp:118-144
 *   fingerprint (FNV-1a), correction scale       mb_common.hpp:123-128, :183-187; mb_common3.hpp j2_scale
 *
 * Generalisation beyond the reference (which is compiled for NP = 2048): the padded slot count is
 * NP(n) = max(2048, 256 * ceil(n / 256)), 256 RNG lanes, ROUNDS = NP / 256 draws per lane per step, slot i = k * 256 + l.
 * For n <= 2048 this is the reference exactly.
 *
 * Pinning: every function is checked in tests/ against the compiled reference headers (oracle/_ref, built in the
 * development container) through committed golden vectors, against the reference-held cases.txt known answers, and
 * against closed forms / invariants (field recomputation, energy identity). See the test files for the mapping.
 */
#ifndef ANC_ORACLE_GOLDEN_H
#define ANC_ORACLE_GOLDEN_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_LANES 256

int orc_np(int n);                                   /* padded slot count NP(n) */

uint64_t orc_splitmix(uint64_t* x);
void orc_rng_seed(uint32_t s[4], uint64_t seed, uint32_t trial, uint32_t lane);
uint32_t orc_rng_next(uint32_t s[4]);

/* Schedule. shape: 0 = geometric t0 -> t1 (reference default), 1 = linear t0 -> t1, 2 = explicit T (Texp, S values). */
typedef struct {
    int S;
    int shape;
    double t0, t1, q, lam, jv, kappa, scale;
    int ramp;
    const double* Texp;
} orc_sched;

/* Per-step tables exactly as mb::tables (scale folded into lambda or kappa). Returns 0, or -1 on bad input. */
int orc_tables(const orc_sched* c, int32_t* fourT, int32_t* q, int64_t* kcorr, int32_t* kconst);

/* One anneal (run_trial_bias semantics, generalised NP). J: n x n used, row stride ldj (J[x*ldj + y]), zero diagonal.
 * bias: n entries or NULL. spins: NP(n)/64 words (bit 1 = +1). nlin: S entries or NULL. h_out: n entries or NULL. */
int orc_run_trial(int n, const int8_t* J, int64_t ldj, const int32_t* bias, int S, const int32_t* fourT, const int32_t* q,
                  const int64_t* kcorr, const int32_t* kconst, uint64_t seed, uint32_t trial, uint64_t* spins,
                  int64_t* sum_sh, int64_t* sum_bs, int64_t* flips, int32_t* nlin, int32_t* h_out);

/* R independent anneals on host threads (nthreads <= 0: all online cores). Output arrays are indexed by replica:
 * spins R x NP(n)/64, sum_sh/sum_bs/flips R, nlin R x S (or NULL). */
int orc_run_batch(int n, const int8_t* J, int64_t ldj, const int32_t* bias, int S, const int32_t* fourT, const int32_t* q,
                  const int64_t* kcorr, const int32_t* kconst, uint64_t seed, const uint32_t* trials, int R, int nthreads,
                  uint64_t* spins, int64_t* sum_sh, int64_t* sum_bs, int64_t* flips, int32_t* nlin);

/* Generators (dense row-major, stride ldj, caller zero-fills). */
/* mb::random_k restricted to the first n spins (gen_vectors_mbb random_kn; n = 2000 gives mb::random_k). If bmax > 0,
 * appends mb_common3 uniform biases in [-bmax, bmax] from the same stream into bias (n entries). */
void orc_gen_random_k(uint64_t seed, double density, int maxabs, int n, int bmax, int8_t* J, int64_t ldj, int32_t* bias);
/* gen_ising random (K bits, pzero, bias bmax with pbig = 1). */
void orc_gen_ising_random(uint64_t seed, int n, int kbits, double pzero, int64_t bmax, int8_t* J, int64_t ldj, int32_t* bias);
/* gen_vectors_mbb bias suite vector (pure function of the matrix fingerprint). */
void orc_gen_mbb_bias(uint64_t fingerprint, int maxabs, int n, int32_t* bias);
/* FNV-1a 64 over the NP x NP (np = 2048 in the reference) dense matrix. */
uint64_t orc_fingerprint(const int8_t* J, int64_t ldj, int np);
/* sum_{x,y < nsum} J^2 / div / div (mb::j2_scale: nsum = div = 2000; mb3::j2_scale: nsum = div = n). */
double orc_j2_scale(const int8_t* J, int64_t ldj, int nsum, int div);

/* Independent scorers. */
int64_t orc_sumw(const int8_t* J, int64_t ldj, int n);                     /* sum_{i<j} -J_ij */
void orc_fields(const int8_t* J, int64_t ldj, int n, const int32_t* bias, const uint64_t* spins, int64_t* h);

#ifdef __cplusplus
}
#endif
#endif
