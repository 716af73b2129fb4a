/*
 * anechoic_mi355x -- C ABI of the MI355X (gfx950) SCA annealing engine with Onsager echo correction.
 *
 * Library: anechoic_amd/_lib/libanechoic_mi355x.so. Plain C types only (no torch types); every status is an int
 * (0 = ok, negative = error) and anc_last_error() returns the thread's last message. No C++ exception crosses the ABI.
 *
 * Semantics are those of the reference golden model (fpga/v80_sca/src/sca_ref_bias.hpp:25-75 run_trial_bias):
 * per-anneal xoshiro128** streams on 256 lanes seeded from (seed, trial, lane), ROUNDS = NP/256 draws per lane per
 * step, Q16.16 threshold decision, n_lin band count, counted (lambda) or temperature-scaled (kappa) Onsager
 * correction, integer field update h += J (s_new - s_old). Results are bit-identical to the golden model on the same
 * (J, bias, tables, seed, trial). NP(n) = max(2048, 256 ceil(n/256)); for n <= 2048 this is the reference's NP = 2048.
 *
 * Memory: J, bias, trial ids and all outputs are DEVICE pointers (caller-owned, e.g. torch tensors); schedules and
 * tables are host arrays. A context owns its device copies of the tables, the engine's internal instance layouts and
 * its exchange workspace; it is not thread-safe, and separate contexts are independent.
 */
#ifndef ANECHOIC_MI355X_H
#define ANECHOIC_MI355X_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ANC_OK 0
#define ANC_ERR_ARG (-1)          /* bad argument (null pointer, size out of range) */
#define ANC_ERR_INSTANCE (-2)     /* J not symmetric, nonzero diagonal, or field bound too large */
#define ANC_ERR_RANGE (-3)        /* tables outside the exact integer range of the engine */
#define ANC_ERR_HIP (-4)          /* HIP runtime error */
#define ANC_ERR_UNSUPPORTED (-5)  /* shape not supported by the requested kernel */
#define ANC_ERR_STATE (-6)        /* missing tables or instance */
#define ANC_ERR_IO (-7)           /* instance or bias file cannot be opened, read or written, or is malformed */

#define ANC_SHAPE_GEOMETRIC 0     /* T_t = t0 (t1/t0)^(t/max(1,S-1)) (reference default) */
#define ANC_SHAPE_LINEAR 1        /* T_t = t0 + (t1-t0) t/max(1,S-1) */
#define ANC_SHAPE_EXPLICIT 2      /* T_t given */

#define ANC_MAX_STEPS 65536
#define ANC_MAX_N 16384

/* Schedule -> tables. Replaces fpga/v80_sca/src/mb_common.hpp:132-156 (mb::Sched, mb::tables; scale folded into
 * lambda or kappa), gpu/onsager_v2_20261007/src/tables.hpp:18-40 (Schedule, build_tables; scale = 1) and
 * fpga/v80_sca/src/sca_ref.hpp:57-68 (make_tables on an explicit T vector). */
typedef struct anc_sched {
    int32_t S;                 /* steps */
    int32_t shape;             /* ANC_SHAPE_* */
    double t0, t1;             /* first/last temperature (geometric, linear) */
    double q;                  /* decision offset */
    double lam;                /* counted Onsager strength lambda (0 = off) */
    double jv;                 /* TEC constant (kconst = -jv) */
    double kappa;              /* temperature-scaled correction (> 0 replaces lambda/jv) */
    double scale;              /* correction scale gamma (1 = reference v6.4 tables) */
    int32_t ramp;              /* linear end ramp of the correction over the last 30% of steps */
    const double* T;           /* ANC_SHAPE_EXPLICIT: S temperatures (host) */
} anc_sched;

/* Fills host arrays fourT, q, kconst (int32[S]) and kcorr (int64[S]). */
int anc_make_tables(const anc_sched* sched, int32_t* fourT, int32_t* q, int64_t* kcorr, int32_t* kconst);

/* Padded slot count NP(n) used for the RNG lane/round structure (sca_ref.hpp:14-21, generalised). */
int anc_np(int n);

const char* anc_last_error(void);
int anc_version(void);

typedef struct anc_ctx anc_ctx;

/* Context on a HIP device. Replaces the reference runner's set-up (gpu/onsager_v2_20261007/src/sca_gpu.cu:563-583,
 * Launcher::setup). NULL on failure (see anc_last_error). */
anc_ctx* anc_create(int device);
void anc_destroy(anc_ctx* ctx);

/* Validates host tables and keeps a device copy in the context. Replaces the table half of kern::Params
 * (sca_gpu.cu:47-63) and the host range guard check_ranges/plan_ranges (sca_gpu.cu:519-528,
 * gpu/multibit_bias_20261008/src/sca_gpu_bias.cu:543-557). */
int anc_set_tables(anc_ctx* ctx, int S, const int32_t* fourT, const int32_t* q, const int64_t* kcorr, const int32_t* kconst);

/* Registers an instance: J is n x n int8 (row stride ldj >= n) and bias n int32 or NULL, both device pointers that
 * must stay valid while the instance is used. The instance is validated on the device (symmetric, zero diagonal) and
 * its field bound max_y sum_x |J_xy| + max |b| is computed; the engine's internal layouts are built from it.
 * Returns an instance id >= 0. Replaces jmat::finish (gpu/multibit_bias_20261008/src/jmat.hpp:35-55) and the device
 * upload of the coupling fragments build_A (sca_gpu.cu:494-516, sca_gpu_bias.cu:510-534). */
int anc_add_instance(anc_ctx* ctx, int n, const int8_t* J, int64_t ldj, const int32_t* bias, void* stream);
int anc_clear_instances(anc_ctx* ctx);
/* Field bound of an instance (>= 0), or a negative status. */
int64_t anc_instance_bound(anc_ctx* ctx, int inst);

/* Synthetic instance on the device: gen_ising random (gpu/multibit_bias_20261008/src/gen_ising.cpp:118-135, pbig >= 1):
 * K-bit couplings (|J| <= 2^(kbits-1) - 1, zero with probability pzero) and biases in [-bmax, bmax], bit-identical to
 * the reference generator for every n (the reference caps n at 2048). J: n x ldj int8 (device; the padding columns
 * are zeroed), bias: n int32 (device) or NULL. No context needed. */
int anc_gen_ising(uint64_t seed, int n, int kbits, double pzero, int bmax, int8_t* J, int64_t ldj, int32_t* bias, void* stream);

/* Caller-allocated device outputs. Replica r writes spins[r * spin_stride ...] (NP(n)/64 words, bit i = 1 iff s_i = +1,
 * as sca::ResultBias::spins), sum_sh[r] = sum_i s_i h_i, sum_bs[r] = sum_i b_i s_i, flips[r], and, when non-NULL,
 * nlin[r * S + t] (the per-step n_lin trace) and fields[r * field_stride + i] (final h_i, i < n). */
typedef struct anc_out {
    uint64_t* spins;
    int64_t spin_stride;
    int64_t* sum_sh;
    int64_t* sum_bs;
    int64_t* flips;
    int32_t* nlin;
    int32_t* fields;
    int64_t field_stride;
} anc_out;

/* Plain kernel: ONE anneal (trial) of instance inst per launch, on a team of workgroups that splits the spin slots.
 * Replaces sca::run_trial_bias (fpga/v80_sca/src/sca_ref_bias.hpp:25-26) and the reference's B=1 launch
 * (sca_gpu.cu:584-592). stream: hipStream_t or NULL. */
int anc_run_plain(anc_ctx* ctx, int inst, uint64_t seed, uint32_t trial, const anc_out* out, void* stream);

/* Batched engine: R anneals of instance inst (trial ids trials[0..R), device; NULL = 0..R-1) in one persistent launch
 * (up to 8 teams of 32 workgroups, each team working through 16- or 32-replica items); the field
 * update is an exact int8 MFMA contraction over all replicas. Replaces the reference's batch launch of chains with
 * trial = trial_offset + chain (sca_gpu.cu:192, :584-592, :700). */
int anc_run_batched(anc_ctx* ctx, int inst, uint64_t seed, const uint32_t* trials, int R, const anc_out* out, void* stream);

/* anc_run_batched without per-call device allocation: the same kernels and results, with the team exchange workspace in a
 * static device array of the library (3.19 MiB per device, resolved at anc_create) instead of the context's. No
 * counterpart in the reference: it is for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc); the context, its tables and instances are set up beforehand.
 * n <= 3072 only (larger instances return ANC_ERR_UNSUPPORTED: their exchange exceeds the array). Every context on a
 * device shares that device's array, and each call leaves its contents behind (the next call clears what must start
 * at zero): calls that overlap on one device race. */
int anc_run_batched_noalloc(anc_ctx* ctx, int inst, uint64_t seed, const uint32_t* trials, int R, const anc_out* out,
                            void* stream);

/* Uneven multi-instance batch in one persistent launch: group g anneals count[g] replicas of instance inst[g] (host
 * arrays); output rows follow the group order, trial ids trials[row] (device; NULL = row). Instances with n <= 2048
 * run on 32-workgroup teams (n <= 1024: 4-round teams, option small_kr4), 2048 < n <= 3072 on 32-workgroup 12-round
 * teams (64-workgroup teams with the flip exchange); when several classes are present one launch hosts them, either
 * in turn on every XCD or on disjoint XCDs (options mixed_ratio, mixed_xcd2). Waits for earlier work on `stream` (the work-item table is reused). No single
 * counterpart in the reference (its kernels take one instance per launch, with runtime n:
 * gpu/multibit_bias_20261008/src/sca_gpu_bias.cu:204-207); semantics per replica are run_trial_bias. */
int anc_run_multi(anc_ctx* ctx, int n_groups, const int32_t* inst, const int32_t* count, uint64_t seed, const uint32_t* trials,
                  const anc_out* out, void* stream);

/* Captures the anc_run_multi launch for these groups (workspace and output clears by fill kernels, then the
 * persistent team launch) as a hipGraph owned by the context, with its own copies of the work-item table and of the
 * current tables. trials (device, may be NULL) and the out arrays are referenced: every replay reads the trial ids
 * and writes the outputs anew. Returns a graph id >= 0 or a negative error code. Replays of one context share its
 * exchange workspace, so they must not overlap with each other or with eager runs (use one stream).
 * anc_clear_instances and anc_destroy free the context's graphs. No counterpart in the reference (its host enqueues
 * one kernel per batch, sca_gpu.cu:584-592; the cached graph is the boundary's context-owned state). */
int anc_graph_multi(anc_ctx* ctx, int n_groups, const int32_t* inst, const int32_t* count, uint64_t seed,
                    const uint32_t* trials, const anc_out* out);
/* Replays a captured run on stream (hipGraphLaunch). */
int anc_graph_launch(anc_ctx* ctx, int graph, void* stream);
/* Frees a captured run (waits for the device first). */
int anc_graph_destroy(anc_ctx* ctx, int graph);

/* Waits for the stream and reports the device-side status of the launches since the last call: the first error any
 * of them raised (exchange time-out) is kept on the device until this call reports it, so a later clean launch
 * cannot hide it. The outputs of a launch that timed out are not valid. Replaces the reference runner's post-launch
 * error check (sca_gpu.cu:708-713). */
int anc_sync_status(anc_ctx* ctx, void* stream);

/* Context options: "xflavor" (3 = XCD-local exchange through L2, the default; 2 = its variant with device-scope stores;
 * 1 = device-scope exchange), "fallback" (0, the default = an exchange time-out is only reported; 1 = the report also
 * switches the context to xflavor 1), "timeout_s" (exchange wait limit), "stall_test" (1 = every team launch waits
 * for a word nothing sets until timeout_s and reports a time-out: tests of the error path; 0 = off), "profile" (1 =
 * record per-phase cycle counters; cannot change while captured runs exist), "multi_nt" (1 or 2 sixteen-replica
 * tiles per work item of the multi-instance runs of n <= 2048 instances; default 2), "multi_nt3" (1 or 2 sixteen-
 * replica tiles per work item of n = 2049..3072 instances with the spin exchanges, in multi-instance runs and as the
 * upper bound of anc_run_batched's tile; default 2), "mixed_ratio" (cost of a work-item step of an n = 2049..3072
 * instance relative to one of an n <= 2048 instance, used to split the XCDs between the two team shapes; default
 * 2.0), "mixed_xcd2" (0 = split by mixed_ratio; else the XCDs forced for the n <= 2048 teams: 1..7, or 2, 4 or 6 with
 * "exchange" 0, whose n = 2049..3072 teams take XCD pairs; or 8 with the spin exchanges: every XCD's team runs its
 * share of both shapes' items in turn), "small_kr4" (1, the default = with the spin exchanges, the multi-instance
 * work items of n <= 1024 instances run on 4-round teams, J' laid out with 1024 slots, in turn before the n <= 2048
 * items on the same teams; 0 = they run as n <= 2048 items), "exchange" (team
 * kernels: 2, the default = every step exchanges the spins, one tagged 2-bit crumb per slot (the 12-round teams of
 * n = 2049..3072 use nibbles), and recomputes the field b + J s with int32 accumulators; 1 = the same with one tagged
 * nibble per slot; 0 = exchanges the flip matrix {-2, 0, 2} as tagged bytes and accumulates h += J ds, the
 * reference's incremental update; identical results), "batched_team" (anc_run_batched, n <= 2048: 16 = 16-member
 * teams with 16-replica items, 32 = 32-member teams; 0, the default = 16 when the instance admits it (crumb exchange,
 * field bound), else 32), "batched_nt" (32-member teams: 1 or 2 sixteen-replica tiles per item; 0 = the default rule),
 * "large_tile" (n > 3072: 128 or 64 replicas per work item of the large kernel; 0, the default = 64 when 128-replica
 * items would leave teams idle). Every launch shape gives identical results.
 * No counterpart in the reference (its CLI flags select kernels at build time). */
int anc_set_option(anc_ctx* ctx, const char* key, double value);
/* Reads an option (the keys above, plus "last_xcd2": XCDs given to 32-workgroup teams by the last anc_run_multi,
 * "last_wide": 1 if the last launch decided with int64 thresholds, chosen when int32 cannot be proven exact,
 * "last_team": team size of the last team launch, "last_nt3": sixteen-replica tiles per work item of the
 * n = 2049..3072 items of the last team launch (0 if it had none), "last_kr4": 1 if the last team launch ran
 * 4-round teams, and "last_large_tile": replicas per work item of the
 * last large-kernel launch).
 * Returns a negative error code for an unknown key. */
double anc_get_option(anc_ctx* ctx, const char* key);

/* Copies the per-workgroup phase counters of the last team launch (8 per workgroup: MFMA phase incl. exchange wait,
 * n_lin gather, LDS + barrier, field assembly, decisions, publish, steps, team*1000+member) into out; returns the
 * number of values. Replaces the reference's SCA_PROFILE clock64 counters (sca_gpu.cu:163-168, :751-758). */
int anc_get_profile(anc_ctx* ctx, long long* out, int count);

/* ---- Differentiable mean-field relaxation (build-defined; no reference counterpart; the closest reference code is
 * the DMFT effective process research/dmft_sca_20261004/dmft.py:16-26, :102-120 and the Onsager term of
 * paper/sections/algorithm.tex:80-95). B instances of n spins, dense J (instance b at J + b*bstride, row stride ldj
 * bytes, a multiple of 16, n * ldj < 2^31) stored as int8 (jbits = 8; ldj >= n), packed by anc_mf_pack_j4 (jbits = 4;
 * ldj >= 8 ceil(n/16)) or by anc_mf_pack_j2 (jbits = 2; ldj >= 4 ceil(n/16)); all storages give identical results.
 * Each product J v is computed exactly on the int8 matrix cores from a block floating-point quantisation of v
 * (q = rint(v 2^s), s = 30 - e with max|v| in [2^(e-1), 2^e)) and rounded once to FP32. FP32 magnetisations, all
 * device pointers:
 *   u_t = J m_t + q m_t - beta_t m_{t-1};  m_{t+1} = tanh(u_t / (2 T_t));
 *   beta_t = coef_t (counted = 0) or coef_t * sum_i (1 - m_{t,i}^2) (counted = 1); beta_0 = 0.
 * Forward stores (m_{t0-1}, m_{t0}) every ckpt steps in ckpt_buf (2 ceil(S/ckpt) B n floats) and m_S in m_final.
 * Backward takes g_S = dL/dm_S and returns dT[t*B+b] (direct dL/dT_t) and dcoef[t*B+b] (dL/dcoef_t) per instance (FP64) and,
 * if g0 is not NULL, dL/dm_0. work: anc_mf_workspace(B, n, ckpt) bytes. */
size_t anc_mf_workspace(int B, int n, int ckpt);
int anc_mf_forward(int B, int n, const int8_t* J, int jbits, int64_t ldj, int64_t bstride, float q, int S, const float* T,
                   const float* coef, int counted, const float* m0, int ckpt, float* ckpt_buf, float* m_final, float* work,
                   void* stream);
int anc_mf_backward(int B, int n, const int8_t* J, int jbits, int64_t ldj, int64_t bstride, float q, int S, const float* T,
                    const float* coef, int counted, int ckpt, const float* ckpt_buf, const float* gS, double* dT, double* dcoef,
                    float* g0, float* work, void* stream);
/* Persistent team kernels: for n <= 2048, a multiple of 4, on a device with >= 256 CUs, anc_mf_forward and
 * anc_mf_backward run each checkpoint segment in one launch (32-workgroup teams on one XCD each, two teams per CU, J
 * kept in registers, the vector exchanged through the XCD's L2 every step) instead of one launch per step; results
 * are identical. Their waits are bounded:
 * anc_mf_status synchronises the stream and returns ANC_ERR_HIP if a wait of the run (since the last anc_mf_forward
 * on this workspace) timed out, ANC_OK otherwise. Process-wide options (no reference counterpart): "team" (1, the
 * default: team kernels where they apply; 0: one launch per step), "timeout_s" (bound of each wait, default 5) and
 * "stall_test" (1: every team workgroup waits on a word nothing sets until the bound, then records the time-out and
 * exits; for tests of the error path). anc_mf_get_option("last_team") is 1 when the last forward or backward call ran
 * the team kernels. */
int anc_mf_status(float* work, int B, int n, int ckpt, void* stream);
int anc_mf_set_option(const char* key, double value);
double anc_mf_get_option(const char* key);   /* -1 for an unknown key */
/* y_b = J_b x_b for every instance (energies and loss gradients). */
int anc_mf_gemv(int B, int n, const int8_t* J, int jbits, int64_t ldj, int64_t bstride, const float* x, float* y, void* stream);
/* Packed 4-bit copy J4 (row stride ldj4, instance stride bstride4) of int8 J for jbits = 4: in each 4-byte group g
 * of a row, byte e holds column 8g+e in its low nibble and column 8g+4+e in its high nibble (two's complement);
 * columns >= n are 0. Entries outside [-8, 7] are added to the device int *bad (zeroed by the caller; the copy is
 * unusable when it reads nonzero). Halves the bytes the mean-field kernels stream per step. */
int anc_mf_pack_j4(int B, int n, const int8_t* J, int64_t ldj, int64_t bstride, uint8_t* J4, int64_t ldj4, int64_t bstride4,
                   int* bad, void* stream);
/* Two-bit copy J2 for jbits = 2 (entries in [-2, 1], e.g. +-1 couplings): in each 4-byte group g of a row (columns
 * 16g .. 16g+15), bits 2k and 2k+1 of byte e hold column 16g+4k+e (two's complement); columns >= n are 0. Entries
 * outside [-2, 1] are added to *bad as for anc_mf_pack_j4. A quarter of the int8 bytes per step. */
int anc_mf_pack_j2(int B, int n, const int8_t* J, int64_t ldj, int64_t bstride, uint8_t* J2, int64_t ldj2, int64_t bstride2,
                   int* bad, void* stream);

/* ---- Instance files, instance statistics and independent scorers. Host memory only; these calls use no GPU. */
#define ANC_FMT_JINT8 0     /* "SCAJINT8", uint32 n, uint32 0, n*n int8 row-major */
#define ANC_FMT_SNOWGPU1 1  /* "SNOWGPU1", uint32 n, n rows of ceil(n/64) uint64 LE words; bit y of row x = 1 <=> J_xy = +1,
                               else -1 (x != y); diagonal bits ignored */
#define ANC_FMT_GSET 2      /* text "n m", then m lines "i j w" (1-indexed); J_ij = J_ji = -w */

/* Reads an instance file into J (host, n rows of stride ldj >= n, columns n..ldj-1 zeroed). Call with J = NULL to
 * read the header only (*n_out = n), then again with the matrix. The matrix must have a zero diagonal and be
 * symmetric. Replaces gpu/multibit_bias_20261008/src/jmat.hpp:57-103 (from_k2000, from_gset, from_jint8, load) and
 * fpga/v80_sca/src/sca_ref.hpp:31-49 (load_graph); n up to ANC_MAX_N. */
int anc_read_instance(const char* path, int format, int32_t* n_out, int8_t* J, int64_t ldj);
/* Writes J (validated as above) in the given format. Replaces jmat.hpp:105-113 (write_jint8) and the SNOWGPU1 writer
 * data/fetch_data.py:38-54 (k2000_bin); the G-set writer emits every pair x < y with J_xy != 0 as w = -J_xy. */
int anc_write_instance(const char* path, int format, int32_t n, const int8_t* J, int64_t ldj);
/* SCABIAS1 files: "SCABIAS1", uint32 n, uint32 0, n int32 with |b| <= 32767; the file's n must equal n. Replace
 * jmat.hpp:164-186 (load_bias, write_bias). */
int anc_read_bias(const char* path, int32_t n, int32_t* bias);
int anc_write_bias(const char* path, int32_t n, const int32_t* bias);

/* Statistics of jmat::finish (jmat.hpp:35-55): sumw = sum_{x<y} -J_xy (cut = (sumw + sum_sh/2)/2), max |J|, field bound
 * max_x sum_y |J_xy|, mean over rows of sum_y J_xy^2, and the number of nonzero entries. Validates J as above. */
typedef struct anc_jstats {
    int64_t sumw;
    int32_t maxabs;
    int64_t hmax;
    double mean_row_sq;
    int64_t nnz;
} anc_jstats;
int anc_instance_stats(int32_t n, const int8_t* J, int64_t ldj, anc_jstats* st);

/* Independent scorers of R replicas (host; spins as 64-bit words, bit i = 1 <=> s_i = +1, spin_stride words per
 * replica), recomputed from J and the spins: cut[r] = sum_{x<y, s_x != s_y} -J_xy (jmat.hpp:130-139 scalar_cut),
 * energy[r] = -sum_{x<y} J_xy s_x s_y - sum_x b_x s_x (:141-151 energy), mismatches[r] = number of x with
 * b_x + sum_y J_xy s_y != fields[r*field_stride + x] (:153-162 field_mismatches). bias may be NULL (zero); any output
 * may be NULL; fields are needed only for mismatches. nthreads 0 = all cores. */
int anc_score(int32_t n, const int8_t* J, int64_t ldj, const int32_t* bias, int32_t R, const uint64_t* spins,
              int64_t spin_stride, const int32_t* fields, int64_t field_stride, int64_t* cut, int64_t* energy,
              int64_t* mismatches, int nthreads);

/* ---- Problem encodings as Ising instances (host memory; no GPU). H(s) = -sum_{i<j} J_ij s_i s_j - sum_i b_i s_i, with
 * x = (1 + s)/2; C = H(all -1) - sigma f(all x = 0) is returned, so that H = sigma f + C on every state.
 * Graph partitioning (sigma = 2), replaces `gen_ising gp` (gpu/multibit_bias_20261008/src/gen_ising.cpp:145-204):
 * random deg-regular graph (pairing model), f = B cut + A (sum x - k)^2, J_ij = -A + B [ij in E], b_i = A (2k - n).
 * J: n rows of stride ldj; edges (optional, n deg / 2 pairs u < v, sorted). */
int anc_gen_gp(uint64_t seed, int32_t n, int32_t deg, int32_t k, int64_t A, int64_t B, int8_t* J, int64_t ldj,
               int32_t* bias, int32_t* edges, int64_t* C);
/* TSP (sigma = 4), replaces `gen_ising tsp` (gen_ising.cpp:205-283): distances 1..wmax, city 0 fixed at position 0,
 * spin (v-1)(cities-1) + (p-1) = city v at position p; (cities-1)^2 spins. W (optional): cities x cities distances. */
int anc_gen_tsp(uint64_t seed, int32_t cities, int32_t wmax, int64_t A, int64_t B, int8_t* J, int64_t ldj, int32_t* bias,
                int32_t* W, int64_t* C);

#ifdef __cplusplus
}
#endif
#endif
