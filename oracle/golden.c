/*
 * ORACLE -- test infrastructure only (the checker for the HIP engine; never linked into the product).
 * Plain-C restatement of the reference golden model. Each function cites the reference code it follows; the
 * header golden.h lists the mapping and the generalisation to NP > 2048.
 */
#define _GNU_SOURCE
#include "golden.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

int orc_np(int n) {
    int np = ((n + ORC_LANES - 1) / ORC_LANES) * ORC_LANES;
    return np < 2048 ? 2048 : np;
}

/* splitmix64 step: fpga/v80_sca/src/sca_ref.hpp:73-78 */
uint64_t orc_splitmix(uint64_t* x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Per-lane seeding: sca_ref.hpp:79-84 (trial bits >= 28 alias lane bits, reproduced as is). */
void orc_rng_seed(uint32_t s[4], uint64_t seed, uint32_t trial, uint32_t lane) {
    uint64_t x = seed ^ ((uint64_t)trial << 20) ^ ((uint64_t)lane << 48) ^ 0x5CA0F00Dull;
    uint64_t a = orc_splitmix(&x), b = orc_splitmix(&x);
    s[0] = (uint32_t)a; s[1] = (uint32_t)(a >> 32); s[2] = (uint32_t)b; s[3] = (uint32_t)(b >> 32);
    if (!(s[0] | s[1] | s[2] | s[3])) s[0] = 1;
}

static inline uint32_t rotl32(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }

/* xoshiro128** (Blackman & Vigna), as sca_ref.hpp:86-90. */
uint32_t orc_rng_next(uint32_t s[4]) {
    const uint32_t r = rotl32(s[1] * 5u, 7) * 9u, t = s[1] << 9;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl32(s[3], 11);
    return r;
}

/* Tables: sca_ref.hpp:57-68 (make_tables) with the schedule rules of mb_common.hpp:141-156 (ramp, scale, kappa). */
int orc_tables(const orc_sched* c, int32_t* fourT, int32_t* q, int64_t* kcorr, int32_t* kconst) {
    const int S = c->S;
    if (S < 1) return -1;
    double* T = (double*)malloc(sizeof(double) * (size_t)S);
    double* L = (double*)malloc(sizeof(double) * (size_t)S);
    if (!T || !L) { free(T); free(L); return -1; }
    const int den = S - 1 > 1 ? S - 1 : 1;
    for (int t = 0; t < S; ++t) {
        if (c->shape == 0) T[t] = c->t0 * pow(c->t1 / c->t0, (double)t / den);          /* mb_common.hpp:136-140 */
        else if (c->shape == 1) T[t] = c->t0 + (c->t1 - c->t0) * ((double)t / den);      /* linear (build-defined) */
        else T[t] = c->Texp[t];
    }
    for (int t = 0; t < S; ++t) {
        L[t] = c->lam;
        if (c->ramp && t >= 0.7 * S) L[t] = c->lam * (S - 1 - t) / fmax(1.0, S - 1 - 0.7 * S);
        if (c->scale != 1.0) L[t] *= c->scale;
    }
    for (int t = 0; t < S; ++t) {
        fourT[t] = (int32_t)llround(4.0 * T[t] * 65536.0);
        q[t] = (int32_t)llround(c->q * 65536.0);
        kcorr[t] = t == 0 ? 0 : (int64_t)llround(L[t] / (2.0 * T[t - 1]) * 16777216.0);
        kconst[t] = t == 0 ? 0 : (int32_t)llround(-c->jv * 65536.0);
    }
    if (c->kappa > 0)
        for (int t = 1; t < S; ++t) {
            const double rf = (c->ramp && t >= 0.7 * S) ? (S - 1 - t) / fmax(1.0, S - 1 - 0.7 * S) : 1.0;
            kconst[t] = (int32_t)llround(c->kappa * c->scale * T[t - 1] * rf * 65536.0);
            kcorr[t] = 0;
        }
    free(T); free(L);
    return 0;
}

/* One anneal: fpga/v80_sca/src/sca_ref_bias.hpp:25-75 with NP = orc_np(n). */
int orc_run_trial(int n, const int8_t* J, int64_t ldj, const int32_t* bias, int S, const int32_t* fourT, const int32_t* q,
                  const int64_t* kcorr, const int32_t* kconst, uint64_t seed, uint32_t trial, uint64_t* spins,
                  int64_t* sum_sh, int64_t* sum_bs, int64_t* flips, int32_t* nlin_out, int32_t* h_out) {
    if (n < 1 || ldj < n) return -1;
    const int np = orc_np(n), rounds = np / ORC_LANES;
    uint32_t (*rng)[4] = (uint32_t(*)[4])malloc(sizeof(uint32_t) * 4 * ORC_LANES);
    int8_t* s = (int8_t*)malloc((size_t)np);
    int8_t* sp = (int8_t*)malloc((size_t)np);
    int32_t* h = (int32_t*)calloc((size_t)np, sizeof(int32_t));
    int* flipped = (int*)malloc(sizeof(int) * (size_t)np);
    if (!rng || !s || !sp || !h || !flipped) { free(rng); free(s); free(sp); free(h); free(flipped); return -1; }
    for (int l = 0; l < ORC_LANES; ++l) orc_rng_seed(rng[l], seed, trial, (uint32_t)l);
    memset(s, 1, (size_t)np); memset(sp, 1, (size_t)np);
    /* initial spins: sca_ref_bias.hpp:31-36 (padding draws consumed) */
    for (int k = 0; k < rounds; ++k)
        for (int l = 0; l < ORC_LANES; ++l) {
            const int i = k * ORC_LANES + l;
            const uint32_t r = orc_rng_next(rng[l]);
            if (i < n) s[i] = (r >> 31) ? 1 : -1;
        }
    /* initial fields: sca_ref_bias.hpp:37-41 */
    for (int y = 0; y < n; ++y) h[y] = bias ? bias[y] : 0;
    for (int x = 0; x < n; ++x) {
        const int8_t* row = J + (size_t)x * (size_t)ldj;
        const int sx = s[x];
        for (int y = 0; y < n; ++y)
            if (y != x) h[y] += (int32_t)row[y] * sx;
    }
    int32_t nlin_prev = 0;
    int64_t nflips = 0;
    for (int t = 0; t < S; ++t) {
        /* correction, decision and n_lin band: sca_ref_bias.hpp:44-56 */
        const int64_t corr = (((int64_t)nlin_prev * kcorr[t]) >> 8) + kconst[t];
        const int32_t f4 = fourT[t], twoT = f4 >> 1;
        int32_t nlin = 0;
        int nf = 0;
        for (int k = 0; k < rounds; ++k)
            for (int l = 0; l < ORC_LANES; ++l) {
                const int i = k * ORC_LANES + l;
                const uint32_t u16 = orc_rng_next(rng[l]) >> 16;
                if (i >= n) continue;
                const int64_t z = ((int64_t)s[i] * h[i]) * 65536 + q[t] - (s[i] == sp[i] ? corr : -corr);
                const int64_t r = (((int64_t)u16 - 32768) * f4) >> 16;
                if (z > -twoT && z < twoT) ++nlin;
                if (z < r) flipped[nf++] = i;
            }
        /* commit and field update: sca_ref_bias.hpp:57-64 */
        memcpy(sp, s, (size_t)n);
        for (int f = 0; f < nf; ++f) s[flipped[f]] = (int8_t)-s[flipped[f]];
        for (int f = 0; f < nf; ++f) {
            const int x = flipped[f];
            const int32_t d = 2 * s[x];
            const int8_t* row = J + (size_t)x * (size_t)ldj;
            for (int y = 0; y < n; ++y) h[y] += d * row[y];
        }
        nflips += nf;
        nlin_prev = nlin;
        if (nlin_out) nlin_out[t] = nlin;
    }
    /* outputs: sca_ref_bias.hpp:69-73 */
    int64_t sh = 0, bs = 0;
    memset(spins, 0, sizeof(uint64_t) * (size_t)(np / 64));
    for (int i = 0; i < n; ++i) {
        sh += (int64_t)s[i] * h[i];
        if (bias) bs += (int64_t)bias[i] * s[i];
        if (s[i] > 0) spins[i >> 6] |= (uint64_t)1 << (i & 63);
    }
    if (h_out) memcpy(h_out, h, sizeof(int32_t) * (size_t)n);
    *sum_sh = sh; *sum_bs = bs; *flips = nflips;
    free(rng); free(s); free(sp); free(h); free(flipped);
    return 0;
}

typedef struct {
    int n; const int8_t* J; int64_t ldj; const int32_t* bias; int S;
    const int32_t *fourT, *q, *kconst; const int64_t* kcorr;
    uint64_t seed; const uint32_t* trials; int R;
    uint64_t* spins; int64_t *sum_sh, *sum_bs, *flips; int32_t* nlin;
    int next; int err;
} batch_job;

static void* batch_worker(void* arg) {
    batch_job* b = (batch_job*)arg;
    const int words = orc_np(b->n) / 64;
    for (;;) {
        const int r = __atomic_fetch_add(&b->next, 1, __ATOMIC_RELAXED);
        if (r >= b->R) break;
        const int rc = orc_run_trial(b->n, b->J, b->ldj, b->bias, b->S, b->fourT, b->q, b->kcorr, b->kconst, b->seed,
                                     b->trials[r], b->spins + (size_t)r * words, b->sum_sh + r, b->sum_bs + r, b->flips + r,
                                     b->nlin ? b->nlin + (size_t)r * b->S : NULL, NULL);
        if (rc) __atomic_store_n(&b->err, rc, __ATOMIC_RELAXED);
    }
    return NULL;
}

/* Replica fan-out over host threads with an atomic work counter (the pattern of gpu/onsager_v2_20261007/src/check_ref.cpp:48-66). */
int orc_run_batch(int n, const int8_t* J, int64_t ldj, const int32_t* bias, int S, const int32_t* fourT, const int32_t* q,
                  const int64_t* kcorr, const int32_t* kconst, uint64_t seed, const uint32_t* trials, int R, int nthreads,
                  uint64_t* spins, int64_t* sum_sh, int64_t* sum_bs, int64_t* flips, int32_t* nlin) {
    batch_job b = {n, J, ldj, bias, S, fourT, q, kconst, kcorr, seed, trials, R, spins, sum_sh, sum_bs, flips, nlin, 0, 0};
    if (nthreads <= 0) nthreads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (nthreads > R) nthreads = R;
    if (nthreads < 1) nthreads = 1;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    if (!th) return -1;
    for (int i = 1; i < nthreads; ++i) pthread_create(&th[i], NULL, batch_worker, &b);
    batch_worker(&b);
    for (int i = 1; i < nthreads; ++i) pthread_join(th[i], NULL);
    free(th);
    return b.err;
}

/* mb::random_k / gen_vectors_mbb random_kn / mb_common3 random_k (mb_common.hpp:81-96, gen_vectors_mbb.cpp:18-34,
 * mb_common3.hpp:88-109). The threshold expression is kept as written, including its double rounding. */
void orc_gen_random_k(uint64_t seed, double density, int maxabs, int n, int bmax, int8_t* J, int64_t ldj, int32_t* bias) {
    uint64_t st = seed;
    const uint64_t thr = (uint64_t)(density * 18446744073709551616.0 - 1.0);
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            const uint64_t a = orc_splitmix(&st), b = orc_splitmix(&st);
            if (a > thr) continue;
            const int mag = 1 + (int)(b % (uint64_t)maxabs);
            const int v = ((b >> 32) & 1) ? mag : -mag;
            J[(size_t)i * ldj + j] = (int8_t)v; J[(size_t)j * ldj + i] = (int8_t)v;
        }
    if (bmax > 0 && bias)
        for (int i = 0; i < n; ++i) bias[i] = (int32_t)(orc_splitmix(&st) % (uint64_t)(2 * bmax + 1)) - bmax;
}

/* gen_ising random: gpu/multibit_bias_20261008/src/gen_ising.cpp:118-135 (pbig = 1). */
void orc_gen_ising_random(uint64_t seed, int n, int kbits, double pzero, int64_t bmax, int8_t* J, int64_t ldj, int32_t* bias) {
    const int M = (1 << (kbits - 1)) - 1;
    uint64_t st = seed ^ 0xA5A5F00D12345678ull, sb = seed ^ 0x5B1A5B1A5B1A5B1Aull;
    const double lim = ldexp(pzero, 64);
    const uint64_t zthr = lim >= 18446744073709551615.0 ? ~0ull : (uint64_t)lim;
    for (int x = 0; x < n; ++x)
        for (int y = x + 1; y < n; ++y) {
            const uint64_t r0 = orc_splitmix(&st), r1 = orc_splitmix(&st);
            int v = 0;
            if (!(pzero > 0 && r0 < zthr)) {
                const int mag = 1 + (int)((r1 >> 1) % (uint64_t)M);
                v = (r1 & 1) ? -mag : mag;
            }
            J[(size_t)x * ldj + y] = (int8_t)v; J[(size_t)y * ldj + x] = (int8_t)v;
        }
    if (bias)
        for (int i = 0; i < n; ++i)
            bias[i] = bmax > 0 ? (int32_t)((int64_t)(orc_splitmix(&sb) % (uint64_t)(2 * bmax + 1)) - bmax) : 0;
}

/* gen_vectors_mbb.cpp:128-140 bias suite vector. */
void orc_gen_mbb_bias(uint64_t fingerprint, int maxabs, int n, int32_t* bias) {
    uint64_t st = 0xB1A5ull ^ fingerprint;
    const int br = maxabs > 1 ? 300 : 10;
    for (int i = 0; i < n; ++i) {
        const uint64_t a = orc_splitmix(&st), b = orc_splitmix(&st);
        int v;
        if (a % 100 < 3) v = 16384 + (int)(b % 16384u);
        else v = (int)(b % (uint64_t)(2 * br + 1)) - br;
        if ((a >> 40) & 1) v = -v;
        bias[i] = v;
    }
}

/* mb_common.hpp:183-187 */
uint64_t orc_fingerprint(const int8_t* J, int64_t ldj, int np) {
    uint64_t h = 1469598103934665603ull;
    for (int x = 0; x < np; ++x)
        for (int y = 0; y < np; ++y) { h ^= (uint8_t)J[(size_t)x * ldj + y]; h *= 1099511628211ull; }
    return h;
}

/* mb_common.hpp:123-128 (nsum = div = 2000) and mb_common3.hpp j2_scale (nsum = div = n). */
double orc_j2_scale(const int8_t* J, int64_t ldj, int nsum, int div) {
    double s = 0;
    for (int x = 0; x < nsum; ++x)
        for (int y = 0; y < nsum; ++y) { const double v = J[(size_t)x * ldj + y]; s += v * v; }
    return s / div / div;
}

int64_t orc_sumw(const int8_t* J, int64_t ldj, int n) {
    int64_t s = 0;
    for (int x = 0; x < n; ++x)
        for (int y = x + 1; y < n; ++y) s -= J[(size_t)x * ldj + y];
    return s;
}

/* h_y = b_y + sum_{x != y} J_xy s_x (jmat.hpp:153-162 recomputation). */
void orc_fields(const int8_t* J, int64_t ldj, int n, const int32_t* bias, const uint64_t* spins, int64_t* h) {
    for (int y = 0; y < n; ++y) h[y] = bias ? bias[y] : 0;
    for (int x = 0; x < n; ++x) {
        const int sx = ((spins[x >> 6] >> (x & 63)) & 1) ? 1 : -1;
        const int8_t* row = J + (size_t)x * ldj;
        for (int y = 0; y < n; ++y)
            if (y != x) h[y] += (int64_t)row[y] * sx;
    }
}
