// Instance files, instance statistics and independent scorers (host memory only; no GPU use).
//
// Formats (the reference's loaders and writers: gpu/multibit_bias_20261008/src/jmat.hpp:86-113 SCAJINT8,
// :164-186 SCABIAS1, :65-84 G-set; fpga/v80_sca/src/sca_ref.hpp:31-49 SNOWGPU1, written by data/fetch_data.py:38-54):
//   SCAJINT8  "SCAJINT8", uint32 n, uint32 0, then n*n int8 row-major. Nothing may follow.
//   SNOWGPU1  "SNOWGPU1", uint32 n, then n rows of ceil(n/64) little-endian uint64 words; bit y of row x is 1 iff
//             J_xy = +1, else J_xy = -1 (x != y, y < n). Diagonal and padding bits are ignored.
//   G-set     text "n m", then m lines "i j w" (1-indexed, i != j, |w| <= 127, no repeated pair); J_ij = J_ji = -w.
//             A further integer after the m edges is an error.
//   SCABIAS1  "SCABIAS1", uint32 n, uint32 0, then n int32 with |b| <= 32767. Nothing may follow.
// Every loaded matrix is validated as jmat::finish does (jmat.hpp:35-55): zero diagonal and symmetric, and its
// statistics (sumw = sum_{x<y} -J_xy, max |J|, field bound max_x sum_y |J_xy|, mean row sum of squares) are those
// of finish. The reference's loaders take n <= 2048 (its compiled NP; SNOWGPU1 exactly n = 2000); these take any
// n <= ANC_MAX_N.
//
// Scorers (jmat.hpp:130-162 scalar_cut, energy, field_mismatches), per replica r with spins s (bit i of the packed
// words = 1 iff s_i = +1): cut = sum_{x<y, s_x != s_y} -J_xy; energy = -sum_{x<y} J_xy s_x s_y - sum_x b_x s_x;
// mismatches = #{x : b_x + sum_y J_xy s_y != fields_x}. They are recomputed from J, never from engine outputs.
#include "anc_internal.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

namespace anc {
namespace {

struct File {
    FILE* f = nullptr;
    explicit File(FILE* p) : f(p) {}
    ~File() {
        if (f) std::fclose(f);
    }
    File(const File&) = delete;
    File& operator=(const File&) = delete;
};

bool at_eof(FILE* f) {
    char c;
    return std::fread(&c, 1, 1, f) != 1;
}

int check_matrix(int n, const int8_t* J, int64_t ldj) {
    for (int x = 0; x < n; ++x) {
        if (J[size_t(x) * ldj + x] != 0) return fail(ANC_ERR_INSTANCE, "nonzero diagonal coupling at %d", x);
        for (int y = x + 1; y < n; ++y)
            if (J[size_t(x) * ldj + y] != J[size_t(y) * ldj + x])
                return fail(ANC_ERR_INSTANCE, "J not symmetric at (%d, %d)", x, y);
    }
    return ANC_OK;
}

int check_dims(int n, const void* J, int64_t ldj) {
    if (n < 1 || n > ANC_MAX_N) return fail(ANC_ERR_ARG, "n = %d outside 1..%d", n, ANC_MAX_N);
    if (!J) return fail(ANC_ERR_ARG, "null matrix");
    if (ldj < n) return fail(ANC_ERR_ARG, "row stride %lld < n = %d", (long long)ldj, n);
    return ANC_OK;
}

// reads the SCAJINT8 / SCABIAS1 header: magic, n, reserved 0
int read_header16(FILE* f, const char* magic, uint32_t& n, const char* what) {
    char m[8];
    uint32_t rsv = 1;
    if (std::fread(m, 1, 8, f) != 8 || std::memcmp(m, magic, 8) != 0 || std::fread(&n, 4, 1, f) != 1 ||
        std::fread(&rsv, 4, 1, f) != 1 || rsv != 0)
        return fail(ANC_ERR_IO, "bad %s header", what);
    return ANC_OK;
}

int read_jint8(FILE* f, int32_t* n_out, int8_t* J, int64_t ldj) {
    uint32_t n = 0;
    if (int rc = read_header16(f, "SCAJINT8", n, "JINT8")) return rc;
    if (n < 1 || n > uint32_t(ANC_MAX_N)) return fail(ANC_ERR_IO, "bad JINT8 header (n = %u outside 1..%d)", n, ANC_MAX_N);
    *n_out = int32_t(n);
    if (!J) return ANC_OK;
    if (int rc = check_dims(int(n), J, ldj)) return rc;
    for (uint32_t x = 0; x < n; ++x) {
        int8_t* row = J + size_t(x) * ldj;
        if (std::fread(row, 1, n, f) != n) return fail(ANC_ERR_IO, "short JINT8 file");
        std::memset(row + n, 0, size_t(ldj - n));
    }
    if (!at_eof(f)) return fail(ANC_ERR_IO, "trailing data in JINT8 file");
    return check_matrix(int(n), J, ldj);
}

int read_snowgpu1(FILE* f, int32_t* n_out, int8_t* J, int64_t ldj) {
    char m[8];
    uint32_t n = 0;
    if (std::fread(m, 1, 8, f) != 8 || std::memcmp(m, "SNOWGPU1", 8) != 0 || std::fread(&n, 4, 1, f) != 1 || n < 1 ||
        n > uint32_t(ANC_MAX_N))
        return fail(ANC_ERR_IO, "bad graph header");
    *n_out = int32_t(n);
    if (!J) return ANC_OK;
    if (int rc = check_dims(int(n), J, ldj)) return rc;
    const size_t words = (n + 63) / 64;
    std::vector<uint64_t> row(words);
    for (uint32_t x = 0; x < n; ++x) {
        if (std::fread(row.data(), 8, words, f) != words) return fail(ANC_ERR_IO, "short graph file");
        int8_t* out = J + size_t(x) * ldj;
        for (uint32_t y = 0; y < n; ++y) out[y] = y == x ? 0 : (((row[y >> 6] >> (y & 63)) & 1) ? 1 : -1);
        std::memset(out + n, 0, size_t(ldj - n));
    }
    return check_matrix(int(n), J, ldj);
}

// G-set text. A first pass (J == nullptr) reads the header only.
int read_gset(FILE* f, int32_t* n_out, int8_t* J, int64_t ldj) {
    long n = 0, e = 0;
    if (std::fscanf(f, "%ld %ld", &n, &e) != 2) return fail(ANC_ERR_IO, "bad G-set header");
    if (n < 1 || n > ANC_MAX_N) return fail(ANC_ERR_IO, "G-set N = %ld out of range 1..%d", n, ANC_MAX_N);
    *n_out = int32_t(n);
    if (!J) return ANC_OK;
    if (int rc = check_dims(int(n), J, ldj)) return rc;
    for (long x = 0; x < n; ++x) std::memset(J + size_t(x) * ldj, 0, size_t(ldj));
    for (long k = 0; k < e; ++k) {
        long i, j, w;
        if (std::fscanf(f, "%ld %ld %ld", &i, &j, &w) != 3) return fail(ANC_ERR_IO, "short G-set file");
        if (i < 1 || j < 1 || i > n || j > n || i == j) return fail(ANC_ERR_IO, "bad G-set edge (%ld, %ld)", i, j);
        if (w < -127 || w > 127) return fail(ANC_ERR_IO, "G-set weight %ld out of int8 range", w);
        int8_t& a = J[size_t(i - 1) * ldj + size_t(j - 1)];
        int8_t& b = J[size_t(j - 1) * ldj + size_t(i - 1)];
        if (a != 0 || b != 0) return fail(ANC_ERR_IO, "duplicate G-set edge (%ld, %ld)", i, j);
        a = b = int8_t(-w);
    }
    long extra;
    if (std::fscanf(f, "%ld", &extra) == 1) return fail(ANC_ERR_IO, "trailing data in G-set file");
    return ANC_OK;
}

// s as +-1 bytes from packed words
void unpack_spins(int n, const uint64_t* w, int8_t* s) {
    for (int i = 0; i < n; ++i) s[i] = ((w[i >> 6] >> (i & 63)) & 1) ? 1 : -1;
}

int32_t dot8(const int8_t* a, const int8_t* b, int len) {
    int32_t acc = 0;
    for (int i = 0; i < len; ++i) acc += int32_t(a[i]) * int32_t(b[i]);
    return acc;
}

int32_t sum8(const int8_t* a, int len) {
    int32_t acc = 0;
    for (int i = 0; i < len; ++i) acc += a[i];
    return acc;
}

}  // namespace
}  // namespace anc

using namespace anc;

extern "C" {

int anc_read_instance(const char* path, int format, int32_t* n_out, int8_t* J, int64_t ldj) {
    clear_error();
    if (!path || !n_out) return fail(ANC_ERR_ARG, "null path or n_out");
    File f(std::fopen(path, format == ANC_FMT_GSET ? "r" : "rb"));
    if (!f.f) return fail(ANC_ERR_IO, "cannot open %s", path);
    switch (format) {
        case ANC_FMT_JINT8: return read_jint8(f.f, n_out, J, ldj);
        case ANC_FMT_SNOWGPU1: return read_snowgpu1(f.f, n_out, J, ldj);
        case ANC_FMT_GSET: return read_gset(f.f, n_out, J, ldj);
        default: return fail(ANC_ERR_ARG, "unknown instance format %d", format);
    }
}

int anc_write_instance(const char* path, int format, int32_t n, const int8_t* J, int64_t ldj) {
    clear_error();
    if (!path) return fail(ANC_ERR_ARG, "null path");
    if (int rc = check_dims(n, J, ldj)) return rc;
    if (int rc = check_matrix(n, J, ldj)) return rc;
    if (format == ANC_FMT_SNOWGPU1)
        for (int x = 0; x < n; ++x)
            for (int y = 0; y < n; ++y)
                if (x != y && J[size_t(x) * ldj + y] != 1 && J[size_t(x) * ldj + y] != -1)
                    return fail(ANC_ERR_ARG, "SNOWGPU1 holds +-1 couplings only (J[%d][%d] = %d)", x, y, J[size_t(x) * ldj + y]);
    if (format == ANC_FMT_GSET)
        for (int x = 0; x < n; ++x)
            for (int y = 0; y < n; ++y)
                if (J[size_t(x) * ldj + y] == -128) return fail(ANC_ERR_ARG, "G-set weights are limited to |w| <= 127");
    if (format != ANC_FMT_JINT8 && format != ANC_FMT_SNOWGPU1 && format != ANC_FMT_GSET)
        return fail(ANC_ERR_ARG, "unknown instance format %d", format);
    File f(std::fopen(path, format == ANC_FMT_GSET ? "w" : "wb"));
    if (!f.f) return fail(ANC_ERR_IO, "cannot create %s", path);
    bool ok = true;
    const uint32_t un = uint32_t(n), zero = 0;
    if (format == ANC_FMT_JINT8) {
        ok = std::fwrite("SCAJINT8", 1, 8, f.f) == 8 && std::fwrite(&un, 4, 1, f.f) == 1 && std::fwrite(&zero, 4, 1, f.f) == 1;
        for (int x = 0; ok && x < n; ++x) ok = std::fwrite(J + size_t(x) * ldj, 1, size_t(n), f.f) == size_t(n);
    } else if (format == ANC_FMT_SNOWGPU1) {
        ok = std::fwrite("SNOWGPU1", 1, 8, f.f) == 8 && std::fwrite(&un, 4, 1, f.f) == 1;
        const size_t words = (size_t(n) + 63) / 64;
        std::vector<uint64_t> row(words);
        for (int x = 0; ok && x < n; ++x) {
            std::fill(row.begin(), row.end(), 0);
            for (int y = 0; y < n; ++y)
                if (y != x && J[size_t(x) * ldj + y] == 1) row[size_t(y) >> 6] |= uint64_t(1) << (y & 63);
            ok = std::fwrite(row.data(), 8, words, f.f) == words;
        }
    } else {
        long m = 0;
        for (int x = 0; x < n; ++x)
            for (int y = x + 1; y < n; ++y) m += J[size_t(x) * ldj + y] != 0;
        ok = std::fprintf(f.f, "%d %ld\n", n, m) > 0;
        for (int x = 0; ok && x < n; ++x)
            for (int y = x + 1; ok && y < n; ++y)
                if (const int v = J[size_t(x) * ldj + y]) ok = std::fprintf(f.f, "%d %d %d\n", x + 1, y + 1, -v) > 0;
    }
    if (!ok) return fail(ANC_ERR_IO, "write to %s failed", path);
    return ANC_OK;
}

int anc_read_bias(const char* path, int32_t n, int32_t* bias) {
    clear_error();
    if (!path || !bias) return fail(ANC_ERR_ARG, "null path or bias");
    File f(std::fopen(path, "rb"));
    if (!f.f) return fail(ANC_ERR_IO, "cannot open %s", path);
    uint32_t fn = 0;
    if (int rc = read_header16(f.f, "SCABIAS1", fn, "bias")) return rc;
    if (int64_t(fn) != int64_t(n)) return fail(ANC_ERR_IO, "bias length %u differs from N = %d", fn, n);
    if (std::fread(bias, 4, fn, f.f) != fn) return fail(ANC_ERR_IO, "short bias file");
    if (!at_eof(f.f)) return fail(ANC_ERR_IO, "trailing data in bias file");
    for (uint32_t i = 0; i < fn; ++i)
        if (bias[i] < -32767 || bias[i] > 32767) return fail(ANC_ERR_IO, "bias outside |b| <= 32767 (b[%u] = %d)", i, bias[i]);
    return ANC_OK;
}

int anc_write_bias(const char* path, int32_t n, const int32_t* bias) {
    clear_error();
    if (!path || !bias || n < 1) return fail(ANC_ERR_ARG, "null path or bias, or n < 1");
    for (int i = 0; i < n; ++i)
        if (bias[i] < -32767 || bias[i] > 32767) return fail(ANC_ERR_ARG, "bias outside |b| <= 32767 (b[%d] = %d)", i, bias[i]);
    File f(std::fopen(path, "wb"));
    if (!f.f) return fail(ANC_ERR_IO, "cannot create %s", path);
    const uint32_t un = uint32_t(n), zero = 0;
    if (std::fwrite("SCABIAS1", 1, 8, f.f) != 8 || std::fwrite(&un, 4, 1, f.f) != 1 || std::fwrite(&zero, 4, 1, f.f) != 1 ||
        std::fwrite(bias, 4, size_t(n), f.f) != size_t(n))
        return fail(ANC_ERR_IO, "write to %s failed", path);
    return ANC_OK;
}

int anc_instance_stats(int32_t n, const int8_t* J, int64_t ldj, anc_jstats* st) {
    clear_error();
    if (!st) return fail(ANC_ERR_ARG, "null stats");
    if (int rc = check_dims(n, J, ldj)) return rc;
    if (int rc = check_matrix(n, J, ldj)) return rc;
    int64_t sumw = 0, hmax = 0, nnz = 0;
    int maxabs = 0;
    double sq = 0;
    for (int x = 0; x < n; ++x) {
        const int8_t* row = J + size_t(x) * ldj;
        int64_t rabs = 0, rsq = 0;
        for (int y = 0; y < n; ++y) {
            const int v = row[y];
            rabs += std::abs(v);
            rsq += v * v;
            nnz += v != 0;
            maxabs = std::max(maxabs, std::abs(v));
            if (y > x) sumw -= v;
        }
        hmax = std::max(hmax, rabs);
        sq += double(rsq);
    }
    st->sumw = sumw;
    st->maxabs = maxabs;
    st->hmax = hmax;
    st->mean_row_sq = sq / n;
    st->nnz = nnz;
    return ANC_OK;
}

int anc_score(int32_t n, const int8_t* J, int64_t ldj, const int32_t* bias, int32_t R, const uint64_t* spins,
              int64_t spin_stride, const int32_t* fields, int64_t field_stride, int64_t* cut, int64_t* energy,
              int64_t* mismatches, int nthreads) {
    clear_error();
    if (int rc = check_dims(n, J, ldj)) return rc;
    if (R < 0 || !spins || spin_stride < (n + 63) / 64) return fail(ANC_ERR_ARG, "bad replica count, spins or spin stride");
    if (mismatches && (!fields || field_stride < n)) return fail(ANC_ERR_ARG, "mismatch count needs fields with stride >= n");
    if (int rc = check_matrix(n, J, ldj)) return rc;
    // per-row sums over y > x (the couplings of each pair once)
    std::vector<int32_t> upper(static_cast<size_t>(n));
    for (int x = 0; x < n; ++x) upper[size_t(x)] = sum8(J + size_t(x) * ldj + x + 1, n - x - 1);
    unsigned nt = nthreads > 0 ? unsigned(nthreads) : std::max(1u, std::thread::hardware_concurrency());
    nt = std::min<unsigned>(nt, unsigned(std::max(1, R)));
    auto work = [&](int r0, int r1) {
        std::vector<int8_t> s(static_cast<size_t>(n));
        for (int r = r0; r < r1; ++r) {
            unpack_spins(n, spins + size_t(r) * spin_stride, s.data());
            int64_t c = 0, pair = 0, bs = 0, bad = 0;
            for (int x = 0; x < n; ++x) {
                const int8_t* row = J + size_t(x) * ldj;
                const int32_t lo = dot8(row, s.data(), x), hi = dot8(row + x + 1, s.data() + x + 1, n - x - 1);
                // sum_{y>x, s_y != s_x} J_xy = (upper_x - s_x hi) / 2
                c -= (upper[size_t(x)] - s[size_t(x)] * hi) / 2;
                pair += s[size_t(x)] * hi;
                const int32_t b = bias ? bias[x] : 0;
                bs += int64_t(b) * s[size_t(x)];
                if (mismatches) bad += int64_t(b) + lo + hi != int64_t(fields[size_t(r) * field_stride + x]);
            }
            if (cut) cut[r] = c;
            if (energy) energy[r] = -pair - bs;
            if (mismatches) mismatches[r] = bad;
        }
    };
    if (nt <= 1) {
        work(0, R);
    } else {
        std::vector<std::thread> th;
        for (unsigned i = 0; i < nt; ++i) th.emplace_back(work, int(int64_t(R) * i / nt), int(int64_t(R) * (i + 1) / nt));
        for (auto& t : th) t.join();
    }
    return ANC_OK;
}

}  // extern "C"
