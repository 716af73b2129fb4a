// Problem encodings as Ising instances (host memory only; no GPU use): graph partitioning and the travelling
// salesman problem in Lucas's QUBO forms, with the reference generator's random streams, so that an instance is
// identical to the reference tool's for the same arguments (gpu/multibit_bias_20261008/src/gen_ising.cpp:145-283,
// `gen_ising gp` and `gen_ising tsp`). Spin s = 2x - 1; H(s) = -sum_{i<j} J_ij s_i s_j - sum_i b_i s_i.
//
// Graph partitioning (gen_ising.cpp:145-171): a random deg-regular simple graph on n vertices from the pairing model
// (n deg stubs, Fisher-Yates shuffle from the last stub with splitmix64 draws mod (i + 1), stream seeded with
// seed ^ 0x6A09E667F3BCC908; consecutive stubs pair up; a self-loop or a repeated edge rejects the whole pairing and
// a fresh stub list is shuffled from the same stream). f(x) = B cut(x) + A (sum x - k)^2;
// J_ij = -A + B [ij in E], b_i = A (2k - n); H = 2 f + C.
//
// TSP (gen_ising.cpp:205-240): cities 0..n-1, distances W_uv = 1 + draw % wmax (stream seed ^ 0x3C6EF372FE94F82B,
// pairs u < v row-major); city 0 fixed at position 0; spin (v-1)(n-1) + (p-1) <=> city v at position p (1..n-1).
// f = A sum_v (1 - sum_p x_vp)^2 + A sum_p (1 - sum_v x_vp)^2 + B length. As a QUBO f = c0 + sum_i c_i x_i +
// sum_{i<j} Q_ij x_i x_j with Q = 2A for two spins of one city or of one position, B W_uv for cities u != v at
// adjacent positions, c_i = -2A (+ B W_0v at position 1, + B W_v0 at position n-1), c0 = 2A (n-1);
// J_ij = -Q_ij, b_i = -(2 c_i + sum_j Q_ij); H = 4 f + C.
//
// C = H(all -1) - sigma f(all -1) is returned with every instance.
#include "anc_internal.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <unordered_set>
#include <vector>

namespace anc {
namespace {

uint64_t splitmix_next(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// H(all spins -1) = -sum_{i<j} J_ij + sum_i b_i
int64_t energy_all_down(int n, const int8_t* J, int64_t ldj, const int32_t* b) {
    int64_t e = 0;
    for (int i = 0; i < n; ++i) {
        e += b[i];
        for (int j = i + 1; j < n; ++j) e -= J[size_t(i) * ldj + j];
    }
    return e;
}

}  // namespace
}  // namespace anc

using namespace anc;

extern "C" {

int anc_gen_gp(uint64_t seed, int32_t n, int32_t deg, int32_t k, int64_t A, int64_t B, int8_t* J, int64_t ldj,
               int32_t* bias, int32_t* edges, int64_t* C) {
    clear_error();
    if (!J || !bias || !C) return fail(ANC_ERR_ARG, "null output");
    if (n < 4 || n > ANC_MAX_N || k < 1 || k >= n || deg < 1 || deg >= n || (int64_t(n) * deg) % 2 || ldj < n)
        return fail(ANC_ERR_ARG, "gp: need 4 <= n <= %d, 1 <= k < n, 1 <= deg < n, n deg even, ldj >= n", ANC_MAX_N);
    if (-A < -127 || -A > 127 || B - A < -127 || B - A > 127) return fail(ANC_ERR_RANGE, "gp: coupling outside int8");
    const int64_t bv = A * (2 * int64_t(k) - n);
    if (bv < INT32_MIN || bv > INT32_MAX) return fail(ANC_ERR_RANGE, "gp: bias outside int32");
    const int64_t nst = int64_t(n) * deg;
    std::vector<int32_t> stubs(static_cast<size_t>(nst));
    std::vector<uint64_t> E;   // edges u < v as u * n + v, in pairing order
    std::unordered_set<uint64_t> seen;
    uint64_t st = seed ^ 0x6A09E667F3BCC908ull;
    for (int attempts = 1;; ++attempts) {
        E.clear();
        seen.clear();
        for (int64_t i = 0; i < nst; ++i) stubs[size_t(i)] = int32_t(i / deg);
        for (int64_t i = nst - 1; i > 0; --i) std::swap(stubs[size_t(i)], stubs[size_t(splitmix_next(st) % uint64_t(i + 1))]);
        bool ok = true;
        for (int64_t i = 0; i < nst; i += 2) {
            int u = stubs[size_t(i)], v = stubs[size_t(i + 1)];
            if (u > v) std::swap(u, v);
            const uint64_t key = uint64_t(u) * uint64_t(n) + uint64_t(v);
            if (u == v || !seen.insert(key).second) { ok = false; break; }
            E.push_back(key);
        }
        if (ok) break;
        if (attempts >= 100000) return fail(ANC_ERR_ARG, "gp: no simple %d-regular graph on %d vertices found", deg, n);
    }
    for (int i = 0; i < n; ++i) {
        std::memset(J + size_t(i) * ldj, 0, size_t(ldj));
        for (int j = 0; j < n; ++j)
            if (j != i) J[size_t(i) * ldj + j] = int8_t(-A);
        bias[i] = int32_t(bv);
    }
    for (uint64_t key : E) {
        const int u = int(key / uint64_t(n)), v = int(key % uint64_t(n));
        J[size_t(u) * ldj + v] = J[size_t(v) * ldj + u] = int8_t(B - A);
    }
    if (edges) {
        std::sort(E.begin(), E.end());
        for (size_t e = 0; e < E.size(); ++e) {
            edges[2 * e] = int32_t(E[e] / uint64_t(n));
            edges[2 * e + 1] = int32_t(E[e] % uint64_t(n));
        }
    }
    *C = energy_all_down(n, J, ldj, bias) - 2 * A * int64_t(k) * k;   // f(all x = 0) = A k^2
    return ANC_OK;
}

int anc_gen_tsp(uint64_t seed, int32_t cities, int32_t wmax, int64_t A, int64_t B, int8_t* J, int64_t ldj, int32_t* bias,
                int32_t* W, int64_t* C) {
    clear_error();
    if (!J || !bias || !C) return fail(ANC_ERR_ARG, "null output");
    const int m = cities - 1;
    if (cities < 3 || int64_t(m) * m > ANC_MAX_N || wmax < 1) return fail(ANC_ERR_ARG, "tsp: need cities >= 3, (cities-1)^2 <= %d, wmax >= 1", ANC_MAX_N);
    const int nv = m * m;
    if (ldj < nv) return fail(ANC_ERR_ARG, "tsp: ldj < (cities-1)^2");
    std::vector<int64_t> Wd(size_t(cities) * cities, 0);
    uint64_t st = seed ^ 0x3C6EF372FE94F82Bull;
    for (int u = 0; u < cities; ++u)
        for (int v = u + 1; v < cities; ++v) Wd[size_t(u) * cities + v] = Wd[size_t(v) * cities + u] = 1 + int64_t(splitmix_next(st) % uint64_t(wmax));
    if (W)
        for (size_t i = 0; i < Wd.size(); ++i) W[i] = int32_t(Wd[i]);
    // Q for spins i = (v, p), j = (v2, p2), positions 1..m
    auto Q = [&](int v, int p, int v2, int p2) -> int64_t {
        if (v == v2 && p == p2) return 0;
        if (v == v2 || p == p2) return 2 * A;                                     // one city twice / one position twice
        if (p2 == p + 1 || p == p2 + 1) return B * Wd[size_t(v) * cities + v2];   // cities v != v2 at adjacent positions
        return 0;
    };
    for (int i = 0; i < nv; ++i) {
        const int v = i / m + 1, p = i % m + 1;
        int8_t* row = J + size_t(i) * ldj;
        std::memset(row, 0, size_t(ldj));
        int64_t rs = 0;
        for (int j = 0; j < nv; ++j) {
            if (j == i) continue;
            const int64_t q = Q(v, p, j / m + 1, j % m + 1);
            if (q > 127 || q < -127) return fail(ANC_ERR_RANGE, "tsp: coupling outside int8");
            row[j] = int8_t(-q);
            rs += q;
        }
        int64_t c = -2 * A;
        if (p == 1) c += B * Wd[size_t(v)];                      // W_0v
        if (p == m) c += B * Wd[size_t(v) * cities];             // W_v0
        const int64_t bi = -(2 * c + rs);
        if (bi < -32767 || bi > 32767) return fail(ANC_ERR_RANGE, "tsp: bias outside 16 bits");
        bias[i] = int32_t(bi);
    }
    *C = energy_all_down(nv, J, ldj, bias) - 4 * (2 * A * int64_t(m));   // f(all x = 0) = c0 = 2 A (cities - 1)
    return ANC_OK;
}

}  // extern "C"
