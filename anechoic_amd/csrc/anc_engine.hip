// Device side of the MI355X SCA/Onsager engine: instance preparation kernels and the persistent team kernel.
//
// Arithmetic contract (reference fpga/v80_sca/src/sca_ref_bias.hpp:25-75, generalised NP):
//   per anneal, 256 xoshiro128** lanes seeded from (seed, trial, lane) (sca_ref.hpp:71-91); slot i = k*256 + l is drawn
//   by lane l in round k; padding draws are consumed; s_i(0) = +1 iff the draw's top bit is set;
//   h(0) = b + J s(0); each step t: corr = ((n_lin(t-1) kcorr[t]) >> 8) + kconst[t];
//   z_i = s_i h_i 2^16 + q[t] - (s_i == s'_i ? corr : -corr); r_i = ((u16 - 32768) fourT[t]) >> 16;
//   n_lin counts -2T < z < 2T (twoT = fourT >> 1); slot i flips iff z_i < r_i; then h += J (s_new - s_old).
//
// Engine layout (MI355X, 8 XCDs x 32 CUs):
//   * A team is 32 workgroups on one XCD (workgroup b runs on XCD b % 8). Team member c owns RNG lanes 8c..8c+7, i.e.
//     the 64 slots x = k*256 + l of those lanes, renumbered x' = l*8 + k so that every member owns one contiguous
//     64-slot block [64c, 64c+64) of the permuted index.
//   * Each member keeps its 64 rows of the permuted coupling matrix J'[y'][x'] resident in VGPRs as int8 MFMA
//     A-fragments (wave w holds the K-quarter x' in [512w, 512w+512)), and its part of the field H[64 x RT] as four
//     per-wave int32 MFMA accumulators (H = bias + sum_w P_w).
//   * One step = decide (256 threads = RT replicas x 8 lanes; thread owns its lane's RNG stream and 8 slots) ->
//     publish the member's flip matrix slice D' = D + tau (int8, D in {-2,0,2}, tau = step parity bit in bit 0 of
//     every byte) directly in MFMA B-fragment layout plus a tagged n_lin partial per replica -> every member reads the
//     D' fragments of its team (self-validating: a byte is current iff its bit 0 equals tau) ->
//     v_mfma_i32_16x16x64_i8 accumulation P_w += J'_w D'_w -> the tau contribution tau*rowsum(J) is removed exactly
//     in the decision threads.
//   * No fence or barrier crosses workgroups: every exchanged word carries its own validity, so relaxed loads and
//     stores suffice (the exchange never relies on ordering between different words).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "anc_team.hpp"

namespace anc {
namespace dev {

// ---------------------------------------------------------------------------------------------------------- prep kernels
// Instance check (jmat.hpp:35-55 rules): J symmetric, zero diagonal; field bound max_y sum_x |J_xy| and max |b|.
// stats: [0] asymmetric entries, [1] nonzero diagonal entries, [2] max row abs sum, [3] max |b|.
__global__ void check_instance(const int8_t* __restrict__ J, int64_t ldj, int n, const int32_t* __restrict__ bias,
                               unsigned long long* stats) {
    const int y = blockIdx.x;
    if (y >= n) return;
    const int8_t* row = J + size_t(y) * size_t(ldj);
    unsigned long long asym = 0, diag = 0, rabs = 0;
    for (int x = threadIdx.x; x < n; x += blockDim.x) {
        const int v = row[x];
        const unsigned m = unsigned(v < 0 ? -v : v);
        rabs += m;
        if (x == y) diag += v != 0;
        else if (v != J[size_t(x) * size_t(ldj) + y]) asym += 1;
    }
    for (int o = 32; o > 0; o >>= 1) {
        asym += __shfl_xor(asym, o); diag += __shfl_xor(diag, o); rabs += __shfl_xor(rabs, o);
    }
    __shared__ unsigned long long red[3][16];
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) { red[0][w] = asym; red[1][w] = diag; red[2][w] = rabs; }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long a = 0, d = 0, r = 0;
        for (int i = 0; i < int(blockDim.x / 64); ++i) { a += red[0][i]; d += red[1][i]; r += red[2][i]; }
        if (a) atomicAdd(stats + 0, a);
        if (d) atomicAdd(stats + 1, d);
        atomicMax(stats + 2, r);
        if (bias) {
            const int b = bias[y];
            atomicMax(stats + 3, (unsigned long long)(b < 0 ? -(long long)b : (long long)b));
        }
    }
}

// J'[y'][x'] = J[y][x] with x' = l*krk + k for x = k*256 + l (zero outside n). krk = the kernel's compile-time round
// count; rounds k >= NP(n)/256 are phantom slots (x >= n), so they are zero as well. One block per permuted row.
__global__ void permute_instance(const int8_t* __restrict__ J, int64_t ldj, int n, int npk, int krk, int8_t* __restrict__ Jp) {
    const int yp = blockIdx.x;
    const int y = (yp % krk) * LANES + yp / krk;
    int8_t* out = Jp + size_t(yp) * size_t(npk);
    for (int xp = threadIdx.x; xp < npk; xp += blockDim.x) {
        const int x = (xp % krk) * LANES + xp / krk;
        out[xp] = (x < n && y < n) ? J[size_t(y) * size_t(ldj) + x] : int8_t(0);
    }
}

// gen_ising random (gpu/multibit_bias_20261008/src/gen_ising.cpp:118-135) on the device: pair p = (x < y) in
// row-major order uses draws 2p and 2p + 1 of the splitmix64 stream seeded with seed ^ 0xA5A5F00D12345678 (the
// stream is counter-based: draw k is mix(state + (k + 1) gamma)); |J| = 1 + (r1 >> 1) % M, negative iff r1 & 1,
// zero iff r0 < zthr. Row x of the grid writes J[x][y] and J[y][x] for y > x, zeros on the diagonal and in the
// padding columns n .. ldj - 1.
__device__ __forceinline__ uint64_t splitmix_at(uint64_t state, uint64_t k) {
    uint64_t x = state + k * 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__global__ void gen_ising_pairs(uint64_t seed, int n, int M, uint64_t zthr, int8_t* __restrict__ J, int64_t ldj) {
    const int x = blockIdx.y;
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= ldj) return;
    int8_t* row = J + size_t(x) * size_t(ldj);
    if (y <= x || y >= n) {
        if (y == x || y >= n) row[y] = 0;
        return;
    }
    const uint64_t st = seed ^ 0xA5A5F00D12345678ull;
    const int64_t p = int64_t(x) * (n - 1) - int64_t(x) * (x - 1) / 2 + (y - x - 1);
    const uint64_t r0 = splitmix_at(st, uint64_t(2 * p + 1)), r1 = splitmix_at(st, uint64_t(2 * p + 2));
    int v = 0;
    if (!(zthr && r0 < zthr)) {
        const int mag = 1 + int((r1 >> 1) % uint64_t(M));
        v = (r1 & 1) ? -mag : mag;
    }
    row[y] = int8_t(v);
    J[size_t(y) * size_t(ldj) + x] = int8_t(v);
}

// Biases of gen_ising random (pbig >= 1): b_i = draw_i % (2 bmax + 1) - bmax from seed ^ 0x5B1A5B1A5B1A5B1A.
__global__ void gen_ising_bias(uint64_t seed, int n, int bmax, int32_t* __restrict__ b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    b[i] = bmax > 0 ? int32_t(int64_t(splitmix_at(seed ^ 0x5B1A5B1A5B1A5B1Aull, uint64_t(i + 1)) % uint64_t(2 * bmax + 1)) - bmax)
                    : 0;
}

// Zeroes n 64-bit words (exchange workspace, status and the accumulated outputs) before a team launch.
__global__ void fill_zero(uint64_t* __restrict__ p, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) p[i] = 0;
}

// Several ranges in one launch: blockIdx.y picks the range.
__global__ void fill_zero_ranges(FillRanges r) {
    const int k = blockIdx.y;
    uint64_t* p = r.p[k];
    const int64_t n = r.n[k];
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) p[i] = 0;
}

}  // namespace dev

// ---------------------------------------------------------------------------------------------------------- launchers
int launch_check_instance(const int8_t* J, int64_t ldj, int n, const int32_t* bias, unsigned long long* stats, hipStream_t st) {
    hipLaunchKernelGGL(dev::check_instance, dim3(n), dim3(256), 0, st, J, ldj, n, bias, stats);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_permute(const int8_t* J, int64_t ldj, int n, int npk, int8_t* Jp, hipStream_t st) {
    hipLaunchKernelGGL(dev::permute_instance, dim3(npk), dim3(256), 0, st, J, ldj, n, npk, npk / 256, Jp);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_gen_ising(uint64_t seed, int n, int M, uint64_t zthr, int bmax, int8_t* J, int64_t ldj, int32_t* bias, hipStream_t st) {
    hipLaunchKernelGGL(dev::gen_ising_pairs, dim3(unsigned((ldj + 255) / 256), unsigned(n)), dim3(256), 0, st, seed, n, M, zthr, J, ldj);
    if (bias) hipLaunchKernelGGL(dev::gen_ising_bias, dim3((n + 255) / 256), dim3(256), 0, st, seed, n, bmax, bias);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_fill_zero(uint64_t* p, int64_t words, hipStream_t st) {
    if (words <= 0) return 0;
    int blocks = int((words + 255) / 256);
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(dev::fill_zero, dim3(blocks), dim3(256), 0, st, p, words);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_fill_zero_ranges(const FillRanges& r, hipStream_t st) {
    if (r.count <= 0) return 0;
    int64_t mx = 0;
    for (int k = 0; k < r.count; ++k) mx = r.n[k] > mx ? r.n[k] : mx;
    if (mx <= 0) return 0;
    int blocks = int((mx + 255) / 256);
    if (blocks > 512) blocks = 512;
    hipLaunchKernelGGL(dev::fill_zero_ranges, dim3(blocks, r.count), dim3(256), 0, st, r);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int kernel_np_for(int np) { return np == 2048 ? 2048 : (np > 2048 && np <= 3072) ? 3072 : 0; }
int team_size_for(int np, int xmode) { const int k = kernel_np_for(np); return k == 2048 ? 32 : k == 3072 ? (xmode ? 32 : 64) : 0; }

int launch_team(const TeamArgs& a, int nt, bool wide, int xflavor, int cus, bool dry, hipStream_t st) {
    if (a.np == 2048) return launch_team32(a, nt, wide, xflavor, cus, dry, st);
    if (a.np == 3072) return launch_team3072(a, nt, wide, xflavor, cus, dry, st);
    return -3;
}

}  // namespace anc
