// Large instances (NP(n) > 3072): instance preparation (fragment-major J', permuted row sums and biases) and the
// launch dispatch over the kernel slot counts 4096, 8192 and 16384 (the kernel itself is in anc_large.hpp).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "anc_large.hpp"

namespace anc {
namespace dev {

// permuted index x' -> slot x (see the layout above); kr_k = KR of the kernel (npk / 256)
__device__ __forceinline__ int large_slot(int xp, int kr_k) {
    const int f = 64 / kr_k;
    const int c = xp >> 8, rho = xp & 255;
    const int mt = rho >> 4, g = (rho >> 2) & 3, e = rho & 3;
    const int j = 4 * mt + e;
    const int lane = 4 * f * c + f * g + j / kr_k;
    return 256 * (j % kr_k) + lane;
}

// Fragment-major J': block (c, kk) writes the 16 KiB of K-block kk of row tile c. A-operand lane (g, ml) of M-tile
// mt holds row 16 mt + ml and the 16 columns of K-pair kp = kk / 2 that match the unpacked spin nibbles: the exchange
// chunk of lane group g holds, as nibble s, the slot 128 kp + 16 (s / 4) + 4 g + s % 4 (the rows of producer lane
// group g), and unpacking puts nibble 2b (K-block 2 kp) or 2b + 1 (K-block 2 kp + 1) into byte b, so byte b of this
// fragment is column 128 kp + 16 (b / 2) + 4 g + 2 (b % 2) + (kk % 2).
__global__ void permute_large(const int8_t* __restrict__ J, int64_t ldj, int n, int npk, int8_t* __restrict__ Jf) {
    const int kr_k = npk / 256, nk = npk / 64;
    const int c = blockIdx.x / nk, kk = blockIdx.x % nk;
    int8_t* out = Jf + (size_t(c) * nk + kk) * 16384;
    for (int i = threadIdx.x; i < 16384; i += blockDim.x) {
        const int mt = i >> 10, lane = (i >> 4) & 63, b = i & 15;
        const int yp = 256 * c + 16 * mt + (lane & 15);
        const int xp = 128 * (kk >> 1) + 16 * (b >> 1) + 4 * (lane >> 4) + 2 * (b & 1) + (kk & 1);
        const int y = large_slot(yp, kr_k), x = large_slot(xp, kr_k);
        out[i] = (y < n && x < n) ? J[size_t(y) * size_t(ldj) + x] : int8_t(0);
    }
}

// Row sums and biases in permuted order (one block per permuted row).
__global__ void rowsum_large(const int8_t* __restrict__ J, int64_t ldj, int n, const int32_t* __restrict__ bias, int npk,
                             int32_t* __restrict__ rs, int32_t* __restrict__ bp) {
    const int xp = blockIdx.x;
    const int y = large_slot(xp, npk / 256);
    int sum = 0;
    if (y < n)
        for (int x = threadIdx.x; x < n; x += blockDim.x) sum += J[size_t(y) * size_t(ldj) + x];
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
    __shared__ int red[16];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int i = 0; i < int(blockDim.x >> 6); ++i) t += red[i];
        rs[xp] = t;
        bp[xp] = (y < n && bias) ? bias[y] : 0;
    }
}

}  // namespace dev

int large_np_for(int np) { return np <= 3072 ? 0 : np <= 4096 ? 4096 : np <= 8192 ? 8192 : np <= 16384 ? 16384 : 0; }

size_t large_xd_bytes(int npk) {
    const int kr = npk / 256;
    const size_t teams = 256 / kr, nk = npk / 64;
    // spins of every team [team][parity][replica half][K-pair][n-tile][lane][16 B], then n_lin counters (4 KiB a team)
    return teams * (2 * 2 * (nk / 2) * 4096 + 4096);
}

int launch_permute_large(const int8_t* J, int64_t ldj, int n, const int32_t* bias, int npk, int8_t* Jf, int32_t* rs,
                         int32_t* bp, hipStream_t st) {
    const int nk = npk / 64, nrt = npk / 256;
    hipLaunchKernelGGL(dev::permute_large, dim3(nrt * nk), dim3(256), 0, st, J, ldj, n, npk, Jf);
    hipLaunchKernelGGL(dev::rowsum_large, dim3(npk), dim3(256), 0, st, J, ldj, n, bias, npk, rs, bp);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_large(const LargeArgs& a, int npk, bool wide, int cus, bool dry, hipStream_t st) {
    const bool c = a.count_nlin != 0;
    switch (npk) {
        case 4096: return c ? launch_large_c16(a, wide, cus, dry, st) : launch_large_k16(a, wide, cus, dry, st);
        case 8192: return c ? launch_large_c32(a, wide, cus, dry, st) : launch_large_k32(a, wide, cus, dry, st);
        case 16384: return c ? launch_large_c64(a, wide, cus, dry, st) : launch_large_k64(a, wide, cus, dry, st);
        default: return -3;
    }
}

}  // namespace anc
