// Declarations shared by the HIP kernels (anc_engine.hip) and the host side (anc_host.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace anc {

// Arguments of the persistent team kernel (one struct, passed by value).
struct TeamArgs {
    const int8_t* Jp;          // permuted J' (np x np, row-major), J'[y'][x'] with x' = l*kr + k for slot x = k*256 + l
    int np;
    int n;
    const int32_t* bias;       // original numbering, n entries, or null
    int S;
    const int32_t* fourT;
    const int32_t* q;
    const int64_t* kcorr;
    const int32_t* kconst;
    uint64_t seed;
    const uint32_t* trials;    // R trial ids, or null: trial = trial0 + r
    uint32_t trial0;
    int R;
    int n_teams;
    uint64_t* xfrag;           // exchange: D' fragments [team][parity][member][nt][64 lanes][16 B]
    uint32_t* xnlin;           // exchange: tagged n_lin partials [team][parity][member][RT]
    int* status;               // [0] error code (1 = exchange timeout), [1] abort flag
    long long timeout_ticks;   // wall-clock ticks before a wait gives up
    uint64_t* spins;
    int64_t spin_stride;
    int64_t* sum_sh;
    int64_t* sum_bs;
    int64_t* flips;
    int32_t* nlin;
    int32_t* fields;
    int64_t field_stride;
};

int launch_check_instance(const int8_t* J, int64_t ldj, int n, const int32_t* bias, unsigned long long* stats, hipStream_t st);
int launch_permute(const int8_t* J, int64_t ldj, int n, int np, int8_t* Jp, hipStream_t st);
int launch_fill_zero(uint64_t* p, int64_t words, hipStream_t st);
// Team kernel for np = 2048: nt = 1 or 2 sixteen-replica tiles per team, wide = int64 thresholds,
// xflavor 0 = XCD-local (L2) exchange, 1 = device-scope exchange. Returns 0, -1 (launch error) or -2 (not co-resident).
int launch_team(const TeamArgs& a, int nt, bool wide, int xflavor, int cus, hipStream_t st);

}  // namespace anc
