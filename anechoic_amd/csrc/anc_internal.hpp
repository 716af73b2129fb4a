// Internal declarations shared by the host-side sources of libanechoic_mi355x.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/anechoic.h"

namespace anc {

// Records a formatted message as the calling thread's last error and returns code.
int fail(int code, const char* fmt, ...);
void clear_error();

inline int padded_slots(int n) {
    int np = ((n + 255) / 256) * 256;
    return np < 2048 ? 2048 : np;
}

struct Tables {
    std::vector<int32_t> fourT, q, kconst;
    std::vector<int64_t> kcorr;
};

// Exact-range plan of the decision arithmetic (see plan() in anc_tables.cpp).
struct Plan {
    bool ok = false;
    bool wide = false;   // int64 thresholds needed
    bool corr24 = false; // int32 plan and every step's |kcorr >> 8| < 2^23 with nmax |kcorr >> 8| < 2^31: the counted
                         // correction n_lin kcorr >> 8 can be formed with 24-bit multiplies
};

int make_tables(const anc_sched& c, int32_t* fourT, int32_t* q, int64_t* kcorr, int32_t* kconst);
int plan(const Tables& tb, int nmax, int64_t hbound, Plan& p);

}  // namespace anc
