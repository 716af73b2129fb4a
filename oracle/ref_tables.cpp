// ORACLE -- test infrastructure only. Wraps v80host::build_tables (gpu/onsager_v2_20261007/src/tables.hpp:24-40), the
// GPU host's flag-built tables. A separate translation unit because tables.hpp includes its own copy of sca_ref.hpp.
#include "tables.hpp"

extern "C" int ref_build_tables(int S, double t0, double t1, double q, double lam, double jv, double kappa, int ramp,
                                int32_t* fourT, int32_t* qo, int64_t* kcorr, int32_t* kconst) {
    v80host::Schedule c; c.S = S; c.t0 = t0; c.t1 = t1; c.q = q; c.lam = lam; c.jv = jv; c.kappa = kappa; c.ramp = ramp != 0;
    const sca::Tables tb = v80host::build_tables(c);
    for (int t = 0; t < S; ++t) { fourT[t] = tb.fourT[t]; qo[t] = tb.q[t]; kcorr[t] = tb.kcorr[t]; kconst[t] = tb.kconst[t]; }
    return 0;
}
