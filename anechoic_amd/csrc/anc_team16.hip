// Instantiations of the 16-member team kernel (anc_team16.hpp): int32 thresholds (with and without the 24-bit
// correction split) / int64 thresholds, exchange flavours 1 and 3, with and without the phase counters.
#include "anc_team.hpp"
#include "anc_team16.hpp"

namespace anc {

template <bool WIDE, int XF, bool C24, bool PROF>
static int launch16_t(const TeamArgs& a, int cus, bool dry, hipStream_t st) {
    const auto kern = dev::team16_kernel<WIDE, XF, C24, PROF>;
    const int grid = ((a.n_teams + 7) / 8) * 8 * 16;
    size_t pad = 0;
    if (int rc = cached_pad((const void*)kern, pad)) return rc;
    if (a.n_teams > 16 || grid > cus) return -2;
    if (dry) return 0;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(dev::THREADS), pad, st, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int XF, bool PROF>
static int launch16_x(const TeamArgs& a, bool wide, int cus, bool dry, hipStream_t st) {
    if (wide) return launch16_t<true, XF, false, PROF>(a, cus, dry, st);
    return a.corr24 ? launch16_t<false, XF, true, PROF>(a, cus, dry, st) : launch16_t<false, XF, false, PROF>(a, cus, dry, st);
}

int launch_team16(const TeamArgs& a, bool wide, int xflavor, int cus, bool dry, hipStream_t st) {
    if (a.prof) return xflavor == 1 ? launch16_x<1, true>(a, wide, cus, dry, st) : launch16_x<3, true>(a, wide, cus, dry, st);
    return xflavor == 1 ? launch16_x<1, false>(a, wide, cus, dry, st) : launch16_x<3, false>(a, wide, cus, dry, st);
}

}  // namespace anc
