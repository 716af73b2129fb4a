// Instantiations of the 16-member team kernel (anc_team16.hpp): int32 / int64 thresholds, exchange flavours 1 and 3.
#include "anc_team.hpp"
#include "anc_team16.hpp"

namespace anc {

template <bool WIDE, int XF>
static int launch16_t(const TeamArgs& a, int cus, bool dry, hipStream_t st) {
    const auto kern = dev::team16_kernel<WIDE, XF>;
    const int grid = ((a.n_teams + 7) / 8) * 8 * 16;
    size_t pad = 0;
    if (int rc = cached_pad((const void*)kern, pad)) return rc;
    if (a.n_teams > 16 || grid > cus) return -2;
    if (dry) return 0;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(dev::THREADS), pad, st, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_team16(const TeamArgs& a, bool wide, int xflavor, int cus, bool dry, hipStream_t st) {
    if (xflavor == 1) return wide ? launch16_t<true, 1>(a, cus, dry, st) : launch16_t<false, 1>(a, cus, dry, st);
    return wide ? launch16_t<true, 3>(a, cus, dry, st) : launch16_t<false, 3>(a, cus, dry, st);
}

}  // namespace anc
