// Instantiations of the 32-workgroup (one XCD, NP = 2048) team kernel: replica tiles 16 / 32, int32 / int64
// thresholds, exchange flavours 1-3.
#include "anc_team.hpp"

namespace anc {

template <int NT, bool WIDE>
static int flavour32(const TeamArgs& a, int xflavor, int cus, bool dry, hipStream_t st) {
    if (xflavor == 2) return launch_team_t<32, 8, NT, WIDE, 2>(a, cus, dry, st);
    if (xflavor == 3) return launch_team_t<32, 8, NT, WIDE, 3>(a, cus, dry, st);
    return launch_team_t<32, 8, NT, WIDE, 1>(a, cus, dry, st);
}

int launch_team32(const TeamArgs& a, int nt, bool wide, int xflavor, int cus, bool dry, hipStream_t st) {
    if (a.xmode == 0) return launch_team32_flip(a, nt, wide, xflavor, cus, dry, st);
    if (a.xmode == 2) return launch_team32_crumb(a, nt, wide, xflavor, cus, dry, st);
    if (nt == 1) return wide ? flavour32<1, true>(a, xflavor, cus, dry, st) : flavour32<1, false>(a, xflavor, cus, dry, st);
    return wide ? flavour32<2, true>(a, xflavor, cus, dry, st) : flavour32<2, false>(a, xflavor, cus, dry, st);
}

}  // namespace anc
