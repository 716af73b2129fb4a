// Instantiations of the fp4 (small-integer couplings) 32-workgroup team kernel: replica tiles 16 / 32, int32 / int64
// thresholds, exchange flavours 1-3.
#include "anc_team.hpp"

namespace anc {

template <int NT, bool WIDE>
static int flavour32_f4(const TeamArgs& a, int xflavor, int cus, bool dry, hipStream_t st) {
    if (xflavor == 2) return launch_team_t<32, 8, NT, WIDE, 2, true>(a, cus, dry, st);
    if (xflavor == 3) return launch_team_t<32, 8, NT, WIDE, 3, true>(a, cus, dry, st);
    return launch_team_t<32, 8, NT, WIDE, 1, true>(a, cus, dry, st);
}

int launch_team32_f4(const TeamArgs& a, int nt, bool wide, int xflavor, int cus, bool dry, hipStream_t st) {
    if (nt == 1) return wide ? flavour32_f4<1, true>(a, xflavor, cus, dry, st) : flavour32_f4<1, false>(a, xflavor, cus, dry, st);
    return wide ? flavour32_f4<2, true>(a, xflavor, cus, dry, st) : flavour32_f4<2, false>(a, xflavor, cus, dry, st);
}

}  // namespace anc
