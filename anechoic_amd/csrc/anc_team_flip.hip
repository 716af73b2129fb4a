// Instantiations of the team kernels with the flip-matrix exchange (option "exchange" 0: D' = D + tau bytes,
// accumulated field): 32-workgroup, 64-workgroup and mixed launches.
#include "anc_team.hpp"

namespace anc {

template <int NT, bool WIDE>
static int flavour32_flip(const TeamArgs& a, int xflavor, int cus, bool dry, hipStream_t st) {
    if (xflavor == 2) return launch_team_t<32, 8, NT, WIDE, 2, 0>(a, cus, dry, st);
    if (xflavor == 3) return launch_team_t<32, 8, NT, WIDE, 3, 0>(a, cus, dry, st);
    return launch_team_t<32, 8, NT, WIDE, 1, 0>(a, cus, dry, st);
}

int launch_team32_flip(const TeamArgs& a, int nt, bool wide, int xflavor, int cus, bool dry, hipStream_t st) {
    if (nt == 1) return wide ? flavour32_flip<1, true>(a, xflavor, cus, dry, st) : flavour32_flip<1, false>(a, xflavor, cus, dry, st);
    return wide ? flavour32_flip<2, true>(a, xflavor, cus, dry, st) : flavour32_flip<2, false>(a, xflavor, cus, dry, st);
}

int launch_team64_flip(const TeamArgs& a, int nt, bool wide, int cus, bool dry, hipStream_t st) {
    if (nt == 1) return wide ? launch_team_t<64, 12, 1, true, 1, 0>(a, cus, dry, st) : launch_team_t<64, 12, 1, false, 1, 0>(a, cus, dry, st);
    return wide ? launch_team_t<64, 12, 2, true, 1, 0>(a, cus, dry, st) : launch_team_t<64, 12, 2, false, 1, 0>(a, cus, dry, st);
}

template <int NT>
static int mixed_wide_flip(const TeamArgs& a1, const TeamArgs& a2, const TeamArgs& a3, int xcd2, int rot, bool wide, bool x3, int cus, bool dry,
                       hipStream_t st) {
    if (wide) return x3 ? launch_mixed_t<NT, NT, true, 3, 0>(a1, a2, a3, xcd2, rot, cus, dry, st) : launch_mixed_t<NT, NT, true, 1, 0>(a1, a2, a3, xcd2, rot, cus, dry, st);
    return x3 ? launch_mixed_t<NT, NT, false, 3, 0>(a1, a2, a3, xcd2, rot, cus, dry, st) : launch_mixed_t<NT, NT, false, 1, 0>(a1, a2, a3, xcd2, rot, cus, dry, st);
}

int launch_team_mixed_flip(const TeamArgs& a1, const TeamArgs& a2, const TeamArgs& a3, int xcd2, int nt, bool wide, int xflavor,
                           int cus, bool dry, hipStream_t st) {
    const bool x3 = xflavor != 1;
    return nt == 1 ? mixed_wide_flip<1>(a1, a2, a3, xcd2, 0, wide, x3, cus, dry, st) : mixed_wide_flip<2>(a1, a2, a3, xcd2, 0, wide, x3, cus, dry, st);
}

}  // namespace anc
