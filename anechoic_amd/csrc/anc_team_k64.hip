// Instantiations of the NP = 3072 team kernels (n = 2049..3072) with the spin exchange: 32 workgroups (one XCD) x 12
// rounds, 16- or 32-replica work items, int32 / int64 thresholds, exchange flavours 1-3. (The flip exchange runs these
// instances on 64-workgroup teams: anc_team_flip.hip.)
#include "anc_team.hpp"

namespace anc {

template <int NT, bool WIDE>
static int flavour3072(const TeamArgs& a, int xflavor, int cus, bool dry, hipStream_t st) {
    if (xflavor == 2) return launch_team_t<32, 12, NT, WIDE, 2, 1>(a, cus, dry, st);
    if (xflavor == 3) return launch_team_t<32, 12, NT, WIDE, 3, 1>(a, cus, dry, st);
    return launch_team_t<32, 12, NT, WIDE, 1, 1>(a, cus, dry, st);
}

int launch_team3072(const TeamArgs& a, int nt, bool wide, int xflavor, int cus, bool dry, hipStream_t st) {
    if (a.xmode == 0) return launch_team64_flip(a, nt, wide, cus, dry, st);
    if (nt == 1) return wide ? flavour3072<1, true>(a, xflavor, cus, dry, st) : flavour3072<1, false>(a, xflavor, cus, dry, st);
    if (nt == 2) return wide ? flavour3072<2, true>(a, xflavor, cus, dry, st) : flavour3072<2, false>(a, xflavor, cus, dry, st);
    return -3;
}

}  // namespace anc
