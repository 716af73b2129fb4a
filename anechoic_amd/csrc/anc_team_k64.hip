// Instantiations of the 64-workgroup (two XCDs, NP = 2304..3072) team kernel: replica tiles 16 / 32, int32 / int64
// thresholds, device-scope exchange (a team spans two XCD L2s).
#include "anc_team.hpp"

namespace anc {

int launch_team64(const TeamArgs& a, int nt, bool wide, int cus, bool dry, hipStream_t st) {
    if (a.xmode == 0) return launch_team64_flip(a, nt, wide, cus, dry, st);
    if (nt == 1) return wide ? launch_team_t<64, 12, 1, true, 1>(a, cus, dry, st) : launch_team_t<64, 12, 1, false, 1>(a, cus, dry, st);
    return wide ? launch_team_t<64, 12, 2, true, 1>(a, cus, dry, st) : launch_team_t<64, 12, 2, false, 1>(a, cus, dry, st);
}

}  // namespace anc
