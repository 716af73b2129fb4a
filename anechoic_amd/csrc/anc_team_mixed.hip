// Instantiations of the mixed launch (both team shapes on disjoint XCDs, uneven batches).
#include "anc_team.hpp"

namespace anc {

template <int NT, int NT3>
static int mixed_wide(const TeamArgs& a1, const TeamArgs& a2, const TeamArgs& a3, int xcd2, int rot, bool wide, bool x3, int cus, bool dry,
                       hipStream_t st) {
    if (wide) return x3 ? launch_mixed_t<NT, NT3, true, 3>(a1, a2, a3, xcd2, rot, cus, dry, st) : launch_mixed_t<NT, NT3, true, 1>(a1, a2, a3, xcd2, rot, cus, dry, st);
    return x3 ? launch_mixed_t<NT, NT3, false, 3>(a1, a2, a3, xcd2, rot, cus, dry, st) : launch_mixed_t<NT, NT3, false, 1>(a1, a2, a3, xcd2, rot, cus, dry, st);
}

int launch_team_mixed(const TeamArgs& a1, const TeamArgs& a2, const TeamArgs& a3, int xcd2, int rot, int nt, int nt3, bool wide,
                      int xflavor, int cus, bool dry, hipStream_t st) {
    if (a2.np != 2048 || a3.np != 3072 || xcd2 < 1 || xcd2 > 8 || (xcd2 == 8 && a2.xmode == 0) || nt3 < 1 || nt3 > 2) return -3;
    if (a2.xmode == 0) {
        if ((8 - xcd2) % 2) return -3;   // 64-workgroup teams take XCD pairs
        return launch_team_mixed_flip(a1, a2, a3, xcd2, nt, wide, xflavor, cus, dry, st);   // (no classes in turn)
    }
    if (a2.xmode == 2) return launch_team_mixed_crumb(a1, a2, a3, xcd2, rot, nt, nt3, wide, xflavor, cus, dry, st);
    const bool x3 = xflavor != 1;
    if (nt == 1) return nt3 == 1 ? mixed_wide<1, 1>(a1, a2, a3, xcd2, rot, wide, x3, cus, dry, st) : mixed_wide<1, 2>(a1, a2, a3, xcd2, rot, wide, x3, cus, dry, st);
    return nt3 == 1 ? mixed_wide<2, 1>(a1, a2, a3, xcd2, rot, wide, x3, cus, dry, st) : mixed_wide<2, 2>(a1, a2, a3, xcd2, rot, wide, x3, cus, dry, st);
}

}  // namespace anc
