// Instantiations of the 32-workgroup (one XCD, NP = 2048) team kernel with the crumb exchange (option `exchange` 2:
// two bits per spin, N = s + 1 + tau): replica tiles 16 / 32, int32 / int64 thresholds, exchange flavours 1-3.
#include "anc_team.hpp"

namespace anc {

template <int NT, bool WIDE>
static int crumb32(const TeamArgs& a, int xflavor, int cus, bool dry, hipStream_t st) {
    if (xflavor == 2) return launch_team_t<32, 8, NT, WIDE, 2, 3>(a, cus, dry, st);
    if (xflavor == 3) return launch_team_t<32, 8, NT, WIDE, 3, 3>(a, cus, dry, st);
    return launch_team_t<32, 8, NT, WIDE, 1, 3>(a, cus, dry, st);
}

int launch_team32_crumb(const TeamArgs& a, int nt, bool wide, int xflavor, int cus, bool dry, hipStream_t st) {
    if (nt == 1) return wide ? crumb32<1, true>(a, xflavor, cus, dry, st) : crumb32<1, false>(a, xflavor, cus, dry, st);
    return wide ? crumb32<2, true>(a, xflavor, cus, dry, st) : crumb32<2, false>(a, xflavor, cus, dry, st);
}

template <int NT, int NT3>
static int mixed_crumb(const TeamArgs& a1, const TeamArgs& a2, const TeamArgs& a3, int xcd2, int rot, bool wide, bool x3, int cus, bool dry,
                       hipStream_t st) {
    if (wide) return x3 ? launch_mixed_t<NT, NT3, true, 3, 3>(a1, a2, a3, xcd2, rot, cus, dry, st) : launch_mixed_t<NT, NT3, true, 1, 3>(a1, a2, a3, xcd2, rot, cus, dry, st);
    return x3 ? launch_mixed_t<NT, NT3, false, 3, 3>(a1, a2, a3, xcd2, rot, cus, dry, st) : launch_mixed_t<NT, NT3, false, 1, 3>(a1, a2, a3, xcd2, rot, cus, dry, st);
}

// nt / nt3: 16-replica tiles of the NP 2048 / NP 3072 work items
int launch_team_mixed_crumb(const TeamArgs& a1, const TeamArgs& a2, const TeamArgs& a3, int xcd2, int rot, int nt, int nt3,
                            bool wide, int xflavor, int cus, bool dry, hipStream_t st) {
    const bool x3 = xflavor != 1;
    if (nt == 1) return nt3 == 1 ? mixed_crumb<1, 1>(a1, a2, a3, xcd2, rot, wide, x3, cus, dry, st) : mixed_crumb<1, 2>(a1, a2, a3, xcd2, rot, wide, x3, cus, dry, st);
    return nt3 == 1 ? mixed_crumb<2, 1>(a1, a2, a3, xcd2, rot, wide, x3, cus, dry, st) : mixed_crumb<2, 2>(a1, a2, a3, xcd2, rot, wide, x3, cus, dry, st);
}

}  // namespace anc
