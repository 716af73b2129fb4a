// Instantiations of the large-instance kernel for a kernel slot count of 4096 (16 rounds per lane), temperature-scaled correction.
#include "anc_large.hpp"

namespace anc {
int launch_large_k16(const LargeArgs& a, bool wide, int cus, bool dry, hipStream_t st) {
    return launch_large_k<16, false>(a, wide, cus, dry, st);
}
}  // namespace anc
