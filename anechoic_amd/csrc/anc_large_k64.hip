// Instantiations of the large-instance kernel for a kernel slot count of 16384 (64 rounds per lane), temperature-scaled correction.
#include "anc_large.hpp"

namespace anc {
int launch_large_k64(const LargeArgs& a, bool wide, int cus, bool dry, hipStream_t st) {
    return launch_large_k<64, false>(a, wide, cus, dry, st);
}
}  // namespace anc
