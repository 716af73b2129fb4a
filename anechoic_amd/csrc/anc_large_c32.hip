// Instantiations of the large-instance kernel for a kernel slot count of 8192 (32 rounds per lane), counted correction or n_lin trace.
#include "anc_large.hpp"

namespace anc {
int launch_large_c32(const LargeArgs& a, bool wide, int cus, bool dry, hipStream_t st) {
    return launch_large_k<32, true>(a, wide, cus, dry, st);
}
}  // namespace anc
