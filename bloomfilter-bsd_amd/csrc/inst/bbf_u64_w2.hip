// Kernels of the sectorized blocked Bloom filter, 64-bit words, w = 2.
#include "../dtl_launch.hpp"

namespace dtlh {
void dtl_reg_u64_w2(std::vector<BbfEntry>& out) { register_sectorized<uint64_t, 2>(out); }
}  // namespace dtlh
