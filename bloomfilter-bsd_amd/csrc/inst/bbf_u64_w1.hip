// Kernels of the sectorized blocked Bloom filter, 64-bit words, w = 1.
#include "../dtl_launch.hpp"

namespace dtlh {
void dtl_reg_u64_w1(std::vector<BbfEntry>& out) { register_sectorized<uint64_t, 1>(out); }
}  // namespace dtlh
