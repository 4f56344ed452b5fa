// Kernels of the sectorized blocked Bloom filter, 32-bit words, w = 2.
#include "../dtl_launch.hpp"

namespace dtlh {
void dtl_reg_u32_w2(std::vector<BbfEntry>& out) { register_sectorized<uint32_t, 2>(out); }
}  // namespace dtlh
