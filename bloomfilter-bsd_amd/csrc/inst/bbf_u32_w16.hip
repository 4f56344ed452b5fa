// Kernels of the sectorized blocked Bloom filter, 32-bit words, w = 16.
#include "../dtl_launch.hpp"

namespace dtlh {
void dtl_reg_u32_w16(std::vector<BbfEntry>& out) { register_sectorized<uint32_t, 16>(out); }
}  // namespace dtlh
