// Kernels of the sectorized blocked Bloom filter, 32-bit words, w = 8.
#include "../dtl_launch.hpp"

namespace dtlh {
void dtl_reg_u32_w8(std::vector<BbfEntry>& out) { register_sectorized<uint32_t, 8>(out); }
}  // namespace dtlh
