// Kernels of the zoned blocked Bloom filter, 32-bit words, w = 16.
#include "../dtl_launch.hpp"

namespace dtlh {
void dtl_reg_z32_w16(std::vector<BbfEntry>& out) { register_zoned<uint32_t, 16>(out); }
}  // namespace dtlh
