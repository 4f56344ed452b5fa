// Kernels of the zoned blocked Bloom filter, 64-bit words, w = 8.
#include "../dtl_launch.hpp"

namespace dtlh {
void dtl_reg_z64_w8(std::vector<BbfEntry>& out) { register_zoned<uint64_t, 8>(out); }
}  // namespace dtlh
