// Kernels of the zoned blocked Bloom filter, 64-bit words, w = 4.
#include "../dtl_launch.hpp"

namespace dtlh {
void dtl_reg_z64_w4(std::vector<BbfEntry>& out) { register_zoned<uint64_t, 4>(out); }
}  // namespace dtlh
