// panslbm2_amd/csrc/collide_aad2_f16.hip -- the AAD collides on the 2D lattices, FP16 storage (FP32
// arithmetic).
#include "collide.h"

namespace lbm {
template hipError_t aad_collide_nd<2, kF16>(const CollideLaunch &cl, const AADArgs &a);
}  // namespace lbm
