// panslbm2_amd/csrc/collide_ad2_f16.hip -- the AD collides on the 2D lattices, FP16 storage (FP32
// arithmetic).
#include "collide.h"

namespace lbm {
template hipError_t ad_collide_nd<2, kF16>(const CollideLaunch &cl, const ADArgs &a);
}  // namespace lbm
