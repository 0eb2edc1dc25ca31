// panslbm2_amd/csrc/collide_ad2_f32.hip -- the AD collides on the 2D lattices, FP32 storage and
// arithmetic.
#include "collide.h"

namespace lbm {
template hipError_t ad_collide_nd<2, kF32>(const CollideLaunch &cl, const ADArgs &a);
}  // namespace lbm
