// panslbm2_amd/csrc/collide_ad3_f32.hip -- the AD collides on the 3D lattices, FP32 storage and
// arithmetic.
#include "collide.h"

namespace lbm {
template hipError_t ad_collide_nd<3, kF32>(const CollideLaunch &cl, const ADArgs &a);
}  // namespace lbm
