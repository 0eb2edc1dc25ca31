// panslbm2_amd/csrc/collide_ad3_f16.hip -- the AD collides on the 3D lattices, FP16 storage (FP32
// arithmetic).
#include "collide.h"

namespace lbm {
template hipError_t ad_collide_nd<3, kF16>(const CollideLaunch &cl, const ADArgs &a);
}  // namespace lbm
