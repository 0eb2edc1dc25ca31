// panslbm2_amd/csrc/collide_ad3.hip -- the AD collides on the 3D lattices (one translation unit
// per equation and dimension so the kernel instantiations compile in parallel).
#include "collide.h"

namespace lbm {
template hipError_t ad_collide_nd<3>(const CollideLaunch &cl, const ADArgs &a);
}  // namespace lbm
