// panslbm2_amd/csrc/collide_ad3.hip -- the AD collides on the 3D lattices, FP64 storage (one
// translation unit per equation, dimension and precision so the kernel instantiations compile in
// parallel).
#include "collide.h"

namespace lbm {
template hipError_t ad_collide_nd<3, kF64>(const CollideLaunch &cl, const ADArgs &a);
}  // namespace lbm
