// panslbm2_amd/csrc/collide_aad3_f16.hip -- the AAD collides on the 3D lattices, FP16 storage (FP32
// arithmetic).
#include "collide.h"

namespace lbm {
template hipError_t aad_collide_nd<3, kF16>(const CollideLaunch &cl, const AADArgs &a);
}  // namespace lbm
