// panslbm2_amd/csrc/collide_aad2_f32.hip -- the AAD collides on the 2D lattices, FP32 storage and
// arithmetic.
#include "collide.h"

namespace lbm {
template hipError_t aad_collide_nd<2, kF32>(const CollideLaunch &cl, const AADArgs &a);
}  // namespace lbm
