// panslbm2_amd/csrc/collide_ns2_f32.hip -- the NS collides on the 2D lattices, FP32 storage and
// arithmetic.
#include "collide.h"

namespace lbm {
template hipError_t ns_collide_nd<2, kF32>(const CollideLaunch &cl, const NSArgs &a);
}  // namespace lbm
