// panslbm2_amd/csrc/collide_ns3_f32.hip -- the NS collides on the 3D lattices, FP32 storage and
// arithmetic.
#include "collide.h"

namespace lbm {
template hipError_t ns_collide_nd<3, kF32>(const CollideLaunch &cl, const NSArgs &a);
}  // namespace lbm
