// panslbm2_amd/csrc/collide_ns2_f16.hip -- the NS collides on the 2D lattices, FP16 storage (FP32
// arithmetic).
#include "collide.h"

namespace lbm {
template hipError_t ns_collide_nd<2, kF16>(const CollideLaunch &cl, const NSArgs &a);
}  // namespace lbm
