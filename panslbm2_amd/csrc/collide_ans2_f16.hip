// panslbm2_amd/csrc/collide_ans2_f16.hip -- the ANS collides on the 2D lattices, FP16 storage (FP32
// arithmetic).
#include "collide.h"

namespace lbm {
template hipError_t ans_collide_nd<2, kF16>(const CollideLaunch &cl, const ANSArgs &a);
}  // namespace lbm
