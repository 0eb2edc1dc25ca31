// panslbm2_amd/csrc/collide_ans3_f32.hip -- the ANS collides on the 3D lattices, FP32 storage and
// arithmetic.
#include "collide.h"

namespace lbm {
template hipError_t ans_collide_nd<3, kF32>(const CollideLaunch &cl, const ANSArgs &a);
}  // namespace lbm
