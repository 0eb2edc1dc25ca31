// panslbm2_amd/csrc/collide_ans3.hip -- the ANS collides on the 3D lattices, FP64 storage (one
// translation unit per equation, dimension and precision so the kernel instantiations compile in
// parallel).
#include "collide.h"

namespace lbm {
template hipError_t ans_collide_nd<3, kF64>(const CollideLaunch &cl, const ANSArgs &a);
}  // namespace lbm
