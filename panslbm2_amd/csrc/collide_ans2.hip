// panslbm2_amd/csrc/collide_ans2.hip -- the ANS collides on the 2D lattices, FP64 storage (one
// translation unit per equation, dimension and precision so the kernel instantiations compile in
// parallel).
#include "collide.h"

namespace lbm {
template hipError_t ans_collide_nd<2, kF64>(const CollideLaunch &cl, const ANSArgs &a);
}  // namespace lbm
