// panslbm2_amd/csrc/collide_ans.hip -- launcher of the ANS (adjointnavierstokes.h) collides (one translation unit per
// equation so the kernel instantiations compile in parallel).
#include "collide.h"

namespace lbm {

hipError_t launch_ans_collide(const CollideLaunch &cl, const ANSArgs &a) {
    double omega = 1.0 / (3.0 * a.nu + 0.5);
    return launch_collide<ANSBody, false>(cl, [&](auto la, auto lb) {
        ANSBody<decltype(la), decltype(lb)> b;
        b.rho = a.rho; b.ux = a.ux; b.uy = a.uy; b.uz = a.uz; b.alpha = a.alpha;
        b.ip = a.ip; b.iux = a.iux; b.iuy = a.iuy; b.iuz = a.iuz; b.imx = a.imx; b.imy = a.imy; b.imz = a.imz;
        b.omega = omega; b.iomega = 1.0 - omega; b.save = a.save;
        return b;
    });
}

}  // namespace lbm
