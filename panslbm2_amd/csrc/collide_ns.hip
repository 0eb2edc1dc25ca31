// panslbm2_amd/csrc/collide_ns.hip -- launcher of the NS (navierstokes.h) collides (one translation unit per
// equation so the kernel instantiations compile in parallel).
#include "collide.h"

namespace lbm {

hipError_t launch_ns_collide(const CollideLaunch &cl, const NSArgs &a) {
    double omega = 1.0 / (3.0 * a.nu + 0.5);
    return launch_collide<NSBody, false>(cl, [&](auto la, auto lb) {
        NSBody<decltype(la), decltype(lb)> b;
        b.rho = a.rho; b.ux = a.ux; b.uy = a.uy; b.uz = a.uz; b.alpha = a.alpha;
        b.omega = omega; b.iomega = 1.0 - omega; b.save = a.save; b.brinkman = a.alpha != nullptr;
        return b;
    });
}

}  // namespace lbm
