// panslbm2_amd/csrc/collide_ad.hip -- launcher of the AD (advection.h) collides (one translation unit per
// equation so the kernel instantiations compile in parallel).
#include "collide.h"

namespace lbm {

hipError_t launch_ad_collide(const CollideLaunch &cl, const ADArgs &a) {
    double omega = 1.0 / (3.0 * a.nu + 0.5);
    return launch_collide<ADBody, true>(cl, [&](auto la, auto lb) {
        ADBody<decltype(la), decltype(lb)> b;
        b.rho = a.rho; b.ux = a.ux; b.uy = a.uy; b.uz = a.uz; b.tem = a.tem; b.qx = a.qx; b.qy = a.qy; b.qz = a.qz;
        b.gsnap = a.gsnap; b.alpha = a.alpha; b.diff = a.diff; b.heat = a.heat;
        b.omegaf = omega; b.iomegaf = 1.0 - omega; b.save = a.save; b.N = cl.g.N;
        return b;
    });
}

}  // namespace lbm
