// panslbm2_amd/csrc/collide_aad.hip -- launcher of the AAD (adjointadvection.h) collides (one translation unit per
// equation so the kernel instantiations compile in parallel).
#include "collide.h"

namespace lbm {

hipError_t launch_aad_collide(const CollideLaunch &cl, const AADArgs &a) {
    double omega = 1.0 / (3.0 * a.nu + 0.5);
    return launch_collide<AADBody, true>(cl, [&](auto la, auto lb) {
        AADBody<decltype(la), decltype(lb)> b;
        b.rho = a.rho; b.ux = a.ux; b.uy = a.uy; b.uz = a.uz; b.alpha = a.alpha; b.tem = a.tem; b.diff = a.diff;
        b.ip = a.ip; b.iux = a.iux; b.iuy = a.iuy; b.iuz = a.iuz; b.imx = a.imx; b.imy = a.imy; b.imz = a.imz;
        b.item = a.item; b.iqx = a.iqx; b.iqy = a.iqy; b.iqz = a.iqz; b.igsnap = a.igsnap;
        b.omegaf = omega; b.iomegaf = 1.0 - omega; b.save = a.save; b.N = cl.g.N;
        return b;
    });
}

}  // namespace lbm
