"""CPU-oracle replica of panslbm2_amd/heatsink2d.py (same workload, same step order, same
residual test), used to check the HIP driver value for value."""
import ctypes

import numpy as np

from _oracle import D2Q9, NC, Lattice, lib, ptr, to_face

D = ctypes.c_double


def run(lx, ly, ad_kind, seed=0, nu=0.1, pr=6.0, umax=0.05, heat=1e-3, alpha_max=1e4, qf=0.01, qg=1.0,
        filter_radius=3.0, eps=1e-6, check_every=100, max_steps=400000):
    O = lib()
    N = lx * ly
    i, j = np.meshgrid(np.arange(lx), np.arange(ly), indexing="xy")
    i, j = i.reshape(-1), j.reshape(-1)
    rng = np.random.default_rng(seed)
    raw = rng.uniform(0.0, 1.0, N)
    s = np.zeros(N)
    O.ol_density_filter(lx, ly, 1, D(filter_radius), ptr(raw), ptr(s))
    dfl = nu / pr
    dso = 10.0 * dfl
    alpha = alpha_max / (ly - 1.0) * qf * (1.0 - s) / (s + qf)
    dads = -alpha_max / (ly - 1.0) * qf * (1.0 + qf) / (s + qf) ** 2
    diff = dso + (dfl - dso) * s * (1.0 + qg) / (s + qg)
    dkds = (dfl - dso) * qg * (1.0 + qg) / (s + qg) ** 2
    hsrc = heat * (1.0 - s)
    H = ly - 1.0
    uin = 4.0 * umax * j * (H - j) / H ** 2
    inlet = ((i == 0) & (j > 0) & (j < ly - 1)).astype(np.int32)
    outlet = ((i == lx - 1) & (j > 0) & (j < ly - 1)).astype(np.int32)
    walls = ((j == 0) | (j == ly - 1)).astype(np.int32)
    flux = (walls | outlet).astype(np.int32)
    objedge = ((j == 0) & (i < lx // 2)).astype(np.int32)
    Ff = lambda k, a, dt=np.float64: to_face(k, lx, ly, 1, a, dt)
    pf, pg = Lattice(D2Q9, lx, ly), Lattice(ad_kind, lx, ly)
    z = lambda: np.zeros(N)
    rho, tem = np.ones(N), z()
    U = [[z(), z()], [z(), z()]]  # U[b] = (ux, uy)
    Q = [[z(), z()], [z(), z()]]
    g = np.zeros((NC[ad_kind], N))
    O.ol_ns_init(pf.h, ptr(rho), ptr(U[0][0]), ptr(U[0][1]), None)
    O.ol_ad_init(pg.h, ptr(tem), ptr(U[0][0]), ptr(U[0][1]), None)
    t = 0
    while True:
        t += 1
        b = t % 2
        O.ol_ad_collide_fc(pf.h, ptr(rho), ptr(U[b][0]), ptr(U[b][1]), None, ptr(alpha), D(nu), pg.h, ptr(tem),
                           ptr(Q[b][0]), ptr(Q[b][1]), None, ptr(diff), 1, ptr(g) if t % check_every == 0 else None,
                           ptr(hsrc))
        if t % check_every == 0:
            ru = O.ol_residual(2, ptr(U[b][0]), ptr(U[b][1]), None, ptr(U[1 - b][0]), ptr(U[1 - b][1]), None, N)
            rq = O.ol_residual(2, ptr(Q[b][0]), ptr(Q[b][1]), None, ptr(Q[1 - b][0]), ptr(Q[1 - b][1]), None, N)
            if (ru < eps and rq < eps) or t >= max_steps:
                break
        O.ol_stream(pf.h)
        O.ol_stream(pg.h)
        O.ol_bc(pf.h, ptr(Ff(D2Q9, walls, np.int32)), 0)
        O.ol_ns_set_u(pf.h, ptr(Ff(D2Q9, inlet, np.int32)), ptr(Ff(D2Q9, uin)), ptr(Ff(D2Q9, z())), ptr(Ff(D2Q9, z())))
        O.ol_ns_set_rho(pf.h, ptr(Ff(D2Q9, outlet, np.int32)), ptr(Ff(D2Q9, np.ones(N))), ptr(Ff(D2Q9, z())),
                        ptr(Ff(D2Q9, z())))
        O.ol_smooth_corner(pf.h)
        O.ol_ad_set_t(pg.h, ptr(Ff(ad_kind, inlet, np.int32)), ptr(Ff(ad_kind, z())), ptr(U[b][0]), ptr(U[b][1]), None)
        O.ol_ad_set_q(pg.h, ptr(Ff(ad_kind, flux, np.int32)), ptr(Ff(ad_kind, z())), ptr(U[b][0]), ptr(U[b][1]), None,
                      ptr(diff), D(0.0))
        O.ol_smooth_corner(pg.h)
    fwd_steps, fb = t, t % 2
    out = {"fwd_f": pf.get(), "fwd_g": pg.get(), "rho": rho.copy(), "tem": tem.copy(), "ux": U[fb][0].copy(),
           "uy": U[fb][1].copy(), "g": g.copy(), "fwd_steps": fwd_steps, "fwd_res": (ru, rq)}
    ux, uy = U[fb]
    ip, imx, imy, item = z(), z(), z(), z()
    IU = [[z(), z()], [z(), z()]]
    IQ = [[z(), z()], [z(), z()]]
    ig = np.zeros((NC[ad_kind], N))
    O.ol_ans_init(pf.h, ptr(ux), ptr(uy), None, ptr(ip), ptr(IU[0][0]), ptr(IU[0][1]), None)
    O.ol_aad_init(pg.h, ptr(ux), ptr(uy), None, ptr(item), ptr(IQ[0][0]), ptr(IQ[0][1]), None)
    otype = Ff(D2Q9, outlet * 2, np.int32)
    t = 0
    while True:
        t += 1
        b = t % 2
        O.ol_aad_collide_fc(pf.h, ptr(rho), ptr(ux), ptr(uy), None, ptr(ip), ptr(IU[b][0]), ptr(IU[b][1]), None,
                            ptr(imx), ptr(imy), None, ptr(alpha), D(nu), pg.h, ptr(tem), ptr(item), ptr(IQ[b][0]),
                            ptr(IQ[b][1]), None, ptr(diff), 1, ptr(ig) if t % check_every == 0 else None)
        if t % check_every == 0:
            ru = O.ol_residual(2, ptr(IU[b][0]), ptr(IU[b][1]), None, ptr(IU[1 - b][0]), ptr(IU[1 - b][1]), None, N)
            rq = O.ol_residual(2, ptr(IQ[b][0]), ptr(IQ[b][1]), None, ptr(IQ[1 - b][0]), ptr(IQ[1 - b][1]), None, N)
            if (ru < eps and rq < eps) or t >= max_steps:
                break
        O.ol_istream(pf.h)
        O.ol_istream(pg.h)
        O.ol_aad_iset_t(pg.h, ptr(Ff(ad_kind, inlet, np.int32)), ptr(ux), ptr(uy), None)
        O.ol_aad_iset_q(pg.h, ptr(Ff(ad_kind, flux, np.int32)), ptr(ux), ptr(uy), None, D(0.0))
        O.ol_aad_iset_q(pg.h, ptr(Ff(ad_kind, objedge, np.int32)), ptr(ux), ptr(uy), None, D(1.0))
        O.ol_smooth_corner(pg.h)
        O.ol_bc(pf.h, ptr(Ff(D2Q9, walls, np.int32)), 1)
        O.ol_ans_iset_u(pf.h, ptr(Ff(D2Q9, inlet, np.int32)), ptr(Ff(D2Q9, uin)), ptr(Ff(D2Q9, z())),
                        ptr(Ff(D2Q9, z())), D(0.0), 1)
        O.ol_aad_iset_rho(pf.h, pg.h, ptr(otype), ptr(rho), ptr(ux), ptr(uy), None, ptr(tem), D(0.0))
        O.ol_smooth_corner(pf.h)
    ab = t % 2
    dfds = z()
    O.ol_sens_brinkman_diffusivity(pg.h, ptr(dfds), ptr(ux), ptr(uy), None, ptr(imx), ptr(imy), None, ptr(dads),
                                   ptr(tem), ptr(item), ptr(IQ[ab][0]), ptr(IQ[ab][1]), None, ptr(g), ptr(ig),
                                   ptr(diff), ptr(dkds))
    dhds = np.full(N, -heat)
    O.ol_sens_heat_source(pg.h, ptr(dfds), ptr(ux), ptr(uy), None, ptr(item), ptr(IQ[ab][0]), ptr(IQ[ab][1]), None,
                          ptr(diff), ptr(dhds))
    objsum = 0.0
    for n in np.nonzero(objedge)[0]:
        objsum += tem[n]
    out.update({"adj_f": pf.get(), "adj_g": pg.get(), "ip": ip, "imx": imx, "imy": imy, "item": item,
                "iqx": IQ[ab][0].copy(), "iqy": IQ[ab][1].copy(), "ig": ig, "dfds": dfds,
                "objective": objsum / int(objedge.sum()), "adj_steps": t, "adj_res": (ru, rq), "gamma": s})
    return out


def run3d(lx, ly, lz, ad_kind, seed=0, nu=0.1, pr=6.0, uin=0.03, heat=1e-3, plate=8, alpha_max=1e4, qf=0.01,
          qg=1.0, filter_radius=2.0, eps=1e-6, check_every=100, max_steps=400000, smooth_corners=False, gamma=None,
          inlet_rho=None):
    """CPU-oracle replica of panslbm2_amd/heatsink3d.py (Heatsink3D, D3Q15 NS): the duct's design,
    coefficients, forward and adjoint step order (with smooth_corners, SmoothCorner on both lattices
    after the boundary calls), residual test and sensitivity."""
    from _oracle import D3Q15

    O = lib()
    N = lx * ly * lz
    n = np.arange(N)
    i, j, k = n % lx, (n // lx) % ly, n // (lx * ly)
    rng = np.random.default_rng(seed)
    raw = np.clip(rng.uniform(0.0, 1.0, N), 0.0, 1.0)
    s = np.zeros(N)
    O.ol_density_filter(lx, ly, lz, D(filter_radius), ptr(raw), ptr(s))
    s[k < plate] = 0.0
    if gamma is not None:  # a given design (e.g. one a design loop produced) instead of the seeded one
        s = np.ascontiguousarray(gamma, dtype=np.float64).copy()
    dfl = nu / pr
    dso = 10.0 * dfl
    alpha = alpha_max / (ly - 1.0) * qf * (1.0 - s) / (s + qf)
    dads = -alpha_max / (ly - 1.0) * qf * (1.0 + qf) / (s + qf) ** 2
    diff = dso + (dfl - dso) * s * (1.0 + qg) / (s + qg)
    dkds = (dfl - dso) * qg * (1.0 + qg) / (s + qg) ** 2
    hsrc = np.where(k < plate, heat, 0.0)
    wall = (j == 0) | (j == ly - 1) | (k == 0) | (k == lz - 1)
    inlet = ((i == 0) & ~wall).astype(np.int32)
    outlet = ((i == lx - 1) & ~wall).astype(np.int32)
    flux = ((i > 0) | wall).astype(np.int32)
    bottom = (k == 0).astype(np.int32)
    F = lambda kind, a, dt=np.float64: to_face(kind, lx, ly, lz, a, dt)
    zf, ones = np.zeros(N), np.ones(N)
    fw, fin, fout = F(D3Q15, wall.astype(np.int32), np.int32), F(D3Q15, inlet, np.int32), F(D3Q15, outlet, np.int32)
    fuin, fz, f1 = F(D3Q15, np.full(N, uin)), F(D3Q15, zf), F(D3Q15, ones)
    gin, gfl, gobj, gz = F(ad_kind, inlet, np.int32), F(ad_kind, flux, np.int32), F(ad_kind, bottom, np.int32), \
        F(ad_kind, zf)
    pf, pg = Lattice(D3Q15, lx, ly, lz), Lattice(ad_kind, lx, ly, lz)
    z = lambda: np.zeros(N)
    rho, tem = np.ones(N), z()
    U = [[z(), z(), z()], [z(), z(), z()]]
    Q = [[z(), z(), z()], [z(), z(), z()]]
    g = np.zeros((NC[ad_kind], N))
    O.ol_ns_init(pf.h, ptr(rho), *[ptr(a) for a in U[0]])
    O.ol_ad_init(pg.h, ptr(tem), *[ptr(a) for a in U[0]])
    res3 = lambda a, b: O.ol_residual(3, *[ptr(x) for x in a], *[ptr(x) for x in b], N)
    t = 0
    while True:
        t += 1
        b = t % 2
        O.ol_ad_collide_fc(pf.h, ptr(rho), *[ptr(a) for a in U[b]], ptr(alpha), D(nu), pg.h, ptr(tem),
                           *[ptr(a) for a in Q[b]], ptr(diff), 1, ptr(g) if t % check_every == 0 else None, ptr(hsrc))
        if t % check_every == 0:
            ru, rq = res3(U[b], U[1 - b]), res3(Q[b], Q[1 - b])
            if (ru < eps and rq < eps) or t >= max_steps:
                break
        O.ol_stream(pf.h)
        O.ol_stream(pg.h)
        O.ol_bc(pf.h, ptr(fw), 0)
        if inlet_rho is None:
            O.ol_ns_set_u(pf.h, ptr(fin), ptr(fuin), ptr(fz), ptr(fz))
        else:
            O.ol_ns_set_rho(pf.h, ptr(fin), ptr(F(D3Q15, np.full(N, inlet_rho))), ptr(fz), ptr(fz))
        O.ol_ns_set_rho(pf.h, ptr(fout), ptr(f1), ptr(fz), ptr(fz))
        O.ol_ad_set_t(pg.h, ptr(gin), ptr(gz), *[ptr(a) for a in U[b]])
        O.ol_ad_set_q(pg.h, ptr(gfl), ptr(gz), *[ptr(a) for a in U[b]], ptr(diff), D(0.0))
        if smooth_corners:
            O.ol_smooth_corner(pf.h)
            O.ol_smooth_corner(pg.h)
    fwd_steps, fb = t, t % 2
    out = {"fwd_f": pf.get(), "fwd_g": pg.get(), "rho": rho.copy(), "tem": tem.copy(), "g": g.copy(),
           "fwd_steps": fwd_steps, "fwd_res": (ru, rq), "gamma": s}
    for a, name in enumerate(("ux", "uy", "uz")):
        out[name] = U[fb][a].copy()
    uf = U[fb]
    ip, item = z(), z()
    IM = [z(), z(), z()]
    IU = [[z(), z(), z()], [z(), z(), z()]]
    IQ = [[z(), z(), z()], [z(), z(), z()]]
    ig = np.zeros((NC[ad_kind], N))
    O.ol_ans_init(pf.h, *[ptr(a) for a in uf], ptr(ip), *[ptr(a) for a in IU[0]])
    O.ol_aad_init(pg.h, *[ptr(a) for a in uf], ptr(item), *[ptr(a) for a in IQ[0]])
    t = 0
    while True:
        t += 1
        b = t % 2
        O.ol_aad_collide_fc(pf.h, ptr(rho), *[ptr(a) for a in uf], ptr(ip), *[ptr(a) for a in IU[b]],
                            *[ptr(a) for a in IM], ptr(alpha), D(nu), pg.h, ptr(tem), ptr(item),
                            *[ptr(a) for a in IQ[b]], ptr(diff), 1, ptr(ig) if t % check_every == 0 else None)
        if t % check_every == 0:
            ru, rq = res3(IU[b], IU[1 - b]), res3(IQ[b], IQ[1 - b])
            if (ru < eps and rq < eps) or t >= max_steps:
                break
        O.ol_istream(pf.h)
        O.ol_istream(pg.h)
        O.ol_aad_iset_t(pg.h, ptr(gin), *[ptr(a) for a in uf])
        O.ol_aad_iset_q(pg.h, ptr(gfl), *[ptr(a) for a in uf], D(0.0))
        O.ol_aad_iset_q(pg.h, ptr(gobj), *[ptr(a) for a in uf], D(1.0))
        O.ol_bc(pf.h, ptr(fw), 1)
        if inlet_rho is None:
            O.ol_ans_iset_u(pf.h, ptr(fin), ptr(fuin), ptr(fz), ptr(fz), D(0.0), 1)
        else:
            O.ol_ans_iset_rho(pf.h, ptr(fin))
        O.ol_ans_iset_rho(pf.h, ptr(fout))
        if smooth_corners:
            O.ol_smooth_corner(pf.h)
            O.ol_smooth_corner(pg.h)
    ab = t % 2
    dfds = z()
    O.ol_sens_brinkman_diffusivity(pg.h, ptr(dfds), *[ptr(a) for a in uf], *[ptr(a) for a in IM], ptr(dads),
                                   ptr(tem), ptr(item), *[ptr(a) for a in IQ[ab]], ptr(g), ptr(ig), ptr(diff),
                                   ptr(dkds))
    objsum = 0.0
    for m in np.nonzero(bottom)[0]:
        objsum += tem[m]
    out.update({"adj_f": pf.get(), "adj_g": pg.get(), "ip": ip, "imx": IM[0], "imy": IM[1], "imz": IM[2],
                "item": item, "iqx": IQ[ab][0].copy(), "iqy": IQ[ab][1].copy(), "iqz": IQ[ab][2].copy(), "ig": ig,
                "dfds": dfds, "objective": objsum / int(bottom.sum()), "adj_steps": t, "adj_res": (ru, rq)})
    return out


def loop3d(lx, ly, lz, iterations, vf=0.4, move=0.05, plate=3, box=None, filter_radius=2.0, stages=None,
           stage_iterations=100, stage_tol=0.01, x0=None, **kw):
    """CPU replica of Heatsink3D.optimize(optimizer="mma") with a Heaviside projection, built from the
    reference's own pieces: HeavisideFilter with the design-region weights (oracle/_ref
    ref3_heaviside_*_region, heatsink3D.cpp:87-93,103), run3d for each evaluation, and the
    reference's MMA (oracle/_ref refmma_*, mma.h) with one volume constraint, as Heatsink3D drives them:
    gamma = projection of x with the base plate solid and fluid outside the box; dF/dgamma zero where
    gamma is set after the filter; sensitivities through the filter; the objective's gradient
    normalised to a largest magnitude of 1; (qf, beta) continuation.  Returns per iteration
    (objective, x, gamma, dfds)."""
    from _oracle import D3Q7, ref

    R = ref()
    if R is None:
        raise RuntimeError("reference harness not built")
    VP, I = ctypes.c_void_p, ctypes.c_int
    R.ref3_heaviside_value_region.argtypes = [I, I, I, D, D, I, I, I, VP, VP]
    R.ref3_heaviside_sensitivity_region.argtypes = [I, I, I, D, D, I, I, I, VP, VP, VP]
    R.refmma_create.restype = VP
    R.refmma_create.argtypes = [I, I, D, VP, VP, VP, VP, VP, D]
    R.refmma_update.argtypes = [VP, I, I, VP, D, VP, VP, VP]
    R.refmma_destroy.argtypes = [VP]
    N = lx * ly * lz
    idx = np.arange(N)
    i, j, k = idx % lx, (idx // lx) % ly, idx // (lx * ly)
    box = tuple(box) if box is not None else (lx, ly, lz)
    in_box = (i < box[0]) & (j < box[1]) & (k < box[2])
    region = in_box & (k >= plate)
    rid = np.nonzero(region)[0]
    n = len(rid)
    stages = stages or ((kw.pop("qf", 0.01), 1.0),)
    stage, cnt = 0, 0
    qf, beta = stages[0]

    def clean(x):
        x = np.clip(np.asarray(x, dtype=np.float64), 0.0, 1.0).copy()
        x[~in_box] = 1.0
        return x

    def gamma_of(x):
        s = np.zeros(N)
        R.ref3_heaviside_value_region(lx, ly, lz, D(filter_radius), D(beta), *box, ptr(x), ptr(s))
        s[~in_box] = 1.0
        s[k < plate] = 0.0
        return s

    def sens(x, d):
        out = np.zeros(N)
        R.ref3_heaviside_sensitivity_region(lx, ly, lz, D(filter_radius), D(beta), *box, ptr(x),
                                            ptr(np.ascontiguousarray(d)), ptr(out))
        return out

    x = clean(np.where(region, 1.0 - vf, 0.0) if x0 is None else x0)
    z = np.zeros(1)
    mma = R.refmma_create(n, 1, D(1.0), ptr(z), ptr(np.full(1, 1e4)), ptr(z), ptr(np.zeros(n)), ptr(np.ones(n)),
                          D(move))
    hist = []
    try:
        for it in range(iterations):
            s = gamma_of(x)
            r = run3d(lx, ly, lz, D3Q7, plate=plate, qf=qf, filter_radius=filter_radius, gamma=s, **kw)
            hist.append((r["objective"], x.copy(), s, r["dfds"]))
            if it + 1 == iterations:
                break
            dfdx = sens(x, np.where(region, r["dfds"], 0.0))[rid]
            dgdx = sens(x, np.where(region, -1.0 / (vf * n), 0.0))[rid]
            g = np.array([(1.0 - s[rid]).sum() / (vf * n) - 1.0])
            dfdx = dfdx / max(np.abs(dfdx).max(), 1e-300)
            xr = np.ascontiguousarray(x[rid])
            R.refmma_update(mma, n, 1, ptr(xr), D(r["objective"]), ptr(dfdx), ptr(g), ptr(np.ascontiguousarray(dgdx)))
            x_old = x.copy()
            x[rid] = xr
            x = clean(x)
            if len(stages) > 1:
                cnt += 1
                dsmax = float(np.abs(x - x_old).max())
                if stage == len(stages) - 1 and dsmax < stage_tol:
                    break
                if dsmax < stage_tol or cnt % stage_iterations == 0:
                    stage, cnt = min(stage + 1, len(stages) - 1), 0
                    qf, beta = stages[stage]
    finally:
        R.refmma_destroy(mma)
    return hist
