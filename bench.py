#!/usr/bin/env python3
"""bench.py -- design evaluations of the 2D forced-convection heatsink on one MI355X.

One step = one design evaluation: forward solve to a 1e-6 step-to-step residual (checked every
100 steps), adjoint solve to the same residual, design sensitivity (see
panslbm2_amd/heatsink2d.py).  Workload: 200 x 400 nodes, D2Q9 Navier-Stokes with Brinkman drag
coupled to D2Q5 advection-diffusion with a volumetric heat source, FP64, seeded synthetic design.

value = MLUPS over the job: nodes x (forward + adjoint time steps) / wall seconds, a coupled
NS+AD node update counting once.  Inputs are resident on the device before timing starts.

Also reported:
  roofline      -- the fused forward step kernel (pull + boundary program + collide) timed with
                   HIP events on the library's stream over 100-step graph launches, against the
                   8 TB/s HBM peak, for this workload (cache resident: ~20 MB per step) and for a
                   4096 x 4096 grid (4.2 GB per step, HBM bound).
  cpu_baseline  -- the reference's own plain C++ (oracle/_ref, OpenMP) running the same forward
                   and adjoint time steps on D2Q9 + D2Q9 (the reference has no D2Q5) on the host
                   cores, a bounded number of steps.
  fp32, fp16    -- the same evaluation with FP32 / FP16 population storage (mixed precision: the
                   reduced precision until the residual reaches its floor or stops falling, FP64 to
                   1e-6), and the reduced-precision fused step on 4096^2.
  batch         -- 64 independent designs (grid sizes from {64x128, 128x256, 200x400, 256x512},
                   heat source and inlet velocity drawn per design) evaluated together, one kernel
                   launch per time step for all of them (FP32 mixed, autotuned workgroup shape),
                   against evaluating them one after another in plain FP64.
  warm_start    -- a 3D design loop (D3Q15 + D3Q7 heat sink, 96 x 96 x 48, solid fraction 0.4,
                   projected-gradient updates: the MMA loop's designs make the adjoint diverge,
                   from the cold start too, after ~6 iterations at this alpha_max) run four ways: cold start in FP64 (the plain GPU
                   form), previous-solution warm start, learned warm start (U-Net on the matrix
                   cores predicting the change of the fields, trained beside the solves), and the
                   learned warm start with FP32 mixed-precision storage; wall seconds and LBM steps
                   per design iteration after the first, every solve to the 1e-6 residual.
  transient     -- the transient natural-convection heat sink of the reference's transient
                   program (141 x 161, 2000 steps forward and backward, the adjoint over the stored
                   trajectory): seconds per evaluation with the trajectory resident in HBM and with
                   checkpoints + recomputed windows.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per node update of the fused forward step, FP64: populations read once and
# written once (D2Q9 + D2Q5 = 14 doubles each way) plus alpha, diffusivity and heat source read.
BYTES_FWD_Q9Q5 = 2 * (9 + 5) * 8 + 3 * 8
BYTES_FWD_Q9Q9 = 2 * (9 + 9) * 8 + 3 * 8
BYTES_FWD_Q9Q5_F32 = 2 * (9 + 5) * 4 + 3 * 8  # FP32 populations, FP64 coefficient fields
BYTES_FWD_Q9Q5_F16 = 2 * (9 + 5) * 2 + 3 * 8  # FP16 populations
# Memory-side traffic of one forward step of the 200x400 workload, from rocprofv3 FETCH_SIZE and
# WRITE_SIZE (KiB per dispatch; profiles/k_collide_fused_heatsink_200x400.txt), FETCH_SIZE scaled by
# 1.920, the factor that makes it read the known bytes of the pull-stream copy kernel
# (profiles/pmc_scale_k_stream_4096x4096.txt).  The counters see traffic below the L2, which on this
# cache-sized step includes Infinity Cache hits.
TRAFFIC_FWD_200x400 = round((5767.5 * 1.920 + 8920.5) * 1024)


def kernel_rate(hs, steps=100, reps=5):
    """Average duration (ms) of one fused forward step, from HIP events around graph launches."""
    # make sure the forward chunk graph exists and state is mid-solve
    hs._chunk("fwd", 100, hs._fwd_step)
    hs.ctx.sync()
    hs.ctx.record(0)
    for _ in range(reps):
        hs._chunk("fwd", 100, hs._fwd_step)
    hs.ctx.record(1)
    ms = hs.ctx.elapsed_ms(0, 1)
    return ms / (reps * hs.check_every)


def big_grid_roofline(ctx, lx=4096, ly=4096, use_graphs=True, prec=0):
    from panslbm2_amd import D2Q5
    from panslbm2_amd.heatsink2d import Heatsink2D

    hs = Heatsink2D(ctx, lx=lx, ly=ly, ad_kind=D2Q5, max_steps=200, filter_radius=1.0, use_graphs=use_graphs,
                    prec=prec)
    L, d = hs.L, hs.d
    from panslbm2_amd import check

    check(L.lbm_ns_initial_condition(hs.pf.h, d["one"].ptr, d["ux0"].ptr, d["uy0"].ptr, None))
    check(L.lbm_ad_initial_condition(hs.pg.h, d["tem"].ptr, d["ux0"].ptr, d["uy0"].ptr, None))
    for s in range(1, 101):
        hs._fwd_step(s)
    ms = kernel_rate(hs, reps=2)
    hs.close()
    return ms, lx * ly


def cpu_baseline(hs, target_seconds=15.0):
    """The reference's plain OpenMP path on the host cores (oracle/_ref), D2Q9 + D2Q9."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np

    from _oracle import lib as oracle_lib
    from _oracle import ref

    R = ref()
    if R is None:
        return None
    lx, ly, N = hs.lx, hs.ly, hs.N
    i, j = hs.i, hs.j
    walls = ((j == 0) | (j == ly - 1)).astype(np.int32)
    inlet = ((i == 0) & (j > 0) & (j < ly - 1)).astype(np.int32)
    outlet = ((i == lx - 1) & (j > 0) & (j < ly - 1)).astype(np.int32)
    flux = (walls | outlet).astype(np.int32)
    objedge = ((j == 0) & (i < lx // 2)).astype(np.int32)
    H = ly - 1.0
    uin = np.ascontiguousarray(4.0 * 0.05 * j * (H - j) / H ** 2)
    alpha = np.ascontiguousarray(hs.d["alpha"].download())
    diff = np.ascontiguousarray(hs.d["diff"].download())
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    secs = np.zeros(2)
    f = R.ref2_heatsink_steps
    f.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double] + [ctypes.c_void_p] * 9
    nt = 20
    f(lx, ly, nt, nt, 0.1, P(alpha), P(diff), P(walls), P(inlet), P(outlet), P(flux), P(objedge), P(uin), P(secs))
    per = secs.sum() / (2 * nt)
    nt = max(20, int(target_seconds / max(per, 1e-9) / 2))
    f(lx, ly, nt, nt, 0.1, P(alpha), P(diff), P(walls), P(inlet), P(outlet), P(flux), P(objedge), P(uin), P(secs))
    mlups = N * 2 * nt / secs.sum() / 1e6
    return {"value": round(mlups, 3), "unit": "MLUPS", "cores": int(oracle_lib().ol_num_threads()),
            "kind": "reference",
            "sample": "%d forward + %d adjoint steps of the same 200x400 heatsink with the reference's plain "
                      "OpenMP functions on D2Q9+D2Q9 (the reference has no D2Q5): fwd %.2f s, adj %.2f s"
                      % (nt, nt, secs[0], secs[1])}


def fp32_extra(ctx, a, r64, prec=None):
    """The benchmark evaluation with FP32 (or FP16) storage (mixed precision), and the reduced-
    precision fused step's bandwidth on 4096^2."""
    from panslbm2_amd import D2Q5, F16, F32
    from panslbm2_amd.heatsink2d import Heatsink2D

    prec = F32 if prec is None else prec
    hs = Heatsink2D(ctx, lx=a.lx, ly=a.ly, ad_kind=D2Q5, max_steps=a.max_steps, use_graphs=not a.no_graphs,
                    prec=prec)
    hs.evaluate()  # captures the graphs
    r = hs.evaluate()
    out = {"mlups": round(hs.N * (r["fwd_steps"] + r["adj_steps"]) / r["seconds"] / 1e6, 1),
           "seconds": round(r["seconds"], 4), "fwd_steps": r["fwd_steps"], "adj_steps": r["adj_steps"],
           "steps_in_reduced_precision": list(r["fp32_steps"]),
           "objective_rel_err_vs_fp64": abs(r["objective"] - r64["objective"]) / abs(r64["objective"]),
           "residuals": [max(r["fwd_residual"]), max(r["adj_residual"])]}
    hs.close()
    if not a.no_big:
        ms, n = big_grid_roofline(ctx, use_graphs=not a.no_graphs, prec=prec)
        ach = (BYTES_FWD_Q9Q5_F32 if prec == F32 else BYTES_FWD_Q9Q5_F16) * n / (ms * 1e-3) / 1e9
        out["roofline_hbm_4096x4096"] = {"achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                         "frac": round(ach / HBM_PEAK_GBS, 4), "us_per_launch": round(ms * 1e3, 1),
                                         "mlups": round(n / (ms * 1e-3) / 1e6, 1)}
    return out


def batch_extra(ctx, n, seed=2024):
    """n designs evaluated together (batched FP32 mixed) and one after another (plain FP64)."""
    import numpy as np

    from panslbm2_amd import D2Q5, F32, F64
    from panslbm2_amd.heatsink2d import Heatsink2D
    from panslbm2_amd.heatsink_batch import HeatsinkBatch2D

    rng = np.random.default_rng(seed)
    sizes = [(64, 128), (128, 256), (200, 400), (256, 512)]
    specs = []
    for i in range(n):
        lx, ly = sizes[rng.integers(0, 4)]
        specs.append(dict(lx=lx, ly=ly, seed=100 + i, heat=float(rng.uniform(5e-4, 2e-3)),
                          umax=float(rng.uniform(0.02, 0.06))))
    hb = HeatsinkBatch2D(ctx, specs, prec=F32, tune=True)
    rb = hb.evaluate()
    hb.close()
    seq_s, seq_updates, objs, div = 0.0, 0, [], []
    for sp in specs:
        h = Heatsink2D(ctx, lx=sp["lx"], ly=sp["ly"], ad_kind=D2Q5, seed=sp["seed"], heat=sp["heat"], umax=sp["umax"],
                       prec=F64)
        q = h.evaluate()
        seq_s += q["seconds"]
        seq_updates += h.N * (q["fwd_steps"] + q["adj_steps"])
        objs.append(q["objective"])
        div.append(q["diverged"])
        h.close()
    err = max(abs(x - y) / abs(y) for x, y in zip(rb["objectives"], objs))
    bdiv = [c == -2 for c in rb["fwd_conv"] + rb["adj_conv"]]
    return {"designs": n, "seconds": round(rb["seconds"], 3), "sequential_fp64_seconds": round(seq_s, 3),
            "speedup_vs_sequential_fp64": round(seq_s / rb["seconds"], 2),
            "mlups": round(rb["node_steps"] / rb["seconds"] / 1e6, 1),
            "sequential_fp64_mlups": round(seq_updates / seq_s / 1e6, 1),
            "steps": [rb["fwd_steps"], rb["adj_steps"]], "steps_in_fp32": list(rb["fp32_steps"]),
            "workgroup_and_xcd_remap": list(rb["tuned"]) if rb["tuned"] else None,
            "max_objective_rel_err_vs_fp64": err,
            "diverged_fp64_sequential": int(sum(div)), "diverged_batched": int(sum(bdiv)),
            "note": "a solve whose residual turns non-finite stops there in both runs (the adjoint of "
                    "designs with u_max near 0.06 diverges in FP64 too); objectives agree to the "
                    "spread a 1e-6 step-to-step residual leaves (~1e-4 of the slowest mode)"}


def transient_extra(ctx, nt=2000):
    """panslbm2_amd/transient2d.py on the reference program's grid (second evaluation timed)."""
    from panslbm2_amd.transient2d import TransientHeatsink2D

    out = {"grid": [141, 161], "steps": nt, "lattices": "D2Q9 NS + D2Q9 AD", "precision": "fp64"}
    for storage in ("full", "checkpoint"):
        drv = TransientHeatsink2D(ctx, 141, 161, nt, seed=1, storage=storage)
        drv.evaluate()
        t0 = time.perf_counter()
        f, _ = drv.evaluate()
        dt = time.perf_counter() - t0
        out[storage] = {"s_per_evaluation": round(dt, 4), "window": drv.K,
                        "mlups": round(drv.N * (2 * nt - 3) / dt / 1e6, 1), "objective": f}
        drv.close()
    return out


def warm_extra(ctx, iters, train_steps=20, use_graphs=True):
    """The 96x96x48 3D design loop, cold vs warm-started (see the module docstring)."""
    import numpy as np

    from panslbm2_amd import F32, F64
    from panslbm2_amd.heatsink3d import Heatsink3D

    out = {"grid": [96, 96, 48], "lattices": "D3Q15 NS + D3Q7 AD", "iterations": iters, "solid_fraction": 0.4,
           "train_steps_per_iteration": train_steps}
    for mode, prec, learn in (("cold_fp64", F64, None), ("previous_solution_fp64", F64, False),
                              ("learned_fp64", F64, True), ("learned_fp32", F32, True)):
        hs = Heatsink3D(ctx, lx=96, ly=96, lz=48, plate=6, prec=prec, use_graphs=use_graphs)
        if learn is not None:
            hs.attach_warm_start(seed=0, steps=train_steps, learn=learn)
        t0 = time.perf_counter()
        hist = hs.optimize(iters, warm=(False if learn is None else None), optimizer="pg")
        total = time.perf_counter() - t0
        later = hist[1:]
        out[mode] = {"s_per_iteration": round(float(np.mean([r["seconds"] + r["train_seconds"] for r in later])), 4),
                     "steps_per_iteration": int(np.mean([r["fwd_steps"] + r["adj_steps"] for r in later])),
                     "total_s": round(total, 2), "objective_last": hist[-1]["objective"],
                     "fallbacks": int(sum(any(r["fallback"]) for r in hist)),
                     "all_converged": bool(all(r["converged"] for r in hist))}
        hs.close()
        print("bench: warm start %s %.3f s/iteration" % (mode, out[mode]["s_per_iteration"]), file=sys.stderr,
              flush=True)
    base = out["cold_fp64"]["s_per_iteration"]
    for mode in ("previous_solution_fp64", "learned_fp64", "learned_fp32"):
        out[mode]["speedup_vs_cold_fp64"] = round(base / out[mode]["s_per_iteration"], 2)
    out["note"] = ("per iteration after the first (the first is a cold start in every mode); the designs of the "
                   "four loops drift apart slightly because each solve stops at the 1e-6 residual")
    return out


def unet_only(ctx, a):
    """The warm-start network at 96x96x48 (4 input, 13 output fields): a.steps forward passes,
    backward passes and Adam updates, host-timed around its own stream, with the forward FLOPs."""
    import numpy as np

    from panslbm2_amd.warmstart import UNet

    dims = (96, 96, 48)
    V = dims[0] * dims[1] * dims[2]
    net = UNet(ctx, dims, 4, 13, seed=0)
    rng = np.random.default_rng(0)
    din = [ctx.array((V,), data=rng.uniform(0, 1, V)) for _ in range(4)]
    dout = [ctx.array((V,), data=rng.standard_normal(V)) for _ in range(13)]
    net.set_input(din)
    net.set_target(dout)
    c1, c2, c3 = 16, 32, 64
    convs = [(0, 4, c1, 27), (0, c1, c1, 27), (1, c1, c2, 27), (1, c2, c2, 27), (2, c2, c3, 27), (2, c3, c3, 27),
             (1, c3 + c2, c2, 27), (1, c2, c2, 27), (0, c2 + c1, c1, 27), (0, c1, c1, 27), (0, c1, 13, 1)]
    flop = sum(2.0 * (V >> (3 * lvl)) * taps * ci * co for lvl, ci, co, taps in convs)
    res = {}
    for name, fn in (("forward", net.forward), ("backward", lambda: net.backward(want_loss=False)),
                     ("adam", lambda: net.adam(1e-4))):
        fn()
        net.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            fn()
        net.synchronize()
        res[name + "_ms"] = round((time.perf_counter() - t0) * 1e3 / a.steps, 3)
    res["forward_gflop"] = round(flop / 1e9, 2)
    res["forward_tflops"] = round(flop / (res["forward_ms"] * 1e-3) / 1e12, 1)
    res["backward_tflops"] = round(2 * flop / (res["backward_ms"] * 1e-3) / 1e12, 1)
    net.close()
    print(json.dumps({"kernel_only": "unet", "grid": list(dims), "passes": a.steps, **res}))


def kernel_only(ctx, a):
    """Launch one kernel kind a.steps times (eager, no graphs) and print its event-timed average."""
    from panslbm2_amd import D2Q5, D2Q9, Lattice, check
    from panslbm2_amd.heatsink2d import Heatsink2D

    if a.kernel_only == "unet":
        return unet_only(ctx, a)
    if a.kernel_only == "stream":
        lat = Lattice(ctx, D2Q9, a.lx, a.ly)
        check(ctx.lib.lbm_stream(lat.h))
        ctx.sync()
        ctx.record(0)
        for _ in range(a.steps):
            check(ctx.lib.lbm_stream(lat.h))
        ctx.record(1)
        ms = ctx.elapsed_ms(0, 1) / a.steps
        nbytes = 2 * 9 * 8 * a.lx * a.ly
    else:
        hs = Heatsink2D(ctx, lx=a.lx, ly=a.ly, ad_kind=D2Q5, max_steps=a.steps, use_graphs=False, filter_radius=1.0)
        L, d = hs.L, hs.d
        check(L.lbm_ns_initial_condition(hs.pf.h, d["one"].ptr, d["ux0"].ptr, d["uy0"].ptr, None))
        check(L.lbm_ad_initial_condition(hs.pg.h, d["tem"].ptr, d["ux0"].ptr, d["uy0"].ptr, None))
        if a.kernel_only == "adj":
            hs.fwd_buf = 0
            hs.adj_prog = hs._record_adjoint(0)
            step = lambda t: hs._adj_step(t, 0)
            nbytes = (2 * (9 + 5) * 8 + 2 * 8 + 4 * 8) * hs.N
        else:
            step = hs._fwd_step
            nbytes = BYTES_FWD_Q9Q5 * hs.N
        for t in range(1, 3):
            step(t)
        ctx.sync()
        ctx.record(0)
        for t in range(3, 3 + a.steps):
            step(t)
        ctx.record(1)
        ms = ctx.elapsed_ms(0, 1) / a.steps
        hs.close()
    print(json.dumps({"kernel_only": a.kernel_only, "grid": [a.lx, a.ly], "launches": a.steps,
                      "us_per_launch": round(ms * 1e3, 3), "algorithmic_bytes": nbytes,
                      "achieved_GBs": round(nbytes / (ms * 1e-3) / 1e9, 1)}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--lx", type=int, default=200)
    ap.add_argument("--ly", type=int, default=400)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-big", action="store_true")
    ap.add_argument("--max-steps", type=int, default=400000, help="cap on forward and on adjoint steps")
    ap.add_argument("--no-graphs", action="store_true", help="launch every step (for profilers)")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-transient", action="store_true")
    ap.add_argument("--batch", type=int, default=64, help="designs in the batched evaluation (0: skip it)")
    ap.add_argument("--warm-iters", type=int, default=10, help="3D design-loop iterations per warm-start mode (0: skip)")
    ap.add_argument("--kernel-only", choices=["fwd", "adj", "stream", "unet"],
                    help="launch only the fused forward step, the fused adjoint step or the pull-stream copy "
                         "--steps times on --lx x --ly (profiling and counter calibration), or time the "
                         "warm-start network's passes at 96x96x48")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (requested --gpus %d)" % a.gpus,
              file=sys.stderr)
        sys.exit(2)

    from panslbm2_amd import D2Q5, D2Q9, Context
    from panslbm2_amd.heatsink2d import Heatsink2D

    ctx = Context(0)
    if a.kernel_only:
        return kernel_only(ctx, a)
    hs = Heatsink2D(ctx, lx=a.lx, ly=a.ly, ad_kind=D2Q5, max_steps=a.max_steps, use_graphs=not a.no_graphs)
    print("bench: heatsink %dx%d D2Q9+D2Q5 FP64, warmup %d, steps %d" % (a.lx, a.ly, a.warmup, a.steps),
          file=sys.stderr, flush=True)
    for w in range(a.warmup):
        r = hs.evaluate()
        print("warmup %d: %.3f s fwd %d adj %d" % (w, r["seconds"], r["fwd_steps"], r["adj_steps"]), file=sys.stderr,
              flush=True)
    ctx.sync()
    t0 = time.perf_counter()
    total_updates = 0
    runs = []
    for s in range(a.steps):
        r = hs.evaluate()
        runs.append(r)
        total_updates += hs.N * (r["fwd_steps"] + r["adj_steps"])
        print("step %d: %.3f s fwd %d adj %d obj %.6e" % (s, r["seconds"], r["fwd_steps"], r["adj_steps"],
                                                           r["objective"]), file=sys.stderr, flush=True)
    ctx.sync()
    wall = time.perf_counter() - t0
    mlups = total_updates / wall / 1e6

    ms_step = kernel_rate(hs)
    achieved = BYTES_FWD_Q9Q5 * hs.N / (ms_step * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": TRAFFIC_FWD_200x400 if (a.lx, a.ly) == (200, 400) else None,
                "kernel": "k_collide_fused<D2Q9,D2Q5,ADBody> (pull + boundary program + corner smoothing + "
                          "collide in one pass), per fused forward step",
                "bytes_per_launch": BYTES_FWD_Q9Q5 * hs.N, "us_per_launch": round(ms_step * 1e3, 3),
                "note": "working set ~20 MB per step: resident in the 256 MB Infinity Cache"}
    extra = {}
    if not a.no_big:
        ms_big, nbig = big_grid_roofline(ctx, use_graphs=not a.no_graphs)
        ach_big = BYTES_FWD_Q9Q5 * nbig / (ms_big * 1e-3) / 1e9
        extra["roofline_hbm_4096x4096"] = {"achieved": round(ach_big, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                           "frac": round(ach_big / HBM_PEAK_GBS, 4),
                                           "us_per_launch": round(ms_big * 1e3, 1),
                                           "mlups": round(nbig / (ms_big * 1e-3) / 1e6, 1)}
    if not a.no_fp32:
        from panslbm2_amd import F16

        extra["fp32"] = fp32_extra(ctx, a, runs[-1])
        extra["fp16"] = fp32_extra(ctx, a, runs[-1], prec=F16)
    if a.batch > 0:
        extra["batch"] = batch_extra(ctx, a.batch)
    if a.warm_iters > 0:
        extra["warm_start"] = warm_extra(ctx, a.warm_iters, use_graphs=not a.no_graphs)
    if not a.no_transient:
        extra["transient"] = transient_extra(ctx)
    # like-for-like with the reference's lattice pairing (D2Q9 advection-diffusion)
    hs9 = Heatsink2D(ctx, lx=a.lx, ly=a.ly, ad_kind=D2Q9, max_steps=a.max_steps, use_graphs=not a.no_graphs)
    r9 = hs9.evaluate()
    extra["mlups_d2q9_d2q9"] = round(hs9.N * (r9["fwd_steps"] + r9["adj_steps"]) / r9["seconds"] / 1e6, 1)
    cpu = None if a.no_cpu else cpu_baseline(hs9)
    hs9.close()

    out = {
        "metric": "MLUPS (fwd+adjoint) & % HBM peak; wall-time to steady, warm vs cold",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall / a.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded uniform design field, density-filtered)",
        "config": {"workload": "2D heatsink design evaluation (forward+adjoint to 1e-6, sensitivity), cold start",
                   "grid": [a.lx, a.ly], "lattices": "D2Q9 NS + D2Q5 AD", "precision": "fp64",
                   "fwd_steps": runs[-1]["fwd_steps"], "adj_steps": runs[-1]["adj_steps"],
                   "wall_s_to_steady_cold": round(runs[-1]["seconds"], 3)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    hs.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
