#!/usr/bin/env python3
"""bench.py -- design evaluations of the 3D heat sink on one MI355X.

One step = one design evaluation: forward solve to a 1e-6 step-to-step residual (checked every
100 steps), adjoint solve to the same residual, design sensitivity (see panslbm2_amd/heatsink3d.py).
Workload: the 3D heat sink at 128 x 128 x 64 nodes (1,048,576), D3Q15 Navier-Stokes with Brinkman
drag coupled to D3Q7 advection-diffusion, inlet u = 0.03 on the x = 0 face, heat source in the
z < 8 base plate, seeded design filtered with radius 2, FP64, cold start.

value = MLUPS over the job: nodes x (forward + adjoint time steps) / wall seconds, a coupled
NS+AD node update counting once.  Inputs are resident on the device before timing starts.

Also reported:
  roofline      -- the fused 3D forward step kernel (pull + boundary program + collide in one pass)
                   timed with HIP events on the library's stream over 100-step graph launches,
                   against the 8 TB/s HBM peak; `traffic` from rocprofv3 FETCH_SIZE / WRITE_SIZE.
  cpu_baseline  -- the reference's own plain C++ (oracle/_ref, OpenMP) running the same forward and
                   adjoint time steps on D3Q15 + D3Q15 (the reference has no D3Q7) at the same grid
                   on the host cores, a bounded number of steps.
  c3_adjoint    -- the fused 3D adjoint step kernel, timed the same way.
  c3_fp16, c3_fp32 -- the same evaluation with FP16 / FP32 population storage (collided in FP32,
                   FP64 once the residual reaches its floor), with the objective's and the
                   sensitivity's error against the FP64 evaluation.
  c3_smoothed   -- the same evaluation with SmoothCorner on both lattices every step (the
                   reference's 3D driver sequence, heatsink3D.cpp:175-176, 215-216), FP64: the
                   fused pass plus the edges and corners on a second stream concurrently with it.
  heatsink2d    -- the 2D heat sink (200 x 400, D2Q9 + D2Q5, FP64): MLUPS, its fused step's roofline,
                   the same kernel on 4096^2 (HBM bound), FP32 / FP16 variants, the reference pairing
                   D2Q9 + D2Q9 and the reference CPU path on it.
  batch         -- 64 independent 2D designs (grid sizes from {64x128, 128x256, 200x400, 256x512},
                   heat source and inlet velocity drawn per design) evaluated together, one kernel
                   launch per time step for all of them (FP32 mixed, autotuned workgroup shape),
                   against evaluating them one after another in plain FP64.
  warm_start    -- the first --warm-iters iterations of the 3D design loop (D3Q15 + D3Q7 heat sink,
                   96 x 96 x 48, MMA, solid fraction 0.4, the reference driver's continuation and
                   design sub-box) run four ways: cold start in FP64 (the plain GPU form),
                   previous-solution warm start, learned warm start (U-Net on the matrix cores
                   predicting the change of the fields, trained beside the solves), and the learned
                   warm start with FP32 mixed-precision storage; wall seconds and LBM steps per
                   design iteration after the first, every solve to the 1e-6 residual, and each
                   mode's last objective and sensitivity against cold FP64 solves of the same design
                   (to 1e-6 and to 1e-9).  --design-loop N runs only this loop, N iterations.
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
# 3D: D3Q15 + D3Q7 populations once each way, alpha, diffusivity and heat source read (forward);
# alpha, diffusivity and the forward rho, ux, uy, uz, T read (adjoint)
BYTES_FWD_Q15Q7 = 2 * (15 + 7) * 8 + 3 * 8
BYTES_ADJ_Q15Q7 = 2 * (15 + 7) * 8 + 2 * 8 + 5 * 8
# Memory-side traffic of one forward step of the 200x400 workload, from rocprofv3 FETCH_SIZE and
# WRITE_SIZE (KiB per dispatch; profiles/k_collide_fused_heatsink_200x400.txt), FETCH_SIZE scaled by
# 1.920, the factor that makes it read the known bytes of the pull-stream copy kernel
# (profiles/pmc_scale_k_stream_4096x4096.txt).  The counters see traffic below the L2, which on this
# cache-sized step includes Infinity Cache hits.
TRAFFIC_FWD_200x400 = round((5767.5 * 1.920 + 8920.5) * 1024)
# The same for the fused 3D forward and adjoint steps at 128x128x64 (profiles/k_collide_fused_3d_128x128x64.txt)
TRAFFIC_FWD_128x128x64 = round((108690.5 * 1.920 + 182828.8) * 1024)
TRAFFIC_ADJ_128x128x64 = round((120205.3 * 1.920 + 180224.0) * 1024)


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


def kernel_rate_adj(hs, reps=5):
    """Average duration (ms) of one fused adjoint step (after an evaluation: the adjoint graph exists)."""
    fb = hs.fwd_buf
    step = lambda t: hs._adj_step(t, fb)
    hs._chunk("adj", 100, step)
    hs.ctx.sync()
    hs.ctx.record(0)
    for _ in range(reps):
        hs._chunk("adj", 100, step)
    hs.ctx.record(1)
    return hs.ctx.elapsed_ms(0, 1) / (reps * hs.check_every)


def host_info():
    """The host the CPU baseline runs on: logical CPUs, model, sockets, cores, OpenMP binding."""
    import subprocess

    info = {"nproc": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True,
                             timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "Model name":
                info["model"] = v
            elif k == "Socket(s)":
                info["sockets"] = int(v)
            elif k == "Core(s) per socket":
                info["cores_per_socket"] = int(v)
            elif k == "Thread(s) per core":
                info["threads_per_core"] = int(v)
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES"):
        info[k] = os.environ.get(k)
    return info


def _omp_env():
    """Thread count = physical cores available to this process (capped by OMP_NUM_THREADS when it is
    set), bound close to cores; must run before the OpenMP runtime starts."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")


def cpu_baseline_3d(hs, target_seconds=15.0):
    """The reference's plain OpenMP path on the host cores (oracle/_ref), D3Q15 + D3Q15, the same
    128 x 128 x 64 duct: forward then adjoint steps, a bounded number."""
    _omp_env()
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np

    from _oracle import lib as oracle_lib
    from _oracle import ref

    R = ref()
    if R is None:
        return None
    lx, ly, lz, N = hs.lx, hs.ly, hs.lz, hs.N
    i, j, k = hs.i, hs.j, hs.k
    wall = (j == 0) | (j == ly - 1) | (k == 0) | (k == lz - 1)
    inlet = ((i == 0) & ~wall).astype(np.int32)
    outlet = ((i == lx - 1) & ~wall).astype(np.int32)
    flux = ((i > 0) | wall).astype(np.int32)
    bottom = (k == 0).astype(np.int32)
    walls = wall.astype(np.int32)
    alpha = np.ascontiguousarray(hs.d["alpha"].download())
    diff = np.ascontiguousarray(hs.d["diff"].download())
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    secs = np.zeros(2)
    f = R.ref3_heatsink_steps
    f.argtypes = [ctypes.c_int] * 5 + [ctypes.c_double] * 2 + [ctypes.c_void_p] * 8
    nt = 4
    f(lx, ly, lz, nt, nt, 0.1, 0.03, P(alpha), P(diff), P(walls), P(inlet), P(outlet), P(flux), P(bottom), P(secs))
    per = secs.sum() / (2 * nt)
    nt = max(4, int(target_seconds / max(per, 1e-9) / 2))
    f(lx, ly, lz, nt, nt, 0.1, 0.03, P(alpha), P(diff), P(walls), P(inlet), P(outlet), P(flux), P(bottom), P(secs))
    mlups = N * 2 * nt / secs.sum() / 1e6
    return {"value": round(mlups, 3), "unit": "MLUPS", "cores": int(oracle_lib().ol_num_threads()),
            "kind": "reference",
            "sample": "%d forward + %d adjoint steps of the same 128x128x64 duct from rest with the reference's "
                      "plain OpenMP functions on D3Q15+D3Q15 (the reference has no D3Q7): fwd %.2f s, adj %.2f s"
                      % (nt, nt, secs[0], secs[1]),
            "host": host_info()}


def c3_precision_extra(ctx, a, r64, dfds64, prec):
    """The 3D evaluation with FP16 / FP32 population storage; error against the FP64 evaluation."""
    import numpy as np

    from panslbm2_amd import D3Q7
    from panslbm2_amd.heatsink3d import Heatsink3D

    hs = Heatsink3D(ctx, a.lx3, a.ly3, a.lz3, ad_kind=D3Q7, max_steps=a.max_steps, use_graphs=not a.no_graphs,
                    prec=prec)
    hs.evaluate()  # captures the graphs
    r = hs.evaluate()
    dfds = hs.d["dfds"].download()
    out = {"mlups": round(hs.N * (r["fwd_steps"] + r["adj_steps"]) / r["seconds"] / 1e6, 1),
           "seconds": round(r["seconds"], 4), "fwd_steps": r["fwd_steps"], "adj_steps": r["adj_steps"],
           "steps_in_reduced_precision": list(r["fp32_steps"]),
           "objective_rel_err_vs_fp64": abs(r["objective"] - r64["objective"]) / abs(r64["objective"]),
           "dfds_max_err_rel_to_max_vs_fp64": float(np.abs(dfds - dfds64).max() / np.abs(dfds64).max()),
           "residuals": [max(r["fwd_residual"]), max(r["adj_residual"])]}
    hs.close()
    return out


def c3_smoothed_extra(ctx, a, r64):
    """The 3D evaluation with SmoothCorner on both lattices (FP64, the default concurrent form)."""
    from panslbm2_amd import D3Q7
    from panslbm2_amd.heatsink3d import Heatsink3D

    hs = Heatsink3D(ctx, a.lx3, a.ly3, a.lz3, ad_kind=D3Q7, max_steps=a.max_steps, use_graphs=not a.no_graphs,
                    smooth_corners=True)
    hs.evaluate()  # captures the graphs
    r = hs.evaluate()
    steps = r["fwd_steps"] + r["adj_steps"]
    out = {"mlups": round(hs.N * steps / r["seconds"] / 1e6, 1), "seconds": round(r["seconds"], 4),
           "fwd_steps": r["fwd_steps"], "adj_steps": r["adj_steps"],
           "us_per_step": round(r["seconds"] / steps * 1e6, 1),
           "unsmoothed_us_per_step": round(r64["seconds"] / (r64["fwd_steps"] + r64["adj_steps"]) * 1e6, 1),
           "objective": r["objective"],
           "note": "edges and corners smoothed every forward and adjoint step; the objective differs from "
                   "the unsmoothed one because the edge and corner states do"}
    hs.close()
    return out


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
    ok = [i for i, (x, y) in enumerate(zip(rb["objectives"], objs)) if np.isfinite(x) and np.isfinite(y)]
    err = max(abs(rb["objectives"][i] - objs[i]) / abs(objs[i]) for i in ok)
    bdiv = [a == -2 or b == -2 for a, b in zip(rb["fwd_conv"], rb["adj_conv"])]  # per design
    return {"designs": n, "seconds": round(rb["seconds"], 3), "sequential_fp64_seconds": round(seq_s, 3),
            "speedup_vs_sequential_fp64": round(seq_s / rb["seconds"], 2),
            "mlups": round(rb["node_steps"] / rb["seconds"] / 1e6, 1),
            "sequential_fp64_mlups": round(seq_updates / seq_s / 1e6, 1),
            "steps": [rb["fwd_steps"], rb["adj_steps"]], "steps_in_fp32": list(rb["fp32_steps"]),
            "workgroup_rows_xcd_remap": list(rb["tuned"]) if rb["tuned"] else None,
            "max_objective_rel_err_vs_fp64": err,
            "diverged_fp64_sequential": int(sum(div)), "diverged_batched": int(sum(bdiv)),
            "fp64_redo_batched": [len(x) for x in rb["fp64_redo"]],
            "note": "designs counted as diverged when their forward or adjoint residual turns "
                    "non-finite (the adjoint of designs with u_max near 0.06 grows without bound in both "
                    "precisions); FP32 solves that diverge are solved again in FP64, batched together "
                    "(fp64_redo_batched: forward, adjoint); objectives of the designs that converge in both runs agree to the spread "
                    "a 1e-6 step-to-step residual leaves (~1e-4 of the slowest mode)"}


def design_batch(ctx, n, iters, seed=2024, vf=0.4, move=0.05, fast_prec="F32", fast_warm=True, plain=True):
    """n 2D heat-sink designs (the batch's grid sizes, heat sources and inlet velocities), each
    optimised for `iters` projected-gradient iterations, two ways: the plain GPU form (one design
    after another, cold FP64 solves) and the fast form (all designs batched in one launch per step,
    FP32 storage carrying the solves, each solve warm-started from its design's previous converged
    solution).  Seconds per design iteration of all n designs, after the first."""
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

    def run(evaluate, hs, label):
        times, steps, objs = [], [], []
        for it in range(iters):
            t0 = time.perf_counter()
            rs = evaluate(it)
            for h, ok in zip(hs, rs["ok"]):
                if ok:
                    h.design_update(vf, move)
            ctx.sync()
            times.append(time.perf_counter() - t0)
            steps.append(rs["node_steps"])
            objs.append(rs["objectives"])
            print("bench: design batch %s it %d %.3f s" % (label, it, times[-1]), file=sys.stderr, flush=True)
        return times, steps, objs

    import panslbm2_amd as pa

    fast_label = "batched_%s_%s" % (fast_prec.lower(), "warm" if fast_warm else "cold")
    hb = HeatsinkBatch2D(ctx, specs, prec=getattr(pa, fast_prec), tune=True)

    def ev_fast(it):
        r = hb.evaluate(warm=fast_warm and it > 0)
        return {"ok": [a > 0 and b > 0 for a, b in zip(r["fwd_conv"], r["adj_conv"])], "objectives": r["objectives"],
                "node_steps": r["node_steps"]}

    tf, sf, of = run(ev_fast, hb.hs, fast_label)
    hb.close()
    if plain:
        hs = [Heatsink2D(ctx, lx=sp["lx"], ly=sp["ly"], ad_kind=D2Q5, seed=sp["seed"], heat=sp["heat"],
                         umax=sp["umax"], prec=F64) for sp in specs]

        def ev_plain(it):
            rs = [h.evaluate(warm=False) for h in hs]
            return {"ok": [not r["diverged"] for r in rs], "objectives": [r["objective"] for r in rs],
                    "node_steps": sum(h.N * (r["fwd_steps"] + r["adj_steps"]) for h, r in zip(hs, rs))}

        tp, sp_, op = run(ev_plain, hs, "plain_fp64_cold")
        for h in hs:
            h.close()
    else:
        tp, sp_, op = tf, sf, of
    later = slice(1, None) if iters > 1 else slice(0, None)
    op, of = np.array(op), np.array(of)
    ok = np.all(np.isfinite(op), axis=0) & np.all(np.isfinite(of), axis=0)
    out = {"designs": n, "iterations": iters, "solid_fraction": vf, "move": move,
           "plain_fp64_cold": {"s_per_iteration": round(float(np.mean(tp[later])), 4),
                               "node_steps_per_iteration": int(np.mean(sp_[later]))},
           fast_label: {"s_per_iteration": round(float(np.mean(tf[later])), 4),
                        "node_steps_per_iteration": int(np.mean(sf[later]))},
           "designs_converging_in_both": int(ok.sum()),
           "mean_objective_first_last": {"plain": [float(op[0, ok].mean()), float(op[-1, ok].mean())],
                                         "fast": [float(of[0, ok].mean()), float(of[-1, ok].mean())]}}
    out["speedup"] = round(out["plain_fp64_cold"]["s_per_iteration"] / out[fast_label]["s_per_iteration"], 2)
    if not plain:
        del out["plain_fp64_cold"], out["speedup"]
    out["note"] = ("per design iteration of all designs after the first; designs whose solves diverge (inlet "
                   "speeds near 0.06) keep their design and are solved again each iteration in both forms")
    return out


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
        # nt - 1 collides forward and nt - 1 backward; checkpointing recomputes each window's forward
        # steps once more during the backward pass, which this figure does not count (so there it
        # is an effective rate)
        recompute = 0 if storage == "full" else nt - 1
        out[storage] = {"s_per_evaluation": round(dt, 4), "window": drv.K,
                        "mlups": round(drv.N * (2 * nt - 2) / dt / 1e6, 1),
                        "recomputed_forward_steps": recompute, "objective": f}
        drv.close()
    return out


DESIGN_LOOP_MODES = {"cold_fp64": ("F64", None), "cold_fp32": ("F32", None),
                     "previous_solution_fp64": ("F64", False), "previous_solution_fp32": ("F32", False),
                     "learned_fp64": ("F64", True), "learned_fp32": ("F32", True)}


def design_loop(ctx, iters, modes, stage_iterations=20, train_steps=20, eps=1e-6):
    """The 96x96x48 topology-optimisation loop: D3Q15 + D3Q7 heat sink, MMA with one volume
    constraint (solid fraction <= 0.4), the reference driver's (qf, beta) continuation
    (heatsink3D.cpp:45) advanced every stage_iterations iterations, its 3/4 design sub-box
    (heatsink3D.cpp:47) with fluid around it, inlet u = 0.01, all-fluid start, every solve to the
    residual eps (1e-6, cap 60,000 steps).  Each mode runs the loop from the same start; after the
    loop the last evaluated design is evaluated again cold in FP64, once to the same residual and
    once to a 100x-1000x tighter one, and the mode's objective and sensitivity are compared with
    both."""
    import numpy as np

    import panslbm2_amd as pa
    from panslbm2_amd.heatsink3d import Heatsink3D

    box = tuple(3 * (n - 1) // 4 + 1 for n in (96, 96, 48))
    cap = 60000 if eps >= 1e-6 else 200000
    tight = min(1e-9, eps * 1e-2)
    kw = dict(lx=96, ly=96, lz=48, plate=6, heaviside_beta=1.0, design_box=box, uin=0.01, max_steps=cap, eps=eps)
    out = {"grid": [96, 96, 48], "lattices": "D3Q15 NS + D3Q7 AD", "iterations": iters, "solid_fraction": 0.4,
           "optimizer": "MMA", "stages": [list(s) for s in Heatsink3D.REFERENCE_STAGES],
           "stage_iterations": stage_iterations, "design_box": list(box), "inlet_u": 0.01,
           "train_steps_per_iteration": train_steps, "residual": eps, "tight_residual": tight}
    for mode in modes:
        prec, learn = DESIGN_LOOP_MODES[mode]
        hs = Heatsink3D(ctx, prec=getattr(pa, prec), **kw)
        if learn is not None:
            hs.attach_warm_start(seed=0, steps=train_steps, learn=learn)
        last = {}

        def cb(r, hs=hs, last=last):
            last.update(dfds=hs.d["dfds"].download(), x=hs.x.copy(), qf=hs.qf, beta=hs.heaviside_beta,
                        objective=r["objective"])
            print("design loop %s it %d stage %d obj %.6f fwd %d adj %d %.3f s%s" % (
                mode, r["iteration"], r["stage"], r["objective"], r["fwd_steps"], r["adj_steps"],
                r["seconds"] + r.get("train_seconds", 0.0), "" if r["converged"] else " NOT CONVERGED"),
                file=sys.stderr, flush=True)

        t0 = time.perf_counter()
        hist = hs.optimize(iters, vf=0.4, move=0.05, optimizer="mma", x0=np.where(hs.region, 1.0, 0.0),
                           stages=Heatsink3D.REFERENCE_STAGES, stage_iterations=stage_iterations,
                           warm=(False if learn is None else None), callback=cb)
        ctx.sync()
        total = time.perf_counter() - t0
        hs.close()
        refs = {}
        for name, e, c in (("cold_fp64", eps, cap), ("tight_fp64", tight, 600000)):
            ref = Heatsink3D(ctx, prec=pa.F64, **dict(kw, eps=e, max_steps=c))
            ref.qf, ref.heaviside_beta = last["qf"], last["beta"]
            ref.set_design(last["x"])
            rr = ref.evaluate(warm=False)
            refs[name] = (rr, ref.d["dfds"].download())
            ref.close()
        rr, d64 = refs["cold_fp64"]
        rt, dt = refs["tight_fp64"]

        def errs(obj, dfds):
            return {"objective_rel_err": abs(obj - rt["objective"]) / abs(rt["objective"]),
                    "dfds_max_rel_err": float(np.max(np.abs(dfds - dt)) / np.max(np.abs(dt)))}

        later = hist[1:] or hist
        out[mode] = {
            "s_per_iteration": round(float(np.mean([r["seconds"] + r.get("train_seconds", 0.0) for r in later])), 4),
            "steps_per_iteration": int(np.mean([r["fwd_steps"] + r["adj_steps"] for r in later])),
            "total_s": round(total, 2), "iterations_run": len(hist), "objective_first": hist[0]["objective"],
            "objective_last": hist[-1]["objective"], "final_stage": hist[-1]["stage"],
            "fallbacks": int(sum(any(r["fallback"]) for r in hist)),
            "all_converged": bool(all(r["converged"] for r in hist)),
            "warm_gain_last": hist[-1].get("warm_gain"),
            "last_design_vs_cold_fp64": {
                "objective_rel_err": abs(last["objective"] - rr["objective"]) / abs(rr["objective"]),
                "dfds_max_rel_err": float(np.max(np.abs(last["dfds"] - d64)) / np.max(np.abs(d64)))},
            "last_design_vs_tight_fp64": errs(last["objective"], last["dfds"]),
            "cold_fp64_vs_tight_fp64": errs(rr["objective"], d64),
            "tight_fp64_steps": [rt["fwd_steps"], rt["adj_steps"]], "tight_fp64_converged": bool(rt["converged"])}
        print("bench: design loop %s %.3f s/iteration" % (mode, out[mode]["s_per_iteration"]), file=sys.stderr,
              flush=True)
    if "cold_fp64" in out:
        base = out["cold_fp64"]["s_per_iteration"]
        for mode in modes:
            out[mode]["speedup_vs_cold_fp64"] = round(base / out[mode]["s_per_iteration"], 2)
    out["note"] = ("seconds per iteration after the first, the warm-start network's training included; the "
                   "loops' designs drift apart slightly because each solve stops at the 1e-6 residual")
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
    if a.kernel_only in ("fwd3", "adj3"):
        from panslbm2_amd import D3Q7
        from panslbm2_amd.heatsink3d import Heatsink3D

        hs = Heatsink3D(ctx, a.lx3, a.ly3, a.lz3, ad_kind=D3Q7, max_steps=a.steps, use_graphs=False,
                        filter_radius=1.0)
        hs.init_forward()
        if a.kernel_only == "adj3":
            hs.fwd_buf = 0
            hs.init_adjoint()
            step = lambda t: hs._adj_step(t, 0)
            nbytes = BYTES_ADJ_Q15Q7 * hs.N
        else:
            step = hs._fwd_step
            nbytes = BYTES_FWD_Q15Q7 * hs.N
        for t in range(1, 3):
            step(t)
        ctx.sync()
        ctx.record(0)
        for t in range(3, 3 + a.steps):
            step(t)
        ctx.record(1)
        ms = ctx.elapsed_ms(0, 1) / a.steps
        hs.close()
        print(json.dumps({"kernel_only": a.kernel_only, "grid": [a.lx3, a.ly3, a.lz3], "launches": a.steps,
                          "us_per_launch": round(ms * 1e3, 3), "algorithmic_bytes": nbytes,
                          "achieved_GBs": round(nbytes / (ms * 1e-3) / 1e9, 1)}))
        return
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


def heatsink2d_extra(ctx, a):
    """The 2D heat sink at 200 x 400 (D2Q9 + D2Q5, FP64): evaluations, its fused forward step's
    roofline, the same kernel on 4096^2, FP32 / FP16 storage, the reference pairing D2Q9 + D2Q9 and
    the reference CPU path on it."""
    from panslbm2_amd import D2Q5, D2Q9, F16
    from panslbm2_amd.heatsink2d import Heatsink2D

    hs = Heatsink2D(ctx, lx=a.lx, ly=a.ly, ad_kind=D2Q5, max_steps=a.max_steps, use_graphs=not a.no_graphs)
    hs.evaluate()
    runs = [hs.evaluate() for _ in range(2)]
    r = runs[-1]
    secs = sum(x["seconds"] for x in runs)
    ups = sum(hs.N * (x["fwd_steps"] + x["adj_steps"]) for x in runs)
    ms_step = kernel_rate(hs)
    ach = BYTES_FWD_Q9Q5 * hs.N / (ms_step * 1e-3) / 1e9
    out = {"grid": [a.lx, a.ly], "lattices": "D2Q9 NS + D2Q5 AD", "precision": "fp64",
           "mlups": round(ups / secs / 1e6, 1), "s_per_evaluation": round(secs / len(runs), 4),
           "fwd_steps": r["fwd_steps"], "adj_steps": r["adj_steps"], "objective": r["objective"],
           "roofline": {"achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(ach / HBM_PEAK_GBS, 4), "us_per_launch": round(ms_step * 1e3, 3),
                        "traffic": TRAFFIC_FWD_200x400 if (a.lx, a.ly) == (200, 400) else None,
                        "note": "working set ~20 MB per step: resident in the 256 MB Infinity Cache"}}
    hs.close()
    if not a.no_big:
        ms_big, nbig = big_grid_roofline(ctx, use_graphs=not a.no_graphs)
        ach_big = BYTES_FWD_Q9Q5 * nbig / (ms_big * 1e-3) / 1e9
        out["roofline_hbm_4096x4096"] = {"achieved": round(ach_big, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                         "frac": round(ach_big / HBM_PEAK_GBS, 4),
                                         "us_per_launch": round(ms_big * 1e3, 1),
                                         "mlups": round(nbig / (ms_big * 1e-3) / 1e6, 1)}
    if not a.no_fp32:
        out["fp32"] = fp32_extra(ctx, a, r)
        out["fp16"] = fp32_extra(ctx, a, r, prec=F16)
    # like-for-like with the reference's lattice pairing (D2Q9 advection-diffusion)
    hs9 = Heatsink2D(ctx, lx=a.lx, ly=a.ly, ad_kind=D2Q9, max_steps=a.max_steps, use_graphs=not a.no_graphs)
    r9 = hs9.evaluate()
    out["mlups_d2q9_d2q9"] = round(hs9.N * (r9["fwd_steps"] + r9["adj_steps"]) / r9["seconds"] / 1e6, 1)
    if not a.no_cpu:
        out["cpu_baseline_d2q9_d2q9"] = cpu_baseline(hs9, target_seconds=8.0)
    hs9.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--lx3", type=int, default=128)
    ap.add_argument("--ly3", type=int, default=128)
    ap.add_argument("--lz3", type=int, default=64)
    ap.add_argument("--lx", type=int, default=200, help="2D grid (and the 2D --kernel-only modes)")
    ap.add_argument("--ly", type=int, default=400)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-big", action="store_true")
    ap.add_argument("--no-2d", action="store_true", help="skip the 2D heat-sink block")
    ap.add_argument("--max-steps", type=int, default=400000, help="cap on forward and on adjoint steps")
    ap.add_argument("--no-graphs", action="store_true", help="launch every step (for profilers)")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-smoothed", action="store_true", help="skip the c3_smoothed extra")
    ap.add_argument("--no-transient", action="store_true")
    ap.add_argument("--batch", type=int, default=64, help="designs in the batched evaluation (0: skip it)")
    ap.add_argument("--warm-iters", type=int, default=10, help="3D design-loop iterations per warm-start mode (0: skip)")
    ap.add_argument("--design-loop", type=int, default=0,
                    help="run only the 96x96x48 MMA design loop for this many iterations per mode and print "
                         "its JSON summary")
    ap.add_argument("--design-modes", default="cold_fp64,previous_solution_fp64,learned_fp64,learned_fp32",
                    help="comma-separated modes of --design-loop: " + ", ".join(DESIGN_LOOP_MODES))
    ap.add_argument("--stage-iterations", type=int, default=20, help="--design-loop iterations per continuation stage")
    ap.add_argument("--design-batch", type=int, default=0,
                    help="run only the batched 2D design loops: this many designs, plain vs FP32 + batched + warm")
    ap.add_argument("--design-batch-iters", type=int, default=8, help="--design-batch iterations")
    ap.add_argument("--design-batch-prec", choices=["F32", "F64"], default="F32",
                    help="--design-batch: storage of the batched form")
    ap.add_argument("--design-batch-cold", action="store_true", help="--design-batch: batched form without warm starts")
    ap.add_argument("--design-eps", type=float, default=1e-6, help="--design-loop steady-state residual")
    ap.add_argument("--kernel-only", choices=["fwd", "adj", "stream", "unet", "fwd3", "adj3"],
                    help="launch only the fused 2D forward / adjoint step (--lx x --ly), the pull-stream copy, "
                         "or the fused 3D forward / adjoint step (--lx3 x --ly3 x --lz3) --steps times "
                         "(profiling and counter calibration), or time the warm-start network's passes at 96x96x48")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (requested --gpus %d)" % a.gpus,
              file=sys.stderr)
        sys.exit(2)

    from panslbm2_amd import D3Q7, F16, F32, Context
    from panslbm2_amd.heatsink3d import Heatsink3D

    ctx = Context(0)
    if a.kernel_only:
        kernel_only(ctx, a)
        ctx.close()
        return
    if a.design_batch > 0:
        print(json.dumps({"design_batch": design_batch(ctx, a.design_batch, a.design_batch_iters,
                                                        fast_prec=a.design_batch_prec,
                                                        fast_warm=not a.design_batch_cold)}))
        ctx.close()
        return
    if a.design_loop > 0:
        modes = [m for m in a.design_modes.split(",") if m]
        for m in modes:
            if m not in DESIGN_LOOP_MODES:
                sys.exit("unknown --design-modes entry %r" % m)
        print(json.dumps({"design_loop": design_loop(ctx, a.design_loop, modes, a.stage_iterations,
                                                      eps=a.design_eps)}))
        ctx.close()
        return
    hs = Heatsink3D(ctx, a.lx3, a.ly3, a.lz3, ad_kind=D3Q7, max_steps=a.max_steps, use_graphs=not a.no_graphs)
    print("bench: 3D heatsink %dx%dx%d D3Q15+D3Q7 FP64, warmup %d, steps %d" % (a.lx3, a.ly3, a.lz3, a.warmup,
                                                                                  a.steps), file=sys.stderr, flush=True)
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
        print("step %d: %.3f s fwd %d adj %d obj %.9e" % (s, r["seconds"], r["fwd_steps"], r["adj_steps"],
                                                           r["objective"]), file=sys.stderr, flush=True)
    ctx.sync()
    wall = time.perf_counter() - t0
    mlups = total_updates / wall / 1e6
    dfds64 = hs.d["dfds"].download()

    ms_step = kernel_rate(hs)
    achieved = BYTES_FWD_Q15Q7 * hs.N / (ms_step * 1e-3) / 1e9
    big = (a.lx3, a.ly3, a.lz3) == (128, 128, 64)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": TRAFFIC_FWD_128x128x64 if big else None,
                "kernel": "k_collide_fused<D3Q15,D3Q7,ADBody> (pull + boundary program + collide in one pass), "
                          "per fused forward step",
                "bytes_per_launch": BYTES_FWD_Q15Q7 * hs.N, "us_per_launch": round(ms_step * 1e3, 3),
                "note": "algorithmic bytes: D3Q15 + D3Q7 populations read and written once (352 B) + alpha, "
                        "diffusivity, heat source (24 B) per node; working set ~0.8 GB per step (HBM)"}
    ms_adj = kernel_rate_adj(hs)
    ach_adj = BYTES_ADJ_Q15Q7 * hs.N / (ms_adj * 1e-3) / 1e9
    extra = {"c3_adjoint": {"kernel": "k_collide_fused<D3Q15,D3Q7,AADBody>, per fused adjoint step",
                            "bytes_per_launch": BYTES_ADJ_Q15Q7 * hs.N, "us_per_launch": round(ms_adj * 1e3, 3),
                            "achieved": round(ach_adj, 1), "frac": round(ach_adj / HBM_PEAK_GBS, 4),
                            "traffic": TRAFFIC_ADJ_128x128x64 if big else None}}
    cpu = None if a.no_cpu else cpu_baseline_3d(hs)
    r64 = runs[-1]
    hs.close()
    if not a.no_fp32:
        extra["c3_fp16"] = c3_precision_extra(ctx, a, r64, dfds64, F16)
        extra["c3_fp32"] = c3_precision_extra(ctx, a, r64, dfds64, F32)
    if not a.no_smoothed:
        extra["c3_smoothed"] = c3_smoothed_extra(ctx, a, r64)
    if not a.no_2d:
        extra["heatsink2d"] = heatsink2d_extra(ctx, a)
    if a.batch > 0:
        extra["batch"] = batch_extra(ctx, a.batch)
    if a.warm_iters > 0:
        extra["warm_start"] = design_loop(ctx, a.warm_iters, [m for m in a.design_modes.split(",") if m],
                                          a.stage_iterations, eps=a.design_eps)
    if not a.no_transient:
        extra["transient"] = transient_extra(ctx)

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
        "config": {"workload": "3D heatsink design evaluation, 128x128x64 D3Q15 NS + D3Q7 AD (forward + adjoint "
                               "to 1e-6, sensitivity), cold start",
                   "grid": [a.lx3, a.ly3, a.lz3], "lattices": "D3Q15 NS + D3Q7 AD", "precision": "fp64",
                   "fwd_steps": r64["fwd_steps"], "adj_steps": r64["adj_steps"],
                   "wall_s_to_steady_cold": round(r64["seconds"], 3), "objective": r64["objective"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
