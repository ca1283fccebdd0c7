"""Benchmark: Voronoi cells per second on one MI355X.

Workload (N=1): 2,000,000 seeds in [0,1]^3, 1M uniform + 1M from an
8-component isotropic Gaussian mixture (means U[0.2,0.8]^3, sigma U[0.01,0.1],
rejection to the cube), seed 2.  One step = one full pass of the cell path on
seeds already resident in HBM: Morton sort, tree, clipping with the security
radius (all tiers; the first pass gathers each cell's neighbours by its own
tree walk, so there is no separate kNN pass at k0 = 0), FP64 volume /
centroid / facet-area integration and the neighbour CSR.  Steps are timed with torch.cuda.synchronize() on both
sides.

The JSON line also carries:
  roofline      the first clipping pass (one launch over every cell, the
                dominant kernel; FP32 corner state, FP64 integration):
                algorithmic FLOPs of that launch (sum over cells of
                7 V + 66 C + 34 T + 20 F: 7 flops per corner-plane test, 66
                per corner from three planes, 34 per integration tetrahedron,
                20 per facet), divided by its duration from HIP events
                recorded on the library's launch stream, against MI355X's
                vector peak of 78.6 T operations/s (FP64, and FP32 without
                packed instructions, which the build excludes);
  fp32_pass     cells the FP32 first pass left to the FP64 first pass, that
                launch's time, and the FP64 first pass timed alone
                (VGC_FP64_FIRST) on the same seeds for comparison;
                traffic (HBM bytes per launch) and valu_busy_frac (share of
                the VALU issue slots in use, the kernel's actual bound) are
                read from the committed rocprofv3 counter summary
                profiles/k_clip_counters_2M_mixture.md (default workload only);
  cpu_baseline  the oracle (a plain-C restatement of geogram's VBW cell +
                security-radius loop with exact predicates) on the host cores,
                on a bounded sample of the same workload;
  ot            semi-discrete transport, 1M uniform seeds, equal masses, to
                1e-4 max relative mass error from w = 0 (every transport leg:
                damped Newton, KMT step rule, CG tolerance from the mass error
                once the steps are full, VGC_OT_INEXACT);
  ot_warm_start 1M 16-Gaussian seeds with log-normal masses: plain start
                against the learned MLP start (features + forward timed;
                newton_cut, time_cut) and against the coarse-plan start
                (every 256th seed solved to 1e-3, extended by a convex soft
                maximum; coarse solve and start timed; coarse_newton_cut,
                coarse_time_cut), and optimize_levels with each level started
                from the convex extension of the previous one (levels_convex);
  batched       256 independent instances (N log-uniform in [2k, 200k]):
                20 batched Lloyd steps and batched transport to 1e-4.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
COUNTERS = os.path.join(ROOT, "profiles", "k_clip_counters_2M_mixture.md")


def profile_figure(key):
    """A `key: value` line of the rocprofv3 counter summary of the clip
    kernel (HBM bytes per launch from FETCH_SIZE + WRITE_SIZE scaled by a
    known-size copy, VALU issue fraction from SQ counters) — None unless the
    summary records the hash of the kernel sources this build compiles, so a
    stale profile is never reported."""
    from geogram_amd.profiling import read_figure
    return read_figure(COUNTERS, key)


def cell_errors(np, g, o, rows):
    """Per-output error of GPU cells g (host arrays: vol, mg, ptr, nbr,
    farea, caller order) against a CPU result o (vol, mg, cnt, nbr, farea;
    row t is cell rows[t]): max relative volume error, max centroid error
    over the cell size, max relative facet-area error (facets above 1e-6 of
    the cell's face scale), and the count of neighbour sets that differ."""
    vo = o["vol"]
    vg = g["vol"][rows]
    nz = vo > 0
    vrel = np.abs(vg[nz] - vo[nz]) / vo[nz]
    cg = g["mg"][rows][nz] / vg[nz, None]
    co = o["mg"][nz] / vo[nz, None]
    cen = np.abs(cg - co).max(axis=1) / np.cbrt(vo[nz])
    # neighbour sets and facet areas, row by row in sorted order
    gc = (g["ptr"][rows + 1] - g["ptr"][rows])
    oc = o["cnt"].astype(np.int64)
    same_cnt = gc == oc
    if len(rows) == len(g["vol"]):
        # the whole workload: the rows are every cell in order
        gn, ga_ = g["nbr"], g["farea"]
    else:
        first = np.cumsum(gc) - gc
        idx = np.repeat(g["ptr"][rows], gc) + (np.arange(int(gc.sum())) - np.repeat(first, gc))
        gn, ga_ = g["nbr"][idx], g["farea"][idx]
    grow = np.repeat(np.arange(len(rows)), gc)
    orow = np.repeat(np.arange(len(rows)), oc)
    mask = np.arange(o["nbr"].shape[1])[None, :] < oc[:, None]
    on, oa = o["nbr"][mask], o["farea"][mask]
    gk = np.lexsort((gn, grow))
    ok = np.lexsort((on, orow))
    if len(gk) == len(ok):
        rowdiff = (gn[gk] != on[ok]) | (grow[gk] != orow[ok])
        bad_rows = np.unique(grow[gk][rowdiff])
        mism = int(len(bad_rows) + np.count_nonzero(~same_cnt & ~np.isin(np.arange(len(rows)),
                                                                           bad_rows)))
        good = ~np.isin(grow[gk], bad_rows) & same_cnt[grow[gk]]
        scale = vo[grow[gk]] ** (2.0 / 3.0)
        fa_o, fa_g = oa[ok], ga_[gk]
        big = good & (fa_o > 1e-6 * scale)
        frel = float((np.abs(fa_g - fa_o)[big] / fa_o[big]).max()) if big.any() else 0.0
    else:
        mism = int(np.count_nonzero(~same_cnt))
        frel = None
    return {"cells": int(len(rows)), "vol_rel_max": float(vrel.max()),
            "centroid_max_over_cell_size": float(cen.max()), "facet_area_rel_max": frel,
            "neighbour_sets_differing": mism}


def reference_mass_error(np, x, w, nu):
    """max_i |m_i - nu_i| / nu_i with the masses m of geogram's own Laguerre
    cells of (x, w) (the reference build, oracle/_ref), or None."""
    import oracle
    if not oracle.ref_available():
        return None
    m = oracle.ref_cells(x, w, threads=min(16, os.cpu_count() or 1))["vol"]
    return float(np.max(np.abs(m - nu) / nu))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=2_000_000)
    ap.add_argument("--k0", type=int, default=0,
                    help="0: neighbours gathered in the clipping kernel; 16..64: kNN lists")
    ap.add_argument("--ot-n", type=int, default=1_000_000,
                    help="seeds of the semi-discrete OT leg (0: skip)")
    ap.add_argument("--ot4-n", type=int, default=1_000_000,
                    help="seeds of the learned-warm-start OT leg (0: skip)")
    ap.add_argument("--c5-instances", type=int, default=256,
                    help="batched leg: independent instances, N log-uniform in [2k, 200k] (0: skip)")
    ap.add_argument("--c5-ot", type=int, default=1, help="batched leg: also the transport solves")
    ap.add_argument("--copy-calib", action="store_true",
                    help="one 256 MiB device-to-device copy before the steps, the known-size "
                         "kernel that scales HBM counters under rocprofv3 --pmc")
    ap.add_argument("--surface-level", type=int, default=7,
                    help="surface leg: icosphere subdivision level (7: 327680 triangles; 0: skip)")
    ap.add_argument("--surface-seeds", type=int, default=100_000)
    ap.add_argument("--periodic-n", type=int, default=500_000,
                    help="periodic leg: uniform seeds on the unit 3-torus (0: skip)")
    ap.add_argument("--newton-iters", type=int, default=10,
                    help="quasi-Newton CVT leg on the 100k CVT workload (0: skip)")
    ap.add_argument("--mesh-n", type=int, default=200_000,
                    help="cell-mesh / RDT export leg (0: skip)")
    ap.add_argument("--cpu-sample", type=int, default=400_000,
                    help="cells of the workload timed on the host with the oracle (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build targets one GPU; --gpus {args.gpus} is not supported "
              "(the path would shard by Morton ranges with a seed halo, see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import geogram_amd as ga

    n = args.n
    x = ga.sample_points(1, 2, n)
    xd = torch.from_numpy(x).cuda()
    ctx = ga.Context(0)

    # algorithmic work of the first clipping pass (deterministic for the
    # workload), counted once outside the timed steps
    work = ctx.cells(xd, k0=args.k0, count_work=True).stats
    if args.copy_calib:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    for _ in range(args.warmup):
        r = ctx.cells(xd, k0=args.k0)
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = ctx.cells(xd, k0=args.k0)
        stats.append(r.stats)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = dt * 1e3 / args.steps
    value = n * args.steps / dt

    sumvol_err = abs(float(r.vol.sum().item()) - 1.0)
    s = stats[-1]
    w = work
    flops_a = 7 * w["a_live"] + 66 * w["a_new"] + 34 * w["a_tets"] + 20 * w["a_facets"]
    clip_ms = float(np.mean([t["t_clip_a_gpu_ms"] for t in stats]))
    achieved = flops_a / (clip_ms * 1e-3) / 1e12
    roofline = {
        "kernel": "k_clip<1,1,false,true,false,false,true> (first clipping pass, every cell, "
                  "FP32 corner state)",
        "bound": "valu",
        "achieved": round(achieved, 4),
        "peak": FP64_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved / FP64_PEAK_TFLOPS, 5),
        "traffic": profile_figure("traffic_bytes_per_launch") if args.n == 2_000_000 and args.k0 == 0 else None,
        "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated; "
                        "from profiles/k_clip_counters_2M_mixture.md when it was taken of "
                        "these kernel sources, else null)",
        "valu_busy_frac": profile_figure("valu_busy_frac") if args.n == 2_000_000 and args.k0 == 0 else None,
        "alg_flops_per_launch": int(flops_a),
        "flops_per_cell": round(flops_a / n, 1),
        "launch_ms_hip_events": round(clip_ms, 3),
    }

    # the FP64 first pass alone on the same seeds (identical outputs)
    fp64_ms = []
    for _ in range(3):
        r64 = ctx.cells(xd, k0=args.k0, fp64_first=True)
        fp64_ms.append(r64.stats["t_clip_a_gpu_ms"])
    same64 = bool(torch.equal(r64.vol, r.vol) and torch.equal(r64.nbr, r.nbr)
                  and torch.equal(r64.farea, r.farea))
    fp32_pass = {"redo_cells": int(s["nb_fp64_redo"]),
                 "redo_gpu_ms": round(float(np.mean([t["t_redo_gpu_ms"] for t in stats])), 3),
                 "fp32_first_pass_gpu_ms": round(clip_ms, 3),
                 "fp64_first_pass_gpu_ms": round(float(min(fp64_ms)), 3),
                 "outputs_identical_to_fp64_pass": same64}
    del r64

    # semi-discrete OT leg: 1M uniform seeds, equal masses, to 1e-4 max
    # relative mass error, plain start (w = 0)
    ot = None
    if args.ot_n > 0:
        m = args.ot_n
        xo = torch.from_numpy(ga.sample_points(0, 3, m)).cuda()
        nu = torch.full((m,), 1.0 / m, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        wo, ost = ctx.ot_solve(xo, nu, tol=1e-4, verbose=True, inexact=True)
        torch.cuda.synchronize()
        tot = time.perf_counter() - t2
        ot = {"n": m, "start": "plain (w = 0)", "time_s": round(tot, 3),
              "newton_iters": int(ost["newton_iters"]), "cg_iters": int(ost["cg_iters"]),
              "cell_passes": int(ost["cell_passes"]), "err0": ost["err0"],
              "final_err": ost["final_err"], "converged": bool(ost["converged"]),
              "t_cells_ms": round(ost["t_cells_ms"], 1), "t_cg_ms": round(ost["t_cg_ms"], 1),
              "mass_err_reference": reference_mass_error(np, xo.cpu().numpy(), wo.cpu().numpy(),
                                                         nu.cpu().numpy())}

    # learned warm start: 1M seeds of a 16-Gaussian mixture with log-normal
    # masses (seed 4), plain start against the MLP start (features + forward
    # included in its time), both to 1e-4 max relative mass error
    ws = None
    if args.ot4_n > 0:
        from geogram_amd import warmstart

        m = args.ot4_n
        x4, nu4 = warmstart.make_instance(4, m)
        x4d, nu4d = torch.from_numpy(x4).cuda(), torch.from_numpy(nu4).cuda()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        w4, sp = ctx.ot_solve(x4d, nu4d, tol=1e-4, verbose=True, inexact=True)
        torch.cuda.synchronize()
        tp = time.perf_counter() - t2
        ws = {"n": m, "plain": {"time_s": round(tp, 3), "newton_iters": int(sp["newton_iters"]),
                                "cell_passes": int(sp["cell_passes"]),
                                "converged": bool(sp["converged"]), "err0": sp["err0"],
                                "mass_err_reference": reference_mass_error(np, x4, w4.cpu().numpy(),
                                                                           nu4)}}
        if os.path.exists(warmstart.PARAMS_FILE):
            params = torch.from_numpy(warmstart.load_params()).cuda()
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            w0 = warmstart.predict_weights(ctx, x4d, nu4d, params=params)
            _, sl = ctx.ot_solve(x4d, nu4d, w0=w0, tol=1e-4, judge_start=True, verbose=True,
                                 inexact=True)
            torch.cuda.synchronize()
            tl = time.perf_counter() - t2
            ws["learned"] = {"time_s": round(tl, 3), "newton_iters": int(sl["newton_iters"]),
                             "cell_passes": int(sl["cell_passes"]),
                             "converged": bool(sl["converged"]), "err0": sl["err0"],
                             "fell_back_to_zero": bool(sl["used_fallback_start"])}
            ws["newton_cut"] = round(1.0 - sl["newton_iters"] / max(1, sp["newton_iters"]), 3)
            ws["time_cut"] = round(1.0 - tl / tp, 3)
        else:
            ws["learned"] = None
        # start from a coarse plan (every 256th seed solved to 1e-3, the fine
        # start its convex soft-max extension; coarse solve and start timed)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        w0c, cinfo = warmstart.coarse_start(ctx, x4d, nu4d)
        _, sc = ctx.ot_solve(x4d, nu4d, w0=w0c, tol=1e-4, judge_start=True, verbose=True,
                             inexact=True)
        torch.cuda.synchronize()
        tc = time.perf_counter() - t2
        ws["coarse"] = {"time_s": round(tc, 3), "newton_iters": int(sc["newton_iters"]),
                        "cell_passes": int(sc["cell_passes"]), "converged": bool(sc["converged"]),
                        "err0": sc["err0"], "fell_back_to_zero": bool(sc["used_fallback_start"]),
                        **{k: v for k, v in cinfo.items() if k != "mu_fine"}}
        ws["coarse_newton_cut"] = round(1.0 - sc["newton_iters"] / max(1, sp["newton_iters"]), 3)
        ws["coarse_time_cut"] = round(1.0 - tc / tp, 3)
        # optimize_levels over BRIO-style levels (each 1/8 of the next, the
        # coarsest at least 8000 seeds), each level from the convex extension
        # of the previous plan (VGC_OT_CONVEX_LEVELS); the seeds permuted
        # into level order before the timed region
        perm, lv = ga.brio_levels(m, min_level=8000, seed=3)
        pt = torch.from_numpy(np.asarray(perm, dtype=np.int64)).cuda()
        x4p, nu4p = x4d[pt].contiguous(), nu4d[pt].contiguous()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        _, sv = ctx.ot_solve_levels(x4p, nu4p, lv, tol=1e-4, inexact=True, convex=True)
        torch.cuda.synchronize()
        tv = time.perf_counter() - t2
        ws["levels_convex"] = {"time_s": round(tv, 3), "levels": [int(v) for v in lv],
                               "newton_iters": [int(st["newton_iters"]) for st in sv],
                               "cell_passes": [int(st["cell_passes"]) for st in sv],
                               "converged": all(bool(st["converged"]) for st in sv),
                               "fell_back_to_zero": [bool(st["used_fallback_start"]) for st in sv]}

    # batched uneven instances: Lloyd (20 steps) and transport to 1e-4 for
    # every instance at once; even instances uniform, odd ones 16-Gaussian;
    # masses equal for k % 4 < 2, log-normal otherwise
    c5 = None
    if args.c5_instances > 0:
        K = args.c5_instances
        rng = np.random.default_rng(5)
        sizes = np.exp(rng.uniform(np.log(2000), np.log(200000), K)).astype(np.int64)
        xs, nus = [], []
        for k, sz in enumerate(sizes):
            xs.append(ga.sample_points(0 if k % 2 == 0 else 2, 5000 + k, int(sz)))
            if k % 4 < 2:
                nuk = np.full(int(sz), 1.0 / sz)
            else:
                nuk = np.exp(np.random.default_rng(k).normal(0.0, 0.5, int(sz)))
                nuk /= nuk.sum()
            nus.append(nuk)
        x5 = torch.from_numpy(np.vstack(xs)).cuda()
        nu5 = torch.from_numpy(np.concatenate(nus)).cuda()
        boxes5 = np.array([[0, 0, 0, 1, 1, 1]] * K, dtype=float)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        _, l5 = ctx.lloyd(x5, 20, inst_n=sizes, boxes=boxes5, autotune=True)
        torch.cuda.synchronize()
        tl5 = time.perf_counter() - t2
        c5 = {"instances": K, "seeds": int(sizes.sum()),
              "lloyd": {"steps": 20, "time_s": round(tl5, 3),
                        "cells_per_s": round(20 * float(sizes.sum()) / tl5, 1),
                        "failed": int(l5["nb_failed"]), "tuned_waves": int(l5["tuned_waves"])}}
        if args.c5_ot:
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            _, s5 = ctx.ot_solve_batch(x5, nu5, sizes, boxes5, tol=1e-4, verbose=True, inexact=True,
                                       autotune=True)
            torch.cuda.synchronize()
            to5 = time.perf_counter() - t2
            its = [s["newton_iters"] for s in s5]
            c5["ot"] = {"time_s": round(to5, 3), "converged": int(sum(s["converged"] for s in s5)),
                        "newton_max": int(max(its)), "newton_mean": round(float(np.mean(its)), 2),
                        "cell_passes": int(s5[0]["cell_passes"]),
                        "t_cells_ms": round(s5[0]["t_cells_ms"], 1), "t_cg_ms": round(s5[0]["t_cg_ms"], 1)}
            # the same batch from each instance's coarse-plan start (every
            # 128th seed, at least 64, solved together to 1e-3; timed)
            from geogram_amd import warmstart

            torch.cuda.synchronize()
            t2 = time.perf_counter()
            w05, ci5 = warmstart.coarse_start_batch(ctx, x5, nu5, sizes, boxes5)
            _, s5c = ctx.ot_solve_batch(x5, nu5, sizes, boxes5, w0=w05, tol=1e-4, verbose=True,
                                        inexact=True, autotune=True)
            torch.cuda.synchronize()
            tc5 = time.perf_counter() - t2
            itc = [s["newton_iters"] for s in s5c]
            c5["ot_coarse_start"] = {
                "time_s": round(tc5, 3), "converged": int(sum(s["converged"] for s in s5c)),
                "newton_max": int(max(itc)), "newton_mean": round(float(np.mean(itc)), 2),
                "cell_passes": int(s5c[0]["cell_passes"]),
                "restarted_from_zero": int(sum(s["used_fallback_start"] for s in s5c)), **ci5}

    # the rows beside the cell pass, each timed on the GPU with its error
    # against the oracle (or the reference build) on the same input
    import oracle

    def gpu_time(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            t2 = time.perf_counter()
            res = fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t2)
        return res, float(np.median(ts))

    surf = None
    if args.surface_level > 0:
        from geogram_amd.configs import icosphere, sphere_seeds

        V, F = icosphere(args.surface_level)
        xs_ = sphere_seeds(args.surface_seeds, 13)
        Vd, Fd, xsd = (torch.from_numpy(V).cuda(), torch.from_numpy(F).cuda(),
                       torch.from_numpy(xs_).cuda())
        (g, mm, sst), ts = gpu_time(lambda: ctx.surface_centroids(xsd, Vd, Fd))
        t2 = time.perf_counter()
        mgo, mo, npoly = oracle.surface_centroids(xs_, V, F)
        to_ = time.perf_counter() - t2
        gh, mh = g.cpu().numpy(), mm.cpu().numpy()
        surf = {"triangles": int(len(F)), "seeds": int(args.surface_seeds),
                "gpu_ms": round(ts * 1e3, 3), "triangles_per_s": round(len(F) / ts, 1),
                "polygons": int(npoly), "items": int(sst["nb_items"]),
                "splits": int(sst["nb_split"]), "redo": int(sst["nb_redo"]),
                "cpu_oracle_s": round(to_, 3), "cpu_oracle_cores": 1,
                "max_rel_mass_err": float(np.abs(mh - mo).max() / mo.max()),
                "max_rel_moment_err": float(np.abs(gh - mgo).max() / np.abs(mgo).max())}

    per = None
    if args.periodic_n > 0:
        xp = ga.sample_points(0, 14, args.periodic_n)
        xpd = torch.from_numpy(xp).cuda()
        rp, tp_ = gpu_time(lambda: ctx.cells(xpd, periodic=True))
        per = {"n": args.periodic_n, "gpu_ms": round(tp_ * 1e3, 3),
               "cells_per_s": round(args.periodic_n / tp_, 1),
               "retries": int(rp.stats["nb_periodic_retries"]),
               "sum_vol_minus_1": abs(float(rp.vol.sum().item()) - 1.0)}
        if oracle.ref_available():
            threads = min(16, os.cpu_count() or 1)
            refp = oracle.ref_cells(xp, threads=threads, periodic=True)
            tref = refp["t_delaunay_s"] + refp["t_cells_s"]
            vg = rp.vol.cpu().numpy()
            per["cpu_reference"] = {"cells_per_s": round(args.periodic_n / tref, 1),
                                    "cores": threads,
                                    "what": "PeriodicDelaunay3d(periodic) + Laguerre cells"}
            per["max_rel_vol_err_vs_reference"] = float(
                (np.abs(vg - refp["vol"]) / refp["vol"]).max())

    newton = None
    if args.newton_iters > 0:
        from geogram_amd.configs import cvt_points

        xc = torch.from_numpy(cvt_points()).cuda()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        _, nst = ctx.cvt_newton(xc, args.newton_iters)
        torch.cuda.synchronize()
        tn = time.perf_counter() - t2
        newton = {"n": int(xc.shape[0]), "iterations": int(nst["iterations"]),
                  "evaluations": int(nst["evaluations"]), "time_s": round(tn, 3),
                  "f0": nst["f0"], "f": nst["f"], "gnorm0": nst["gnorm0"], "gnorm": nst["gnorm"]}

    meshleg = None
    if args.mesh_n > 0:
        xm = ga.sample_points(0, 15, args.mesh_n)
        (tets), tr = gpu_time(lambda: ctx.rdt(xm), reps=2)
        (cm), tc_ = gpu_time(lambda: ctx.cell_mesh(xm), reps=2)
        meshleg = {"n": args.mesh_n, "rdt_ms": round(tr * 1e3, 3), "tets": int(len(tets)),
                   "cell_mesh_ms": round(tc_ * 1e3, 3),
                   "note": "host arrays in and out (PCIe included)"}

    # CPU baseline and per-output errors.  The reference itself (geogram's
    # Delaunay + ConvexCell path, oracle/_ref) runs the whole workload on the
    # host cores when it is built; the oracle (plain-C restatement) runs a
    # sample otherwise.  Errors of the GPU cells: against the reference on
    # every cell (planes in absolute coordinates, as geogram forms them), and
    # against the oracle in the seed-centred frame on the sample.
    cpu, errors = None, {}
    if args.cpu_sample > 0:
        rh = {"vol": r.vol.cpu().numpy(), "mg": r.mg.cpu().numpy(),
              "ptr": r.nbr_ptr.cpu().numpy().astype(np.int64), "nbr": r.nbr.cpu().numpy(),
              "farea": r.farea.cpu().numpy()}
        threads = min(16, os.cpu_count() or 1)
        if oracle.ref_available():
            ref = oracle.ref_cells(x, threads=threads)
            tc = ref["t_delaunay_s"] + ref["t_cells_s"]
            cpu = {"value": round(n / tc, 1), "unit": "cells/s", "cores": threads,
                   "kind": "reference",
                   "sample": f"the whole workload ({n} cells) through geogram's fastest CPU path "
                             f"(PeriodicDelaunay3d + VBW::ConvexCell + 6 box clips, compute_RVD "
                             f"cube mode; {ref['t_delaunay_s']:.2f} s Delaunay + "
                             f"{ref['t_cells_s']:.2f} s cells on {threads} threads)"}
            errors["vs_reference"] = cell_errors(np, rh, ref, np.arange(n))
        sample = np.random.default_rng(0).choice(n, min(args.cpu_sample, n), replace=False)
        sample.sort()
        oracle.set_threads(threads)
        oracle.cells(x[:1000], k0=20)  # load
        t1 = time.perf_counter()
        oabs = oracle.cells(x, k0=20, subset=sample.astype(np.int32))
        to = time.perf_counter() - t1
        port = {"value": round(len(sample) / to, 1), "unit": "cells/s", "cores": threads,
                "kind": "port",
                "sample": f"{len(sample)} random cells of the same 2M workload, kd-tree over all "
                          f"2M seeds included ({to:.2f} s)"}
        if cpu is None:
            cpu = port
        else:
            cpu["port"] = port
        ocen = oracle.cells(x, k0=20, subset=sample.astype(np.int32), centred=True)
        errors["vs_oracle_absolute_frame"] = cell_errors(np, rh, oabs, sample)
        errors["vs_oracle_centred_frame"] = cell_errors(np, rh, ocen, sample)

    out = {
        "metric": "Voronoi cells/sec; 1M-seed OT time to 1e-4 mass err (learned warm start vs plain)",
        "value": round(value, 1),
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "2M seeds in [0,1]^3 (1M uniform + 1M 8-Gaussian mixture, seed 2), "
                        "one pass of Voronoi cells: volume, centroid, neighbours, facet areas",
            "n": n,
            "k0": args.k0,
        },
        "roofline": roofline,
        "fp32_pass": fp32_pass,
        "cpu_baseline": cpu,
        "errors": errors or None,
        "ot": ot,
        "ot_warm_start": ws,
        "batched": c5,
        "surface": surf,
        "periodic": per,
        "cvt_newton": newton,
        "mesh_export": meshleg,
        "checks": {"sum_vol_minus_1": sumvol_err, "failed": int(s["nb_failed"]),
                   "uncertain": int(s["nb_uncertain"])},
        "phases_ms": {k: round(s[k], 3) for k in ("t_build_ms", "t_knn_ms", "t_clip_ms",
                                                  "t_pass_bc_ms", "t_compact_ms", "t_total_ms",
                                                  "t_clip_a_gpu_ms", "t_knn_gpu_ms")},
        "tiers": {k: int(s[k]) for k in ("nb_pass_b", "nb_pass_c", "nb_pass_e", "radius_entries")},
        "work": {k: int(w[k]) for k in ("nb_tests", "nb_clips", "a_tests", "a_live", "a_new",
                                        "a_tets", "a_facets")},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
