#!/usr/bin/env python3
"""Throughput of the full SPH step (neighbour rebuild + both pair passes +
symplectic update + CFL Δt) on one MI355X.

Workload: the large 2D dam break — a 4.0 m x 4.0 m water column at dx = 0.002
(2000 x 2000 = 4,000,000 fluid particles) in a 16 m x 8 m tank with 3 wall
layers (47,991 particles), hydrostatic initial density, dist = H (a rebuild
every step), adaptive CFL = 0.2 Δt.  A step is one `stepsolve!` iteration over
the whole scene with all state resident in HBM.

The headline `value` is the FP32-split path (FP32 pair terms summed in FP64,
density as ρ-ρ0, FP64 positions and Δt); the FP64 path runs on
the same GPU in the same invocation and is reported beside it.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpucelllistsph.jl_amd"))

import gclsph as G  # noqa: E402
from gclsph import scenes as S  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# MI355X VALU issue peak: 256 CUs x 4 SIMDs, a wave64 instruction every 4
# cycles per SIMD, at the 2.4 GHz peak engine clock (1 wave64 instruction per
# CU per cycle; the FP32 vector peak without packed math, 78.6 TFLOP/s in FMA)
VALU_PEAK_GIPS = 256 * 2.4
# The first call of more than 15 steps on a problem or batch times the pass
# variants on 15 steps (4 code-generation variants at 3 block sizes, the
# occupancy cap, the two fastest again) after up to 16 untimed ones (a fresh
# device warms up over its first dozen steps); a warmup call of 32 steps
# keeps that tuning, fully warmed, out of the timed call.
TUNE_STEPS = 32


def log(msg: str) -> None:
    """Progress of the legs on stderr (stdout carries only the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(dim: int, split: bool) -> dict:
    """Compulsory HBM bytes per particle for each stage (canonical field sizes:
    positions 8·D, velocity 8·D or 4·D, density 8 or 4, ptype 4, cell key 4,
    permutation 4); neighbour reuse through caches counts zero."""
    x = 8 * dim
    v = (4 if split else 8) * dim
    r = 4 if split else 8
    pt = 4
    return {
        "rebuild": 4 + x + 4 + 4 + 2 * (x + v + r + pt),  # bin: read x, write key; reorder: key+perm, state r/w
        "pass1": (x + v + r + pt) + (x + v + r),  # read x, v, ρ, ptype; write xΔt½, vΔt½, ρΔt½
        "pass2": (x + x + v + r + pt + v + r) + (x + v + r),  # read xⁿ, xΔt½, vΔt½, ρΔt½, ptype, vⁿ, ρⁿ; write x, v, ρ
    }


def run_gpu(scene, precision: str, steps: int, warmup: int, stage_split: bool = True):
    """Times `steps` steps of one stepsolve! call as production runs it: the
    pass variants already tuned and the step replayed as a captured graph
    (the warmup call tunes, a 2-step call captures).  A second call of the
    same length with per-stage HIP events (eager launches) gives the split."""
    cl = G.GPUCellList(scene.points, scene.cellsize, scene.dist)
    prob = G.SPHProblem.from_params(cl, scene.params, scene.rho, scene.ptype, precision=precision)
    tl = G.default_timelims(prob)
    w = max(warmup, TUNE_STEPS)
    G.stepsolve_(prob, w, timestepping=True, timelims=tl)
    G.stepsolve_(prob, 2, timestepping=True, timelims=tl)
    t0 = time.perf_counter()
    upd, _ = G.stepsolve_(prob, steps, timestepping=True, timelims=tl)
    t1 = time.perf_counter()
    out = {"seconds": t1 - t0, "updaten": upd, "dt": prob.dt, "warmup_steps": w + 2,
           "pass_variants": prob.pass_variants()}
    if stage_split:
        prob.set_timing(True)
        t2 = time.perf_counter()
        G.stepsolve_(prob, steps, timestepping=True, timelims=tl)
        out["timed_split_seconds"] = time.perf_counter() - t2
        out["timing"] = prob.timing()
        prob.set_timing(False)
    return out


def accuracy(scene, steps: int = 1000) -> dict:
    """FP32-split against FP64 after `steps` adaptive steps from the same
    initial state, per field (max and 99.99 % quantile), beside the FP64
    path's own sensitivity: FP64 started from positions perturbed by ~1 ulp."""
    def run(prec, pert=0.0):
        pts = scene.points
        if pert:
            pts = tuple(p * (1 + pert * (S.uniform01(5, len(p), k) - 0.5)) for k, p in enumerate(pts))
        cl = G.GPUCellList(pts, scene.cellsize, scene.dist)
        prob = G.SPHProblem.from_params(cl, scene.params, scene.rho, scene.ptype, precision=prec)
        G.stepsolve_(prob, steps, timestepping=True)
        out = (np.array(G.get_points(prob)), np.array(G.get_velocity(prob)), G.get_density(prob))
        del prob, cl
        return out

    x64, v64, r64 = run("f64")
    xp, vp, rp = run("f64", 4e-16)
    x32, v32, r32 = run("f32split")
    dx = scene.params.dx
    vmax = float(np.abs(v64).max())

    def stats(err, scale):
        return {"max": float(err.max() / scale), "p99.99": float(np.quantile(err, 0.9999) / scale)}

    ex = np.abs(x32 - x64).max(axis=0)
    ev = np.abs(v32 - v64).max(axis=0)
    er = np.abs(r32 / r64 - 1)
    cx = np.abs(xp - x64).max(axis=0)
    cv = np.abs(vp - v64).max(axis=0)
    cr = np.abs(rp / r64 - 1)
    return {
        "steps": steps,
        "f32split_vs_f64": {"x_over_dx": stats(ex, dx), "v_over_vmax": stats(ev, vmax), "rho_rel": stats(er, 1.0),
                            "particles_x_beyond_1e-4dx": int((ex > 1e-4 * dx).sum())},
        "f64_vs_f64_1ulp_perturbed": {"x_over_dx": stats(cx, dx), "v_over_vmax": stats(cv, vmax),
                                      "rho_rel": stats(cr, 1.0),
                                      "particles_x_beyond_1e-4dx": int((cx > 1e-4 * dx).sum())},
        "vmax": vmax,
        "targets": {"rho_rel_max_1e-5": bool(er.max() < 1e-5),
                    "x_max_1e-4dx": bool(ex.max() < 1e-4 * dx),
                    "x_max_note": "the maximum position bound is below FP64's own 1-ulp sensitivity on this "
                                  "scene (chaotic floor-adjacent row); FP32-split is compared with that floor"},
    }


def run_ensemble(steps: int, warmup: int):
    """The batched ensemble: 64 seeded 2D dam breaks of uneven size (FP32-split),
    all stepped by one batch (one launch per stage, a captured step graph),
    against the same problems stepped one by one."""
    scs = S.config4()
    probs = []
    for sc in scs:
        cl = G.GPUCellList(sc.points, sc.cellsize, sc.dist)
        probs.append(G.SPHProblem.from_params(cl, sc.params, sc.rho, sc.ptype, precision="f32split"))
    total = sum(sc.n for sc in scs)
    batch = G.SPHBatch(probs)
    w = max(warmup, TUNE_STEPS)
    G.batch_stepsolve_(batch, w, timestepping=True)
    G.batch_stepsolve_(batch, 2, timestepping=True)
    t0 = time.perf_counter()
    G.batch_stepsolve_(batch, steps, timestepping=True)
    tb = time.perf_counter() - t0
    for p in probs:
        G.stepsolve_(p, w, timestepping=True)
        G.stepsolve_(p, 2, timestepping=True)
    t0 = time.perf_counter()
    for p in probs:
        G.stepsolve_(p, steps, timestepping=True)
    t1 = time.perf_counter() - t0
    tuned = batch.pass_variants()
    del batch
    return {"scenes": len(scs), "pass_variants": tuned, "particles": total, "min_particles": min(sc.n for sc in scs),
            "max_particles": max(sc.n for sc in scs), "steps": steps,
            "batched_particle_steps_per_s": total * steps / tb, "batched_ms_per_step": tb / steps * 1e3,
            "one_by_one_particle_steps_per_s": total * steps / t1, "one_by_one_ms_per_step": t1 / steps * 1e3,
            "speedup": t1 / tb,
            "workload": "64 2D dam breaks, dx in {0.005,0.01,0.02}, 0.5-1.5 x 1.0-2.0 m columns, 4x3 m tanks, "
                        "FP32-split, adaptive CFL"}


def read_profile(path: str) -> dict:
    """`key: value` lines of the committed rocprofv3 PMC summary of the dominant
    kernel: HBM bytes per dispatch (FETCH_SIZE / WRITE_SIZE scaled by the
    copy-kernel calibration), the VALU busy fraction and the wave64 VALU
    instructions per dispatch (SQ_INSTS_VALU)."""
    vals = {}
    if not os.path.exists(path):
        return vals
    with open(path) as fh:
        for line in fh:
            k, sep, v = line.partition(":")
            if sep and k.strip() in ("traffic_bytes_per_dispatch", "valu_busy_per_dispatch",
                                     "valu_instructions_per_dispatch"):
                try:
                    vals[k.strip()] = float(v.split()[0])
                except (ValueError, IndexError):
                    pass
    return vals


def cpu_baseline(scene, seconds_budget: float = 15.0, max_steps: int = 30):
    """The reference's step on all host cores: oracle/cpu_omp.cpp, an OpenMP
    per-particle gather restating the reference's per-pair arithmetic in FP64
    (checked against the serial oracle in tests/).  Sample: consecutive
    adaptive steps of the same workload from its initial state, after one
    untimed step, until about `seconds_budget` of CPU time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # test/check infrastructure, used here only as the CPU baseline

    omp = O.OmpSPHProblem(scene.points, scene.cellsize, scene.dist, scene.params, scene.rho, scene.ptype)
    omp.stepsolve(1, True)
    steps, t0 = 0, time.perf_counter()
    while steps < max_steps and (steps == 0 or time.perf_counter() - t0 < seconds_budget):
        omp.stepsolve(1, True)
        steps += 1
    dt = time.perf_counter() - t0
    threads = omp.threads
    return {"value": scene.n * steps / dt, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(),
            "sample": f"{steps} adaptive steps (each with its neighbour rebuild) of the same {scene.n}-particle "
                      f"scene, FP64, OpenMP per-particle gather on {threads} threads (oracle/cpu_omp.cpp, "
                      f"the reference's per-pair arithmetic): {dt:.2f} s after one untimed step"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--dx", type=float, default=0.002)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-f64", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble run")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the 1000-step FP32-split vs FP64 per-field error leg")
    ap.add_argument("--accuracy-steps", type=int, default=1000)
    ap.add_argument("--scene", choices=("2d", "3d", "ensemble", "celllist"), default="2d",
                    help="2d: the headline line (default); 3d: the 883,376-particle 3D dam break with every "
                         "physics term, FP32-split and FP64; ensemble: only the batched-ensemble leg; celllist: "
                         "update! (binning + pair list) of a cell list over the 4 M-particle scene")
    ap.add_argument("--cl-points", choices=("dambreak", "uniform200k"), default="dambreak",
                    help="points of --scene celllist: the 4 M-particle dam break, or the 200,000 uniform points "
                         "of the reference's own update!/partialupdate! benchmark")
    ap.add_argument("--walls", choices=("fixture", "minus1"), default="fixture",
                    help="wall typing of the 2D dam break: the reference fixtures' (inner layer 1, outer 0) "
                         "or ptype -1 for every wall particle")
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "pass2_f32split_traffic.txt"))
    ap.add_argument("--calibrate-copy", type=float, default=0.0,
                    help="copy a buffer of this many GiB three times and exit (HBM counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy > 0:
        import torch

        n = int(args.calibrate_copy * (1 << 30)) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the SPH step is built for a single MI355X "
              "(multi-GPU sharding is a design note, not implemented)", file=sys.stderr)
        sys.exit(2)

    if args.scene == "celllist" and args.cl_points == "uniform200k":
        pts = S.config1_points()
        d = S.CONFIG1_DIST
        cl = G.GPUCellList(pts, (d, d), d)
        rng = np.random.default_rng(7)
        pts = tuple(p + rng.normal(0.0, 0.001, p.shape) for p in pts)  # one N(0, 0.001^2) jitter
        cl.points = pts
        res = {}
        for name, fn in (("update!", G.update_), ("partialupdate!", G.partialupdate_)):
            for _ in range(args.warmup):
                fn(cl)
            t0 = time.perf_counter()
            for _ in range(args.steps):
                fn(cl)
            res[name] = (time.perf_counter() - t0) / args.steps * 1e3
        # the reference's own README benchmark of this system (BenchmarkTools medians, its GPU unstated)
        ref = {"update!": 43.195, "partialupdate!": 42.673}
        print(json.dumps({"metric": "cell-list update! (binning + pair list) per second", "value": 1e3 / res["update!"],
                          "unit": "updates/s", "ms_per_update": res["update!"],
                          "ms_per_partialupdate": res["partialupdate!"], "steps": args.steps,
                          "reference_readme_median_ms": ref,
                          "config": {"workload": "GPUCellList over 200,000 uniform points in [0,1)^2 (seed 12345), "
                                                 "jittered once by N(0, 0.001^2), FP64, cell size = dist = 0.016",
                                     "particles": cl.n, "pairs": cl.pairsn}}))
        return
    if args.scene == "celllist":
        sc = S.config3(dx=args.dx, walls=-1 if args.walls == "minus1" else "fixture")
        cl = G.GPUCellList(sc.points, sc.cellsize, sc.dist)
        for _ in range(args.warmup):
            G.update_(cl)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            G.update_(cl)
        dt = (time.perf_counter() - t0) / args.steps
        print(json.dumps({"metric": "cell-list update! (binning + pair list) per second", "value": 1.0 / dt,
                          "unit": "updates/s", "ms_per_update": dt * 1e3, "steps": args.steps,
                          "config": {"workload": f"GPUCellList over the 2D dam break dx={args.dx}, FP64 points, "
                                                 "cell size = dist = H", "particles": sc.n, "pairs": cl.pairsn}}))
        return
    if args.scene == "ensemble":
        print(json.dumps({"metric": "particle-steps/s (batched ensemble)", "ensemble": run_ensemble(args.steps,
                                                                                                     args.warmup)}))
        return
    if args.scene == "3d":
        sc3 = S.config2(dx=0.01)
        r = {prec: run_gpu(sc3, prec, args.steps, args.warmup) for prec in ("f32split", "f64")}
        print(json.dumps({
            "metric": "particle-steps/s (full SPH step, 3D dam break)",
            "config": {"workload": "3D dam break dx=0.01: 1.0x1.0x0.5 m block in a 3.2x1.5x0.5 m tank, nu, "
                                   "surface tension, DPC, XSPH, CSPM, adaptive CFL", "particles": sc3.n},
            **{prec: {"value": sc3.n * args.steps / v["seconds"], "ms_per_step": v["seconds"] / args.steps * 1e3,
                      "stage_ms_per_step": {k: t[0] / args.steps for k, t in v["timing"].items() if t[1] > 0}}
               for prec, v in r.items()}}))
        return
    scene = S.config3(dx=args.dx, walls=-1 if args.walls == "minus1" else "fixture")
    n = scene.n
    log(f"scene: {n} particles; FP32-split leg")
    res32 = run_gpu(scene, "f32split", args.steps, args.warmup)
    log(f"FP32-split {res32['seconds'] / args.steps * 1e3:.3f} ms/step; FP64 leg")
    res64 = None if args.no_f64 else run_gpu(scene, "f64", args.steps, args.warmup)

    # value and ms_per_step: the production call (captured step graph);
    # stage split and the dominant kernel's duration: the eager call with HIP
    # events around every stage on the problem's stream
    value = n * args.steps / res32["seconds"]
    ms_per_step = res32["seconds"] / args.steps * 1e3
    bytes_pp = algorithmic_bytes(2, True)
    stages = res32["timing"]
    dom = max(("pass1", "pass2"), key=lambda k: stages[k][0])
    dom_ms = stages[dom][0] / max(stages[dom][1], 1)
    achieved = bytes_pp[dom] * n / (dom_ms * 1e-3) / 1e9
    prof = read_profile(args.traffic_file) if dom == "pass2" else {}
    traffic = prof.get("traffic_bytes_per_dispatch")
    out = {
        "metric": "particle-steps/s (full SPH step) FP32-split vs FP64 on 1 MI355X; HBM GB/s",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        # untimed steps actually run before the timed call: at least TUNE_STEPS
        # (the autotuner's), then the 2-step call that captures the step graph
        "warmup_steps_run": res32["warmup_steps"],
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair sums + f64 positions/dt (FP32-split)",
        "data": "synthetic",
        "config": {
            "workload": f"2D dam break dx={args.dx}: 4.0x4.0 m column ({int(scene.info['fluid'])} fluid) in a "
                        f"16x8 m tank ({int(scene.info['boundary'])} wall, "
                        f"{'ptype -1' if args.walls == 'minus1' else 'fixture typing: inner layer 1, outer 0'}), "
                        "WendlandC2, dist=H, adaptive CFL=0.2",
            "particles": n,
            "rebuilds_in_timed_steps": res32["updaten"],
            "precision": "FP32-split",
            "timed_path": "captured step graph, tuned pass variants",
            "pass_variants": res32["pass_variants"],
        },
        # ms per timed step (a stage that does not run every step, such as CSPM
        # every cspmn steps under "corrections", is spread over all steps) and
        # how many of the timed steps ran it
        "stage_ms_per_step": {k: v[0] / args.steps for k, v in stages.items() if v[1] > 0},
        "stage_launches": {k: v[1] for k, v in stages.items() if v[1] > 0},
        "stage_split_ms_per_step": res32["timed_split_seconds"] / args.steps * 1e3,
        "roofline": {
            "bound": "hbm",
            "kernel": f"k_{dom}",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_particle": bytes_pp[dom],
            "valu_busy": prof.get("valu_busy_per_dispatch"),
            # the roof that binds the pair passes: VALU issue (instructions per
            # launch from the committed SQ counters / the same measured time)
            "valu": None if "valu_instructions_per_dispatch" not in prof else {
                "instructions_per_launch": prof["valu_instructions_per_dispatch"],
                "achieved": prof["valu_instructions_per_dispatch"] / (dom_ms * 1e-3) / 1e9,
                "peak": VALU_PEAK_GIPS, "unit": "G wave64 VALU instructions/s",
                "frac": prof["valu_instructions_per_dispatch"] / (dom_ms * 1e-3) / 1e9 / VALU_PEAK_GIPS},
            "note": "pair passes are co-limited by vector-L1 access rate and VALU issue, not HBM: "
                    "achieved = algorithmic bytes / measured kernel time; "
                    "valu_busy (share of SIMD cycles issuing VALU) and traffic come from the committed PMC profile",
        },
    }
    if res64 is not None:
        v64 = n * args.steps / res64["seconds"]
        out["f64"] = {"value": v64, "ms_per_step": res64["seconds"] / args.steps * 1e3,
                      "pass_variants": res64["pass_variants"],
                      "stage_ms_per_step": {k: v[0] / args.steps for k, v in res64["timing"].items() if v[1] > 0}}
        out["f32split_over_f64"] = value / v64
    if not args.no_ensemble:
        log("ensemble leg")
        out["ensemble"] = run_ensemble(args.steps, args.warmup)
    if not args.no_accuracy:
        log(f"accuracy leg: {args.accuracy_steps} steps FP64, FP64 perturbed, FP32-split")
        out["accuracy"] = accuracy(scene, args.accuracy_steps)
    if not args.no_cpu:
        log("CPU baseline leg")
        out["cpu_baseline"] = cpu_baseline(scene)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
