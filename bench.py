#!/usr/bin/env python3
"""Oseen3D 32^3 Q2/P1disc: FGMRES(30) + GMG cell-Vanka solve to rtol 1e-10 on one MI355X.

One step = one complete solve of the Oseen system with the operator resident in HBM:
the value-dependent set-up (batched LU of every 85x85 Vanka patch on levels 8^3..32^3 and
the dense LU of the 4^3 level) followed by FGMRES from x = 0 until the restart-cycle
relative residual drops below 1e-10, each FGMRES iteration preconditioned by one V(2,2)
cycle.  ``value`` is seconds per solve (lower is better) of the mixed-precision form (by
default bf16 patch factors, FP32 smoother and level operators, FP64 FGMRES and coarse
solve); the plain FP64 form is timed beside it.

A separate profiling pass then times every hot-kernel launch with HIP events on the
library's stream and divides the kernel's algorithmic bytes by its time (roofline), and
the FP64 CPU oracle (reference-faithful: patch matrices re-assembled and LU-factored in
every sweep, coarse matrix refactored at every visit) is timed for a bounded number of
FGMRES iterations on the host cores and scaled to the GPU's iteration count.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Oseen3D 32^3 Q2/P1disc solve s to 1e-10: mixed/batched vs FP64 GMG-Vanka"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


USE_GRAPH = 1
SHARED_BLOCK = 1
PATCH_INVERSE = 1


def solver_params(precision: int):
    from to_merge_master_parmoon_amd import _abi
    # FGMRES(30), SC_LIN_RED_FACTOR_SADDLE = 1e-10, SC_LIN_RES_NORM_MIN_SADDLE = 0,
    # SC_LIN_MAXIT_SADDLE = 1000, SC_AMG_PREC_IT = 1, V(2,2), damp = alpha = beta = 1
    return _abi.default_params(sc_gmres_restart=30, sc_lin_red_factor_saddle=1e-10,
                               sc_lin_res_norm_min_saddle=0.0, sc_lin_maxit_saddle=1000,
                               sc_pre_smooth_saddle=2, sc_post_smooth_saddle=2, precision=precision,
                               use_graph=USE_GRAPH, shared_block=SHARED_BLOCK, patch_inverse=PATCH_INVERSE)


def run_gpu(prob, precision: int, steps: int, warmup: int, log):
    from to_merge_master_parmoon_amd.multigrid import DeviceVector, build_hierarchy
    t0 = time.time()
    mg = build_hierarchy(prob, solver_params(precision))
    log(f"hierarchy on device ({time.time() - t0:.1f}s), levels {mg.n_dof}")
    n = mg.n_dof[-1]
    rhs = DeviceVector(mg, n, prob.rhs(-1))
    zero = DeviceVector(mg, n, np.zeros(n))
    sol = DeviceVector(mg, n)
    lib, ctx = mg._lib, mg._ctx

    def step():
        mg._check(lib.pmg_memcpy(ctx, sol.ptr, zero.ptr, 8 * n, 2))
        mg.refactor()
        return mg.fgmres_device(sol, rhs)

    for w in range(warmup):
        it, res = step()
        log(f"warmup {w}: {it} iterations, res {res:.3e}")
    mg.synchronize()
    times, iters = [], []
    for k in range(steps):
        mg.synchronize()
        t = time.perf_counter()
        it, res = step()
        mg.synchronize()
        times.append(time.perf_counter() - t)
        iters.append(it)
        log(f"step {k}: {times[-1] * 1e3:.1f} ms, {it} iterations, res {res:.3e}")
    x = sol.download()
    hist = mg.fgmres_history()
    # profiling pass: per-launch HIP events (synchronizing; not part of the timed steps)
    mg.reset_timers()
    mg.set_profiling(True)
    step()
    mg.set_profiling(False)
    tm = mg.timers()
    return dict(times=times, iters=iters, x=x, hist=hist, timers=tm, mg=mg)


def run_batched(args, log):
    """The batched workload: independent perturbed-cube Oseen problems of 8^3..24^3 cells
    (problem.batched_specs) solved in one batch context, each FGMRES step one merged
    V-cycle for all problems; timed beside the same problems solved one context at a time."""
    from to_merge_master_parmoon_amd import problem
    from to_merge_master_parmoon_amd.multigrid import DeviceVector, NSEBatch, build_hierarchy
    prec = {"fp64": 0, "fp32": 1, "bf16": 2}[args.precision]
    specs = problem.batched_specs(args.batch_count)
    t0 = time.time()
    probs = []
    for i, (n, nu, f) in enumerate(specs):
        probs.append(problem.batch_instance(n, nu, f, i))
        if (i + 1) % 8 == 0:
            log(f"generated {i + 1}/{len(specs)} problems ({time.time() - t0:.1f}s)")
    ndof = [p.rhs(-1).size for p in probs]
    log(f"{len(probs)} problems, {sum(ndof)} dofs in total")

    def params_for(p):
        sp = solver_params(p)
        sp.sc_lin_maxit_saddle = args.batch_maxit
        return sp

    def batched(p):
        t = time.time()
        B = NSEBatch(probs, params=params_for(p))
        log(f"batch context ({time.time() - t:.1f}s)")
        rh = [DeviceVector(B, ndof[i], probs[i].rhs(-1)) for i in range(len(probs))]
        zs = [DeviceVector(B, ndof[i], np.zeros(ndof[i])) for i in range(len(probs))]
        import ctypes as C
        nb = len(probs)
        sp = (C.c_void_p * nb)(*[z.ptr for z in zs])
        rp = (C.c_void_p * nb)(*[r.ptr for r in rh])
        it, res, stv = (C.c_int * nb)(), (C.c_double * nb)(), (C.c_int * nb)()
        times = []
        for k in range(args.warmup + args.steps):
            for z in zs:
                z.upload(np.zeros(z.n))
            B.synchronize()
            t = time.perf_counter()
            B.refactor()
            B._lib.pmg_batch_fgmres_solve(B._ctx, sp, rp, it, res, stv, 1)
            B.synchronize()
            dt = time.perf_counter() - t
            if k >= args.warmup:
                times.append(dt)
            log(f"batch {['fp64', 'fp32', 'bf16'][p]} step {k}: {dt * 1e3:.1f} ms, "
                f"converged {sum(1 for s in stv if s == 0)}/{nb}")
        xs = [z.download() for z in zs]
        B.close()
        return float(np.mean(times)), list(it), list(stv), xs

    t_b, it_b, st_b, x_b = batched(prec)
    t_64 = it_64 = None
    if prec != 0 and not args.no_fp64_ref:
        t_64, it_64, _, x_64 = batched(0)
    # one problem at a time, same precision, each in its own context
    seq = 0.0
    for i, pr in enumerate(probs):
        mg = build_hierarchy(pr, params_for(prec))
        rhs = DeviceVector(mg, ndof[i], pr.rhs(-1))
        sol = DeviceVector(mg, ndof[i], np.zeros(ndof[i]))
        mg.synchronize()
        t = time.perf_counter()
        mg.refactor()
        st = mg._lib.pmg_fgmres_solve(mg._ctx, sol.ptr, rhs.ptr, None, None, 1)
        mg.synchronize()
        seq += time.perf_counter() - t
        mg.close()
    log(f"one at a time: {seq * 1e3:.1f} ms")
    conv = sum(1 for s in st_b if s == 0)
    cfg = {"workload": f"batched_oseen_{len(probs)}_problems_8-24^3_q2p1disc_fgmres30_vcycle22",
           "problems": len(probs), "dofs": int(sum(ndof)), "precision": args.precision,
           "maxit": args.batch_maxit, "converged": conv,
           "iterations": it_b, "one_at_a_time_s": round(seq, 4), "speedup_vs_one_at_a_time": round(seq / t_b, 3)}
    if t_64 is not None:
        cfg.update({"fp64_batch_s": round(t_64, 4), "speedup_vs_fp64": round(t_64 / t_b, 3),
                    "fp64_iterations": it_64})
    return {"metric": "Batched Oseen Q2/P1disc solves s to 1e-10 (all problems, one batch)",
            "value": round(t_b, 5), "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(t_b * 1e3, 3), "higher_is_better": False, "scaling": "strong",
            "vs_baseline": None, "dtype": args.precision, "data": "synthetic", "config": cfg}


def run_reassembly(args, log):
    """One fixed-point step's device work before the solve on the 32^3 hierarchy: velocity at
    the Gauss points of every level (pmg_eval_velocity), device re-assembly of A11..A33
    (pmg_assemble_oseen) and the value-dependent set-up (pmg_refactor); timed beside the host
    assembly of the same blocks by the problem generator (one process, all host threads)."""
    import ctypes as C
    from to_merge_master_parmoon_amd import problem
    from to_merge_master_parmoon_amd.multigrid import DeviceVector, build_hierarchy
    prec = {"fp64": 0, "fp32": 1, "bf16": 2}[args.precision]
    t0 = time.time()
    prob = problem.oseen3d(args.n)
    host_assembly_s = time.time() - t0
    mg = build_hierarchy(prob, solver_params(prec))
    L = prob.n_levels
    for l in range(L):
        mg.level_geometry(l, prob.cell_vertices(l))
    lib, ctx = mg._lib, mg._ctx
    u = [DeviceVector(mg, mg.n_dof[l], np.random.default_rng(l).standard_normal(mg.n_dof[l])) for l in range(L)]
    bq = [DeviceVector(mg, prob.descs[l].n_cells * 81) for l in range(L)]

    def step(refactor=True):
        for l in range(L):
            mg._check(lib.pmg_eval_velocity(ctx, l, u[l].ptr, bq[l].ptr, 1))
            mg._check(lib.pmg_assemble_oseen(ctx, l, bq[l].ptr, 1e-2, 0.0, 1))
        if refactor:
            mg.refactor()

    for _ in range(args.warmup):
        step()
    mg.synchronize()
    times, times_asm = [], []
    for k in range(args.steps):
        mg.synchronize()
        t = time.perf_counter()
        step(refactor=False)
        mg.synchronize()
        times_asm.append(time.perf_counter() - t)
        t = time.perf_counter()
        step()
        mg.synchronize()
        times.append(time.perf_counter() - t)
        log(f"step {k}: eval + assembly {times_asm[-1] * 1e3:.2f} ms, with set-up {times[-1] * 1e3:.2f} ms")
    mg.close()
    nnz = prob.descs[-1].a_nnz
    cells = prob.descs[-1].n_cells
    return {"metric": f"Oseen3D {args.n}^3 Q2/P1disc fixed-point step device work s (eval + re-assembly + set-up)",
            "value": round(float(np.mean(times)), 5), "unit": "s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(1e3 * float(np.mean(times)), 3), "higher_is_better": False,
            "scaling": "strong", "vs_baseline": None, "dtype": "f64 assembly", "data": "synthetic",
            "config": {"workload": f"oseen3d_{args.n}^3_reassembly_{L}_levels", "precision": args.precision,
                       "eval_plus_assembly_s": round(float(np.mean(times_asm)), 5),
                       "finest_cells": cells, "finest_nnz_per_block": nnz,
                       "host_generator_assembly_s": round(host_assembly_s, 3),
                       "host_generator_threads": os.cpu_count()}}


def measured_traffic(kernel: str, precision: str, n: int):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary (calibrated against a copy kernel), for the workload it was measured on."""
    if precision not in ("fp32", "bf16") or n != 32:   # the mixed paths share k_vanka_res<float>
        return None, None
    path = ROOT / "profiles" / f"k_vanka_res_oseen32_{precision}_hbm_traffic.md"
    if not path.exists():
        return None, None
    for line in path.read_text().splitlines():
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch:" and parts[1] == kernel:
            return float(parts[2]), str(path.relative_to(ROOT))
    return None, None


def roofline_from(tm: dict, precision: str, n: int):
    kern = {k: v for k, v in tm.items() if isinstance(v, dict) and v["calls"] > 0}
    solve_kernels = {k: v for k, v in kern.items() if k not in ("patch_lu", "coarse_lu")}
    dom = max(solve_kernels, key=lambda k: solve_kernels[k]["ms"])
    v = kern[dom]
    avg_ms = v["ms"] / v["calls"]
    bytes_per_launch = v["bytes"] / v["calls"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    breakdown = {k: dict(ms=round(x["ms"], 3), calls=int(x["calls"]),
                         avg_us=round(1e3 * x["ms"] / x["calls"], 2),
                         gbs=round(x["bytes"] / (x["ms"] * 1e-3) / 1e9, 1) if x["ms"] > 0 else None)
                 for k, x in kern.items()}
    names = {"vanka_res": "k_vanka_res", "vanka_solve": "k_vanka_solve", "defect": "k_type4<2>",
             "matvec": "k_type4<0/1>", "transfer": "k_csr", "coarse_solve": "k_trsv_fwd/bwd"}
    traffic, src = measured_traffic(names.get(dom, dom), precision, n)
    roof = {"bound": "hbm", "kernel": names.get(dom, dom), "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None if traffic is None else round(traffic),
            "bytes_per_launch": round(bytes_per_launch), "avg_launch_us": round(avg_ms * 1e3, 2)}
    if src:
        roof["traffic_source"] = src
    return roof, breakdown


def cpu_baseline(prob, gpu_iters: int, cpu_iters: int, log):
    sys.path.insert(0, str(ROOT / "oracle"))
    import pmg_oracle
    threads = min(16, os.cpu_count() or 1)
    pmg_oracle.set_threads(threads)
    prm = solver_params(0)
    prm.sc_lin_maxit_saddle = cpu_iters
    orc = pmg_oracle.Oracle(prm, prob.descs, refactor=1)
    f = prob.rhs(-1).copy()
    x = np.zeros_like(f)
    t = time.perf_counter()
    it, res, st = orc.fgmres(x, f)
    dt = time.perf_counter() - t
    orc.close()
    per_it = dt / max(it, 1)
    log(f"cpu oracle: {it} iterations in {dt:.1f}s ({per_it:.2f} s/iteration, {threads} threads)")
    return {"value": round(per_it * gpu_iters, 3), "unit": "s", "cores": threads, "kind": "port",
            "sample": f"{it} FGMRES iterations of the reference-faithful FP64 oracle on the 32^3 problem "
                      f"(per-sweep patch re-assembly + LU, per-visit coarse LU) timed: {dt:.2f} s; "
                      f"scaled to the GPU's {gpu_iters} iterations"}


def rel_err(x, ref, n_u3):
    """relative L2 differences of the velocity block and of the pressure block; the
    pressure is compared modulo its constant mode (mean of the P1disc constants), which
    the all-Dirichlet Oseen operator leaves undetermined"""
    du = np.linalg.norm(x[:n_u3] - ref[:n_u3]) / max(np.linalg.norm(ref[:n_u3]), 1e-300)
    px, pr = x[n_u3:].copy(), ref[n_u3:].copy()
    px[0::4] -= px[0::4].mean()
    pr[0::4] -= pr[0::4].mean()
    dp = np.linalg.norm(px - pr) / max(np.linalg.norm(pr), 1e-300)
    return float(du), float(dp)


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=32, help="cells per edge of the finest level")
    ap.add_argument("--precision", choices=["fp64", "fp32", "bf16"], default="bf16",
                    help="smoother/level-operator arithmetic of the measured solve (FGMRES is FP64)")
    ap.add_argument("--cpu-iters", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64-ref", action="store_true", help="skip the FP64 GPU solve timed beside it")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--workload", choices=["oseen32", "batched", "reassembly"], default="oseen32",
                    help="batched: the independent-problems workload; reassembly: one fixed-point step's "
                         "device re-assembly and set-up (instead of the headline solve)")
    ap.add_argument("--batch-count", type=int, default=64)
    ap.add_argument("--batch-maxit", type=int, default=100)
    ap.add_argument("--no-graph", action="store_true", help="launch every V-cycle kernel by kernel")
    ap.add_argument("--lu-sweeps", action="store_true",
                    help="mixed precision: stored LU factors and triangular sweeps instead of patch inverses")
    ap.add_argument("--nine-blocks", action="store_true",
                    help="stream all nine A blocks even where A11 = A22 = A33 and the coupling blocks vanish")
    args = ap.parse_args()
    global USE_GRAPH, SHARED_BLOCK, PATCH_INVERSE
    USE_GRAPH = 0 if args.no_graph else 1
    SHARED_BLOCK = 0 if args.nine_blocks else 1
    PATCH_INVERSE = 0 if args.lu_sweeps else 1
    log = (lambda *a: None) if args.quiet else (lambda *a: print("[bench]", *a, file=sys.stderr, flush=True))
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): this solver has no multi-GPU "
              "path (one problem, or one batch of independent problems, per GPU).", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched":
        print(json.dumps(run_batched(args, log)), flush=True)
        return
    if args.workload == "reassembly":
        print(json.dumps(run_reassembly(args, log)), flush=True)
        return
    from to_merge_master_parmoon_amd import problem
    prec = {"fp64": 0, "fp32": 1, "bf16": 2}[args.precision]
    t0 = time.time()
    prob = problem.oseen3d(args.n)
    sz = prob.sizes(-1)
    n_u3 = 3 * sz["n_udof"]
    log(f"problem {args.n}^3: {sz['n_dof']} dofs, nnz(A) {sz['a_nnz']} ({time.time() - t0:.1f}s)")
    ref = None
    if prec != 0 and not args.no_fp64_ref:
        log("FP64 GPU solve (plain form, timed beside the measured one)")
        ref = run_gpu(prob, 0, max(1, args.steps), min(args.warmup, 1), log)
        ref["mg"].close()
    log(f"measured solve: {args.precision}")
    g = run_gpu(prob, prec, args.steps, args.warmup, log)
    ms = 1e3 * float(np.mean(g["times"]))
    roof, breakdown = roofline_from(g["timers"], args.precision, args.n)
    iters = int(g["iters"][-1])
    cfg = {"workload": f"oseen3d_{args.n}^3_q2p1disc_fgmres30_vcycle22", "levels": len(g['mg'].n_dof),
           "dofs": sz["n_dof"], "precision": args.precision, "nu": 1e-2,
           "fgmres_iterations": iters, "final_residual": float(g["hist"][-1]) if len(g["hist"]) else None,
           "setup_ms": round(max(g["timers"]["patch_lu"]["ms"], g["timers"]["coarse_lu"]["ms"]), 3)}
    if ref is not None:
        fp64_ms = 1e3 * float(np.mean(ref["times"]))
        du, dp = rel_err(g["x"], ref["x"], n_u3)
        cfg.update({"fp64_solve_s": round(fp64_ms / 1e3, 5), "fp64_iterations": int(ref["iters"][-1]),
                    "speedup_vs_fp64": round(fp64_ms / ms, 3),
                    "rel_err_vs_fp64": {"velocity": du, "pressure": dp}})
    out = {
        "metric": METRIC,
        "value": round(ms / 1e3, 5),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32 smoother+level operator / f64 FGMRES+coarse",
                  "bf16": "bf16 patch data / f32 smoother / f64 FGMRES+coarse"}[args.precision],
        "data": "synthetic",
        "config": cfg,
        "roofline": roof,
        "kernels": breakdown,
    }
    out["cpu_baseline"] = None if args.no_cpu else cpu_baseline(prob, iters, args.cpu_iters, log)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
