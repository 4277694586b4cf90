#!/usr/bin/env python3
"""Oseen3D 32^3 Q2/P1disc: FGMRES(30) + GMG cell-Vanka solve to rtol 1e-10 on one MI355X.

One step = one complete solve of the Oseen system with the operator resident in HBM:
the value-dependent set-up (batched local data of every 85x85 Vanka patch on levels
8^3..32^3 and the dense factors of the 4^3 level) followed by FGMRES from x = 0 until the
restart-cycle relative residual drops below 1e-10, each FGMRES iteration preconditioned by
one V(2,2) cycle.  ``value`` is seconds per solve (lower is better) of the mixed-precision
form (by default bf16 patch data, FP32 smoother and level operators, FP64 FGMRES and
coarse solve); the plain FP64 form is timed beside it, both with the shared velocity block
(one A value stream) and with all nine A blocks streamed.

A separate profiling pass times every hot-kernel launch with HIP events on the library's
stream; the roofline is the finest level's local-residual kernel (algorithmic bytes per
launch / average launch time).  The FP64 CPU oracle (reference-faithful: patch matrices
re-assembled and LU-factored in every sweep, coarse matrix refactored at every visit) then
solves the same system to 1e-10 on the host cores, with a 1-core sample and a
cached-factor variant; its solution gives the GPU forms' errors against the oracle.  The
batched workload (64 independent problems) is measured once per precision at the end.
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


def solver_params(precision: int, shared_block: int | None = None, patch_inverse: int | None = None,
                  refine_steps: int = 0):
    from to_merge_master_parmoon_amd import _abi
    # FGMRES(30), SC_LIN_RED_FACTOR_SADDLE = 1e-10, SC_LIN_RES_NORM_MIN_SADDLE = 0,
    # SC_LIN_MAXIT_SADDLE = 1000, SC_AMG_PREC_IT = 1, V(2,2), damp = alpha = beta = 1
    return _abi.default_params(sc_gmres_restart=30, sc_lin_red_factor_saddle=1e-10,
                               sc_lin_res_norm_min_saddle=0.0, sc_lin_maxit_saddle=1000,
                               sc_pre_smooth_saddle=2, sc_post_smooth_saddle=2, precision=precision,
                               use_graph=USE_GRAPH,
                               shared_block=SHARED_BLOCK if shared_block is None else shared_block,
                               patch_inverse=PATCH_INVERSE if patch_inverse is None else patch_inverse,
                               bf16_refine_steps=refine_steps)


AUTOTUNE = True
TUNING_FILE = None


def run_gpu(prob, precision: int, steps: int, warmup: int, log, shared_block=None, patch_inverse=None,
            solve_only_steps: int = 0, profile: bool = True, refine_steps: int = 0):
    """`steps` timed solves including the value-dependent set-up (pmg_refactor), then
    `solve_only_steps` solves with the factors kept (their difference is the set-up's cost
    on the critical path), then one profiling pass.  Before the warm-up the launch shapes
    are tuned on the device (pmg_autotune, once per hierarchy, outside the timed steps)."""
    from to_merge_master_parmoon_amd.multigrid import DeviceVector, build_hierarchy
    t0 = time.time()
    mg = build_hierarchy(prob, solver_params(precision, shared_block, patch_inverse, refine_steps))
    log(f"hierarchy on device ({time.time() - t0:.1f}s), levels {mg.n_dof}, "
        f"shared block {[mg.shared_block(l) for l in range(mg.n_levels)]}")
    shapes = None
    key = f"fp{['64', '32', '32bf16'][precision]}_sb{mg.shared_block(mg.n_levels - 1)}"
    if TUNING_FILE and Path(TUNING_FILE).exists() and key in json.loads(Path(TUNING_FILE).read_text()):
        from to_merge_master_parmoon_amd.multigrid import load_tuning
        tmp = Path(TUNING_FILE).with_suffix(".one.json")
        tmp.write_text(json.dumps(json.loads(Path(TUNING_FILE).read_text())[key]))
        shapes = load_tuning(mg, str(tmp))
        tmp.unlink()
        log(f"launch shapes loaded from {TUNING_FILE}: {shapes}")
    elif AUTOTUNE:
        t0 = time.time()
        shapes = mg.autotune(20)
        log(f"launch shapes tuned ({time.time() - t0:.2f}s): {shapes}")
        if TUNING_FILE:
            saved = json.loads(Path(TUNING_FILE).read_text()) if Path(TUNING_FILE).exists() else {}
            saved[key] = {"levels": shapes}
            Path(TUNING_FILE).write_text(json.dumps(saved, indent=1))
    n = mg.n_dof[-1]
    rhs = DeviceVector(mg, n, prob.rhs(-1))
    zero = DeviceVector(mg, n, np.zeros(n))
    sol = DeviceVector(mg, n)
    lib, ctx = mg._lib, mg._ctx

    def step(refactor=True):
        mg._check(lib.pmg_memcpy(ctx, sol.ptr, zero.ptr, 8 * n, 2))
        if refactor:
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
    solve_only = []
    for k in range(solve_only_steps):
        mg.synchronize()
        t = time.perf_counter()
        step(refactor=False)
        mg.synchronize()
        solve_only.append(time.perf_counter() - t)
    if solve_only:
        log(f"solve with the factors kept: {1e3 * float(np.mean(solve_only)):.1f} ms")
    tm = None
    if profile:
        # per-launch HIP events (synchronizing; not part of the timed steps)
        mg.reset_timers()
        mg.set_profiling(True)
        step()
        mg.set_profiling(False)
        tm = mg.timers()
    return dict(times=times, iters=iters, x=x, hist=hist, timers=tm, mg=mg, solve_only=solve_only, shapes=shapes)


def batched_problems(count: int, log):
    from to_merge_master_parmoon_amd import problem
    specs = problem.batched_specs(count)
    t0 = time.time()
    probs = []
    for i, (n, nu, f) in enumerate(specs):
        probs.append(problem.batch_instance(n, nu, f, i))
        if (i + 1) % 8 == 0:
            log(f"generated {i + 1}/{len(specs)} problems ({time.time() - t0:.1f}s)")
    return specs, probs


def batched_solve(probs, precision: int, maxit: int, steps: int, warmup: int, log):
    """One batch context for all problems; per step pmg_refactor (patch data of the merged
    levels + every problem's coarse factors) + pmg_batch_fgmres_solve from x = 0."""
    import ctypes as C
    from to_merge_master_parmoon_amd.multigrid import DeviceVector, NSEBatch
    ndof = [p.rhs(-1).size for p in probs]
    sp_ = solver_params(precision)
    sp_.sc_lin_maxit_saddle = maxit
    t = time.time()
    B = NSEBatch(probs, params=sp_)
    log(f"batch context ({time.time() - t:.1f}s)")
    if AUTOTUNE:
        B.refactor()
        log(f"batch launch shapes tuned: {B.autotune(20)}")
    rh = [DeviceVector(B, ndof[i], probs[i].rhs(-1)) for i in range(len(probs))]
    zs = [DeviceVector(B, ndof[i], np.zeros(ndof[i])) for i in range(len(probs))]
    nb = len(probs)
    sp = (C.c_void_p * nb)(*[z.ptr for z in zs])
    rp = (C.c_void_p * nb)(*[r.ptr for r in rh])
    it, res, stv = (C.c_int * nb)(), (C.c_double * nb)(), (C.c_int * nb)()
    times = []
    for k in range(warmup + steps):
        for z in zs:
            z.upload(np.zeros(z.n))
        B.synchronize()
        t = time.perf_counter()
        B.refactor()
        B._lib.pmg_batch_fgmres_solve(B._ctx, sp, rp, it, res, stv, 1)
        B.synchronize()
        dt = time.perf_counter() - t
        if k >= warmup:
            times.append(dt)
        log(f"batch {['fp64', 'fp32', 'bf16'][precision]} step {k}: {dt * 1e3:.1f} ms, "
            f"converged {sum(1 for s in stv if s == 0)}/{nb}")
    xs = [z.download() for z in zs]
    B.close()
    return float(np.mean(times)), list(it), list(stv), xs


def batched_summary(specs, probs, prec: int, maxit: int, steps: int, warmup: int, log, fp64_ref=True):
    t_b, it_b, st_b, x_b = batched_solve(probs, prec, maxit, steps, warmup, log)
    out = {"problems": len(probs), "dofs": int(sum(p.rhs(-1).size for p in probs)),
           "precision": ["fp64", "fp32", "bf16"][prec], "maxit": maxit, "solve_s": round(t_b, 4),
           "converged": sum(1 for s in st_b if s == 0), "iterations": it_b}
    stalled = [i for i, s in enumerate(st_b) if s != 0]
    out["stalled"] = [{"problem": i, "n": specs[i][0], "nu": round(specs[i][1], 5),
                       "dirichlet_faces": "all six" if specs[i][2] == 63 else "x=0,y=0,z=0", "status": st_b[i]}
                      for i in stalled]
    if prec != 0 and fp64_ref:
        t_64, it_64, st_64, x_64 = batched_solve(probs, 0, maxit, steps, warmup, log)
        errs = []
        for i, p in enumerate(probs):
            if st_b[i] == 0 and st_64[i] == 0:
                nu_ = p.descs[-1].n_udof
                du, dp = rel_err(x_b[i], x_64[i], 3 * nu_)
                errs.append(max(du, dp))
        out.update({"fp64_solve_s": round(t_64, 4), "speedup_vs_fp64": round(t_64 / t_b, 3),
                    "fp64_converged": sum(1 for s in st_64 if s == 0),
                    "same_iterations_as_fp64": it_b == it_64, "same_status_as_fp64": st_b == st_64,
                    "max_rel_err_vs_fp64_converged": max(errs) if errs else None})
    return out


def run_batched(args, log):
    """The batched workload: independent perturbed-cube Oseen problems of 8^3..24^3 cells
    (problem.batched_specs) solved in one batch context, each FGMRES step one merged
    V-cycle for all problems; timed beside the same problems solved one context at a time."""
    from to_merge_master_parmoon_amd.multigrid import DeviceVector, build_hierarchy
    prec = {"fp64": 0, "fp32": 1, "bf16": 2}[args.precision]
    specs, probs = batched_problems(args.batch_count, log)
    ndof = [p.rhs(-1).size for p in probs]
    log(f"{len(probs)} problems, {sum(ndof)} dofs in total")
    summ = batched_summary(specs, probs, prec, args.batch_maxit, args.steps, args.warmup, log,
                           fp64_ref=not args.no_fp64_ref)
    # one problem at a time, same precision, each in its own context
    seq = 0.0
    for i, pr in enumerate([] if args.no_one_at_a_time else probs):
        sp_ = solver_params(prec)
        sp_.sc_lin_maxit_saddle = args.batch_maxit
        mg = build_hierarchy(pr, sp_)
        rhs = DeviceVector(mg, ndof[i], pr.rhs(-1))
        sol = DeviceVector(mg, ndof[i], np.zeros(ndof[i]))
        mg.synchronize()
        t = time.perf_counter()
        mg.refactor()
        mg._lib.pmg_fgmres_solve(mg._ctx, sol.ptr, rhs.ptr, None, None, 1)
        mg.synchronize()
        seq += time.perf_counter() - t
        mg.close()
    log(f"one at a time: {seq * 1e3:.1f} ms")
    cfg = {"workload": f"batched_oseen_{len(probs)}_problems_8-24^3_q2p1disc_fgmres30_vcycle22", **summ}
    if not args.no_one_at_a_time:
        cfg.update({"one_at_a_time_s": round(seq, 4), "speedup_vs_one_at_a_time": round(seq / summ["solve_s"], 3)})
    return {"metric": "Batched Oseen Q2/P1disc solves s to 1e-10 (all problems, one batch)",
            "value": summ["solve_s"], "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(summ["solve_s"] * 1e3, 3), "higher_is_better": False, "scaling": "strong",
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


def run_channel(args, log):
    """The time-dependent channel workload: [0,2.5] x [0,0.41]^2, 96 x 16 x 16 hexahedra
    (4 levels), Re = 100 parabolic inflow with a seeded 1 % perturbation, Crank-Nicolson
    dt = 0.01 for `--channel-steps` steps with 3 fixed-point Oseen iterations each, state in
    device memory (to_merge_master_parmoon_amd.timeloop); the measured precision beside
    the FP64 loop, whose solution at every step gives the error."""
    from to_merge_master_parmoon_amd import problem
    from to_merge_master_parmoon_amd.timeloop import ThetaTimeLoop, perturbed_inflow
    prec = {"fp64": 0, "fp32": 1, "bf16": 2}[args.precision]
    prob = problem.channel()
    g = perturbed_inflow(prob)
    nu, dt = 0.0041, 0.01

    def loop(p_):
        sp_ = solver_params(p_)
        sp_.internal_project_pressure = 0   # natural outflow: the pressure is determined
        tl = ThetaTimeLoop(prob, sp_, nu, dt, dirichlet=g)
        tl.step()   # warm-up step (graph capture, first-touch), then restart from rest
        tl.close()
        tl = ThetaTimeLoop(prob, sp_, nu, dt, dirichlet=g)
        out = []
        for k in range(args.channel_steps):
            t, its = tl.step()
            out.append((t, its, tl.solution()))
            log(f"channel {['fp64', 'fp32', 'bf16'][p_]} step {k}: {t * 1e3:.1f} ms, FGMRES iterations {its}")
        tl.close()
        return out

    meas = loop(prec)
    ref = loop(0) if prec != 0 else meas
    n_u3 = 3 * prob.descs[-1].n_udof
    steps = []
    for k, ((t, its, x), (t64, its64, x64)) in enumerate(zip(meas, ref)):
        du = np.linalg.norm(x[:n_u3] - x64[:n_u3]) / np.linalg.norm(x64[:n_u3])
        dp = np.linalg.norm(x[n_u3:] - x64[n_u3:]) / max(np.linalg.norm(x64[n_u3:]), 1e-300)
        steps.append({"step": k, "s": round(t, 5), "fp64_s": round(t64, 5), "iterations": its,
                      "fp64_iterations": its64, "rel_err_vs_fp64": {"velocity": float(du), "pressure": float(dp)}})
    mean = float(np.mean([st["s"] for st in steps]))
    mean64 = float(np.mean([st["fp64_s"] for st in steps]))
    cfg = {"workload": "tnse3d_channel_96x16x16_q2p1disc_cn_dt0.01_picard3", "dofs": prob.sizes(-1)["n_dof"],
           "levels": prob.n_levels, "re": 100, "nu": nu, "steps": args.channel_steps, "picard_per_step": 3,
           "precision": args.precision, "fp64_step_s": round(mean64, 5), "speedup_vs_fp64": round(mean64 / mean, 3),
           "max_rel_err_vs_fp64": max(max(st["rel_err_vs_fp64"].values()) for st in steps), "per_step": steps}
    return {"metric": "TNSE3D channel 96x16x16 Crank-Nicolson step s (3 fixed-point Oseen solves)",
            "value": round(mean, 5), "unit": "s", "n_gpus": 1, "steps": args.channel_steps, "warmup": 1,
            "ms_per_step": round(mean * 1e3, 3), "higher_is_better": False, "scaling": "strong",
            "vs_baseline": None, "dtype": args.precision, "data": "synthetic", "config": cfg}


def measured_traffic(kernel: str, precision: str, n: int):
    """HBM bytes per launch of `kernel` on the finest level from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (calibrated against a copy kernel)."""
    if precision not in ("fp32", "bf16") or n != 32:   # the mixed paths share k_vanka_res<float>
        return None, None
    path = ROOT / "profiles" / f"k_vanka_res_oseen32_{precision}_hbm_traffic.md"
    if not path.exists():
        return None, None
    for line in path.read_text().splitlines():
        parts = line.split()
        if len(parts) == 3 and parts[0] == "finest_level_traffic_bytes_per_launch:" and parts[1] == kernel:
            return float(parts[2]), str(path.relative_to(ROOT))
    return None, None


def general_type4_res_bytes(prob, vs: int = 4) -> float:
    """Algorithmic bytes of one finest-level local-residual launch (one color) in the general
    type-4 format: every A row entry with its nine values (3x3 block) and a column index,
    BT and B entries with three values and an index, no credit for equal or zero blocks;
    plus per patch dof the right-hand side read, the local residual written, its dof index
    and permutation byte; averaged over the colors (each color holds 1/8 of the cells)."""
    d = prob.descs[-1]
    a = prob.arrays(prob.n_levels - 1)
    na = np.diff(a["a_rowptr"]).astype(np.float64)
    nbt = np.diff(a["bt_rowptr"]).astype(np.float64)
    nb = np.diff(a["b_rowptr"]).astype(np.float64)
    act = np.arange(d.n_udof) < d.n_active
    vel = np.where(act, na * (9 * vs + 4) + nbt * (3 * vs + 4) + 16, na * (3 * vs + 4) + 8)
    pre = nb * (3 * vs + 4) + 8
    ug, pg = a["u_global"], a["p_global"]
    total = (vel[ug].sum() + ug.size * (3 * vs + 3 * vs + 4 + 3) + pre[pg].sum() + pg.size * (vs + vs + 4 + 1)
             + d.n_cells * (4 + 31 * 4))
    from to_merge_master_parmoon_amd.multigrid import color_cells
    nc, _, _ = color_cells(d.n_cells, ug, d.n_udof)
    return float(total) / nc


def roofline_from(tm: dict, precision: str, n: int, prob):
    """Roofline of the finest level's local-residual kernel from the profiling pass."""
    kern = {k: v for k, v in tm.items() if isinstance(v, dict) and v["calls"] > 0}
    v = tm["vanka_res_finest"]
    avg_ms = v["ms"] / v["calls"]
    bytes_per_launch = v["bytes"] / v["calls"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    gen = general_type4_res_bytes(prob)
    breakdown = {k: dict(ms=round(x["ms"], 3), calls=int(x["calls"]),
                         avg_us=round(1e3 * x["ms"] / x["calls"], 2),
                         gbs=round(x["bytes"] / (x["ms"] * 1e-3) / 1e9, 1) if x["ms"] > 0 else None)
                 for k, x in kern.items()}
    traffic, src = measured_traffic("k_vanka_res", precision, n)
    roof = {"bound": "hbm", "kernel": "k_vanka_res (finest level, one color per launch)",
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None if traffic is None else round(traffic),
            "bytes_per_launch": round(bytes_per_launch), "avg_launch_us": round(avg_ms * 1e3, 2),
            # the same launch's bytes in the general type-4 format (nine A values per entry,
            # no credit for equal or vanishing blocks): what the nine-block kernel streams
            "bytes_per_launch_general_type4": round(gen)}
    if src:
        roof["traffic_source"] = src
    return roof, breakdown


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(prob, threads: int, log):
    """The reference-faithful FP64 oracle (per-sweep patch re-assembly by linear search and
    LU, per-visit coarse LU, OpenMP over the cells of a color and the defect rows) solves
    the system to 1e-10 on `threads` host threads; then the same with cached factors (a
    stronger CPU form, for context) and one V-cycle of the faithful form on one thread."""
    import tempfile
    sys.path.insert(0, str(ROOT / "oracle"))
    import pmg_oracle
    tmp = tempfile.mkdtemp(prefix="pmg_oracle_")
    t = time.time()
    native = pmg_oracle.build_native(tmp)
    flags = pmg_oracle.CXXFLAGS_NATIVE if native else pmg_oracle.CXXFLAGS_PREBUILT
    if native:
        pmg_oracle.use_library(native)
    log(f"oracle build for this host ({flags}): {time.time() - t:.1f}s")
    prm = solver_params(0)
    f = prob.rhs(-1).copy()
    pmg_oracle.set_threads(threads)
    orc = pmg_oracle.Oracle(prm, prob.descs, refactor=1)
    x = np.zeros_like(f)
    t = time.perf_counter()
    it, res, st = orc.fgmres(x, f)
    dt = time.perf_counter() - t
    log(f"cpu oracle (faithful, {threads} threads): {it} iterations in {dt:.2f}s, status {st}")
    # one V-cycle on one thread (bounded 1-core sample) and on all threads
    v = f.copy()
    t = time.perf_counter()
    orc.vcycle(v)
    vc_n = time.perf_counter() - t
    pmg_oracle.set_threads(1)
    t = time.perf_counter()
    orc.vcycle(v)
    vc_1 = time.perf_counter() - t
    orc.close()
    log(f"cpu oracle V-cycle: {vc_n:.2f}s on {threads} threads, {vc_1:.2f}s on 1 thread")
    pmg_oracle.set_threads(threads)
    t = time.perf_counter()
    orc_c = pmg_oracle.Oracle(prm, prob.descs, refactor=0)
    xc = np.zeros_like(f)
    itc, resc, stc = orc_c.fgmres(xc, f)
    dtc = time.perf_counter() - t
    orc_c.close()
    log(f"cpu oracle (cached factors, {threads} threads): {itc} iterations in {dtc:.2f}s")
    base = {"value": round(dt, 3), "unit": "s", "cores": threads, "kind": "port",
            "sample": f"complete FGMRES(30) solve to rtol 1e-10 of the 32^3 Oseen system ({it} iterations, "
                      f"status {st}) by the reference-faithful FP64 oracle (colored Vanka with per-sweep patch "
                      f"re-assembly and LU, per-visit dense coarse LU), OpenMP on {threads} threads",
            "cpu_model": cpu_model(), "compiler_flags": flags, "iterations": it,
            "one_core_vcycle_s": round(vc_1, 3), "threads_vcycle_s": round(vc_n, 3),
            "cached_factor_solve_s": round(dtc, 3), "cached_factor_iterations": itc}
    return base, x, xc


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
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=32, help="cells per edge of the finest level")
    ap.add_argument("--precision", choices=["fp64", "fp32", "bf16"], default="bf16",
                    help="smoother/level-operator arithmetic of the measured solve (FGMRES is FP64)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="host threads of the CPU baseline (default: the CPU share of this process, at most 16)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64-ref", action="store_true", help="skip the FP64 GPU solve timed beside it")
    ap.add_argument("--no-nine-blocks", action="store_true", help="skip the nine-block timings")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched workload's fields")
    ap.add_argument("--no-refine", action="store_true", help="skip the timings with local refinement steps")
    ap.add_argument("--no-autotune", action="store_true", help="keep the built-in launch shapes")
    ap.add_argument("--tuning-file", default=None,
                    help="JSON of launch shapes: loaded when it holds this form's shapes, else tuned and saved there")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--workload", choices=["oseen32", "batched", "reassembly", "channel"], default="oseen32",
                    help="batched: the independent-problems workload; reassembly: one fixed-point step's "
                         "device re-assembly and set-up; channel: the Crank-Nicolson channel time loop "
                         "(instead of the headline solve)")
    ap.add_argument("--channel-steps", type=int, default=20)
    ap.add_argument("--batch-count", type=int, default=64)
    ap.add_argument("--batch-maxit", type=int, default=100)
    ap.add_argument("--no-one-at-a-time", action="store_true",
                    help="batched workload: skip the same problems solved one context at a time")
    ap.add_argument("--no-graph", action="store_true", help="launch every V-cycle kernel by kernel")
    ap.add_argument("--lu-sweeps", action="store_true",
                    help="mixed precision: stored LU factors and triangular sweeps instead of patch inverses")
    ap.add_argument("--nine-blocks", action="store_true",
                    help="stream all nine A blocks even where A11 = A22 = A33 and the coupling blocks vanish")
    args = ap.parse_args()
    global USE_GRAPH, SHARED_BLOCK, PATCH_INVERSE, AUTOTUNE, TUNING_FILE
    USE_GRAPH = 0 if args.no_graph else 1
    AUTOTUNE = not args.no_autotune
    TUNING_FILE = args.tuning_file
    SHARED_BLOCK = 0 if args.nine_blocks else 1
    PATCH_INVERSE = 0 if args.lu_sweeps else 1
    log = (lambda *a: None) if args.quiet else (lambda *a: print("[bench]", *a, file=sys.stderr, flush=True))
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): this build measures one GPU; "
              "the multi-GPU decomposition is described in DESIGN.md and not built.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched":
        print(json.dumps(run_batched(args, log)), flush=True)
        return
    if args.workload == "reassembly":
        print(json.dumps(run_reassembly(args, log)), flush=True)
        return
    if args.workload == "channel":
        print(json.dumps(run_channel(args, log)), flush=True)
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
        ref = run_gpu(prob, 0, max(1, args.steps), min(args.warmup, 1), log, profile=False)
        ref["mg"].close()
    log(f"measured solve: {args.precision}")
    g = run_gpu(prob, prec, args.steps, args.warmup, log, solve_only_steps=max(1, args.steps))
    g["mg"].close()
    ms = 1e3 * float(np.mean(g["times"]))
    roof, breakdown = roofline_from(g["timers"], args.precision, args.n, prob)
    iters = int(g["iters"][-1])
    solve_only_ms = 1e3 * float(np.mean(g["solve_only"]))
    cfg = {"workload": f"oseen3d_{args.n}^3_q2p1disc_fgmres30_vcycle22", "levels": prob.n_levels,
           "dofs": sz["n_dof"], "precision": args.precision, "nu": 1e-2,
           "shared_block": SHARED_BLOCK, "patch_inverse": PATCH_INVERSE,
           "local_solves": ("stored block form (A^-1, A^-1 BT_c, [-S^-1 B_c A^-1 | S^-1]) on 32^3, full inverses on 16^3/8^3"
                            if prec and PATCH_INVERSE and SHARED_BLOCK else
                            "stored full inverses" if prec and PATCH_INVERSE else "LU factors and triangular sweeps"),
           "coarse_solve": ("FP64 block form (A^-1, S^-1 and products), two GEMV phases per visit"
                            if prec and SHARED_BLOCK else "FP64 LU of the reduced system, triangular sweeps"),
           "fgmres_iterations": iters, "final_residual": float(g["hist"][-1]) if len(g["hist"]) else None,
           "solve_with_factors_kept_s": round(solve_only_ms / 1e3, 5),
           "setup_on_critical_path_ms": round(ms - solve_only_ms, 3),
           "patch_setup_ms": round(g["timers"]["patch_lu"]["ms"], 3),
           "coarse_setup_ms_second_stream": round(g["timers"]["coarse_lu"]["ms"], 3),
           "autotuned_launch_shapes": g["shapes"]}
    if ref is not None:
        fp64_ms = 1e3 * float(np.mean(ref["times"]))
        du, dp = rel_err(g["x"], ref["x"], n_u3)
        cfg.update({"fp64_solve_s": round(fp64_ms / 1e3, 5), "fp64_iterations": int(ref["iters"][-1]),
                    "speedup_vs_fp64": round(fp64_ms / ms, 3),
                    "rel_err_vs_fp64": {"velocity": du, "pressure": dp}})
    if prec and PATCH_INVERSE and not args.no_refine:
        # FP32 refinement steps after every local solve with the stored (BF16) data
        cfg["refine_steps"] = {}
        for k in (1, 2, 3):
            r = run_gpu(prob, prec, 2, 1, log, profile=False, refine_steps=k)
            r["mg"].close()
            rk = {"solve_s": round(float(np.mean(r["times"])), 5), "iterations": int(r["iters"][-1])}
            if ref is not None:
                du, dp = rel_err(r["x"], ref["x"], n_u3)
                rk["rel_err_vs_fp64"] = {"velocity": du, "pressure": dp}
            cfg["refine_steps"][str(k)] = rk
            log(f"{k} refinement steps: {1e3 * rk['solve_s']:.1f} ms, {rk['iterations']} iterations")
    if not args.no_nine_blocks and SHARED_BLOCK:
        # all nine A blocks streamed (the general type-4 form of the operator, no credit for
        # A11 = A22 = A33 or the vanishing coupling blocks)
        nine, nine_res = {}, None
        for p_, name in ((0, "fp64"), (prec, args.precision)) if prec else ((0, "fp64"),):
            r = run_gpu(prob, p_, 2, 1, log, shared_block=0, profile=(p_ == prec))
            r["mg"].close()
            nine[name] = 1e3 * float(np.mean(r["times"]))
            log(f"nine blocks, {name}: {nine[name]:.1f} ms")
            if p_ == prec:
                nine_res = r["timers"]["vanka_res_finest"]
        cfg["nine_blocks"] = {f"{k}_solve_s": round(v / 1e3, 5) for k, v in nine.items()}
        if nine_res and nine_res["calls"]:
            # the nine-block residual kernel streams the general type-4 bytes itself
            us = 1e3 * nine_res["ms"] / nine_res["calls"]
            b = nine_res["bytes"] / nine_res["calls"]
            cfg["nine_blocks"]["k_vanka_res_finest"] = {"avg_launch_us": round(us, 2), "bytes_per_launch": round(b),
                                                        "achieved_gbs": round(b / us / 1e3, 1),
                                                        "frac": round(b / us / 1e3 / HBM_PEAK_GBS, 4)}
        if prec and ref is not None:
            cfg["nine_blocks"]["speedup_vs_fp64_nine_blocks"] = round(nine["fp64"] / nine[args.precision], 3)
            cfg["nine_blocks"]["speedup_vs_fp64_nine_blocks_of_the_headline"] = round(nine["fp64"] / ms, 3)
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
    if not args.no_cpu:
        try:
            ncpu = len(os.sched_getaffinity(0))
        except AttributeError:
            ncpu = os.cpu_count() or 1
        threads = args.cpu_threads or min(16, ncpu)
        base, x_orc, x_orc_cached = cpu_baseline(prob, threads, log)
        out["cpu_baseline"] = base
        cfg["speedup_vs_cpu"] = round(base["value"] / (ms / 1e3), 1)
        du, dp = rel_err(g["x"], x_orc_cached, n_u3)
        cfg["rel_err_vs_oracle"] = {"velocity": du, "pressure": dp}
        if ref is not None:
            du, dp = rel_err(ref["x"], x_orc_cached, n_u3)
            cfg["fp64_rel_err_vs_oracle"] = {"velocity": du, "pressure": dp}
        du, dp = rel_err(x_orc, x_orc_cached, n_u3)
        cfg["oracle_faithful_vs_cached"] = {"velocity": du, "pressure": dp}
    else:
        out["cpu_baseline"] = None
    if not args.no_batch:
        specs, probs = batched_problems(args.batch_count, log)
        cfg["batched"] = batched_summary(specs, probs, prec, args.batch_maxit, 1, 1, log,
                                         fp64_ref=not args.no_fp64_ref)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
