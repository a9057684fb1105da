#!/usr/bin/env python3
"""Benchmark: seconds per nonlinear Blatter-Pattyn velocity solve on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c5|c3|c2|c1] [--guess learned|none]

A step is one complete nonlinear (Picard) velocity solve, run until the relative
L2 velocity error against the FP64 reference solution is <= 1e-6.  The FP64
reference is the same problem with every linear system solved to an FP64
relative residual of 1e-11 and the Picard loop run until the velocity changes
by less than 1e-10 (cism_amd/fp64_reference.py; computed once, before the
timed region, not timed).

Every leg calls cism_glissade_velo_higher_solve: the geometry (thck, usrf, topg,
flwa, beta_external) goes in, the masks, staggered fields, surface gradient, beta,
loads, learned guess and the Picard loop run on the device, the velocity comes out.
Timed: device seconds from resident inputs to the solution on the device (`value`);
the PCIe-inclusive wall time of the same call is reported beside it.
  headline  learned initial guess (features, network, decode on the GPU) + fast
            path: FP32 matrix + FP32 Chron-Gear inner solves inside FP64
            iterative refinement, inexact-Picard linear tolerances and
            Anderson-accelerated Picard iterations
  beside    the fast path from a zero guess (what the network gains); the fast
            path under the reference's own stopping rule (pre-solve L2 < 1e-4)
  baseline  plain FP64 path: the reference algorithm (Chron-Gear, SIA
            preconditioner, tolerance 1e-8, 200 iterations), same stopping rule

Printed: one JSON line.  `roofline` covers the dominant kernel of the headline
solve (the FP32 27-point matvec), measured with HIP events on the context's
stream; `cpu_baseline` times the CPU oracle (plain C restatement, 1 core) on a
bounded sample and extrapolates with the GPU's iteration counts.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SYM_SLOTS = 14
FP64_VALU_PEAK_TFLOPS = 78.6
STOP_REL_ERR = 1.0e-6


def workload(name):
    """(problem, description, surrogate weight family or None)"""
    from cism_amd import configs
    if name == "c1":
        return configs.dome(), "dome 31x31x10, dx 2 km, no-slip, diag precond", None
    if name == "c2":
        return (configs.ismip_hom_a(ewn=160, nsn=160, nz=21, L=80000.0),
                "ISMIP-HOM A 160x160x21 periodic, L=80 km, frozen bed, Chron-Gear + SIA precond",
                None)
    if name == "c3":
        return (configs.grf_sheet(ewn=512, nsn=512, nz=11, dx=1000.0, seed=1),
                "GRF sheet 512x512x11, dx 1 km, lognormal beta, Chron-Gear + SIA precond", "grf")
    if name == "c5":
        return (configs.ice_sheet(ewn=1200, nsn=1400, nz=11, dx=1000.0, seed=5),
                "ice sheet 1200x1400x11, dx 1 km, grounded (13 floating cells; margins at most "
                "300 m thick, lateral ocean pressure where they meet the sea), cold ice A = 1e-17, "
                "multiscale thickness, bed relief, lognormal beta, Chron-Gear + SIA precond", "sheet")
    raise SystemExit(f"unknown workload {name}")


def owned_active_nodes(p):
    av = p.active_vertex[2:p.nx - 2, 2:p.ny - 2] if p.periodic else p.active_vertex[1:p.nx - 2, 1:p.ny - 2]
    return int(av.sum()) * p.nz


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(p, lu, lv, fp64_stats, all_cores=False):
    """Times the CPU oracle on a bounded sample: one assembly and a short Chron-Gear run on
    the assembled system; extrapolates to a full solve with the GPU FP64 run's counts.
    all_cores: the OpenMP build of the same restatement on every core OpenMP is given."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.set_all_cores(all_cores)
    try:
        return _cpu_sample(O, p, lu, lv, fp64_stats, all_cores)
    finally:
        O.set_all_cores(False)


def _cpu_sample(O, p, lu, lv, fp64_stats, all_cores):
    t0 = time.perf_counter()
    st, A, _ = O.assemble(p, p.uvel, p.vvel)
    t_asm = time.perf_counter() - t0
    bu, bv = lu.copy(order="F"), lv.copy(order="F")
    O.dirichlet(p, A, bu, bv, p.uvel, p.vvel)
    for a in A:
        O.halo(p, a)
    n_it = 40 if p.nnodes < 2_000_000 else 8     # keeps the sample near 10-30 s of CPU work
    xu, xv = p.uvel.copy(order="F"), p.vvel.copy(order="F")
    t0 = time.perf_counter()
    O.pcg(p, A, bu, bv, xu, xv, p.whichprecond, chrongear=True, maxiters=n_it, tolerance=1e-30)
    t_pcg = (time.perf_counter() - t0) / n_it
    n_pic, n_lin = fp64_stats["picard_iters"], fp64_stats["linear_iters"]
    est = n_pic * t_asm + n_lin * t_pcg
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)) if all_cores else 1
    what = f"OpenMP, {cores} threads" if all_cores else "1 thread"
    return {"value": round(est, 3), "unit": "s/solve", "cores": cores, "kind": "port",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "sample": (f"oracle (C, -O3, {what}): 1 assembly {t_asm:.2f} s + {n_it} Chron-Gear "
                       f"iterations {t_pcg * 1e3:.1f} ms/it, extrapolated to the GPU FP64 run's "
                       f"{n_pic} Picard / {n_lin} PCG iterations")}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c5")
    ap.add_argument("--guess", default="learned", choices=["learned", "none"])
    ap.add_argument("--weights", default=None,
                    help="network weights (default: the workload family's file, if trained)")
    ap.add_argument("--anderson", type=int, default=4,
                    help="Anderson acceleration depth of the fast path's Picard loop (0 off)")
    ap.add_argument("--inner-rtol", type=float, default=1e-2,
                    help="relative residual of each FP32 inner solve of the refinement")
    ap.add_argument("--coarse", type=int, default=1, choices=[0, 1],
                    help="aggregation coarse correction in the FP32 inner solves' preconditioner")
    ap.add_argument("--skip-final", type=int, default=1, choices=[0, 1],
                    help="end a linear solve on its FP32 inner solve's estimate when that meets half "
                         "the tolerance (cism_solve_opts.skip_final_residual)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="time only the FP32 matvec and the copy calibration kernel (for counters)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-fast", type=int, default=0, metavar="N",
                    help="for kernel profiles: run only the learned-guess fast solve N times, "
                         "for a fixed number of Picard iterations (--picard), without the "
                         "reference solve and the other legs")
    ap.add_argument("--picard", type=int, default=14,
                    help="Picard iterations of --profile-fast (the headline solve takes 14)")
    ap.add_argument("--baseline-steps", type=int, default=1)
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (this build runs the "
              "whole domain on one MI355X; multi-GPU sharding is a design note only)",
              file=sys.stderr)
        return 2

    from cism_amd import gpu as G
    from cism_amd import surrogate as SG
    from cism_amd.fp64_reference import reference_solution

    def log(m):
        print(f"[bench] {m}", file=sys.stderr, flush=True)

    p, desc, family = workload(a.workload)
    log(f"workload {a.workload}: {desc}; inputs sha256 {inputs_hash(p)}")
    ctx = G.Context.for_problem(p)
    lu, lv = ctx.loads(p)
    nodes = owned_active_nodes(p)

    # both timed paths run until their error against the reference is <= 1e-6; the Picard
    # cap only guards against divergence
    cap = 500

    if a.kernels_only:
        # one short mixed-precision Picard iteration leaves an FP32 matrix; then only the
        # kernels under study run
        ctx.set_linear_solver(5, 1e-8)
        ctx.velo_higher_solve_3d(p, lu, lv, precision=1, maxiter_nonlinear=1)
        ctx.set_linear_solver(200, 1e-8)
        ms32 = ctx.time_kernel(1, a.steps)
        msu = ctx.time_kernel(4, a.steps)
        msc = ctx.time_kernel(3, a.steps)
        ms_asm = ctx.time_kernel(2, max(1, a.steps // 10))
        ms64f = ctx.time_kernel(5, a.steps)
        ms64 = ctx.time_kernel(0, a.steps)
        ms64o = ctx.time_kernel(6, a.steps)
        print(json.dumps({"workload": a.workload, "owned_active_nodes": nodes,
                          "fp32_matvec_ms": ms32, "fp32_update_ms": msu, "copy_1GiB_ms": msc,
                          "assembly_ms": ms_asm, "fp64_matvec_ms": ms64, "fp64_matvec_full_ms": ms64f,
                          "fp64_matvec_sym_other_order_ms": ms64o, "launches": a.steps}))
        ctx.close()
        return 0

    if a.profile_fast:
        weights = a.weights or (SG.default_weights(family) if family else None)
        guess = 0
        if weights is not None and os.path.exists(weights):
            params, meta = SG.load_weights(weights)
            ctx.learned_guess_load(params, int(meta.get("nz", p.nz)),
                                   1 if meta.get("target") == "log1p" else 0, float(meta["uscale"]))
            guess = 1
        bp = G.BasalPhysics(p.whichbabc) if p.whichbabc != 6 else None
        for s in range(a.profile_fast):
            r = ctx.glissade_velo_higher_solve(p, bp=bp, beta_external=p.extra.get("beta_external"),
                                               initial_guess=guess, precision=1, adaptive_linear=1,
                                               inner_rtol=a.inner_rtol, maxiter_nonlinear=a.picard,
                                               anderson_m=a.anderson, coarse=a.coarse,
                                               skip_final_residual=a.skip_final, efvs=False)
            log(f"fast solve {s}: {r['seconds_device']:.3f} s device, {r['picard_iters']} Picard, "
                f"{r['fp32_matvecs']} FP32 + {r['fp64_matvecs']} FP64 matvecs")
        ctx.close()
        return 0

    # large grids: FP64-residual-certified refinement (same fixed point, much faster)
    ref_prec = 1 if p.nnodes > 4_000_000 else 0
    ref = reference_solution(ctx, p, lu, lv, log=log, precision=ref_prec, velo_tol=1e-10)
    ctx.set_reference_velocity(ref["uvel"], ref["vvel"])

    weights = a.weights or (SG.default_weights(family) if family else None)
    use_guess = a.guess == "learned" and weights is not None and os.path.exists(weights)
    if use_guess:
        params, meta = SG.load_weights(weights)
        ctx.learned_guess_load(params, int(meta.get("nz", p.nz)),
                               1 if meta.get("target") == "log1p" else 0, float(meta["uscale"]))
    bp = G.BasalPhysics(p.whichbabc) if p.whichbabc != 6 else None
    bext = p.extra.get("beta_external")

    # the model's velocity arrays, kept across solves as a Fortran caller keeps its own
    vel_out = tuple(np.zeros((p.nz, p.nx - 1, p.ny - 1), order="F") for _ in range(2))

    def solve(guess, stop=STOP_REL_ERR, precision=1):
        """one solve from the geometry (cism_glissade_velo_higher_solve): the cell fields go
        in, every input of the Picard loop is computed on the device, the velocity comes out.
        seconds = device time from resident inputs to the solution on the device (inputs,
        loads, learned guess, Picard loop); seconds_wall adds the host transfers (with a
        learned guess the solution lands in the persistent velocity arrays; the zero guess
        uploads p's zero velocity)."""
        kw = dict(precision=1, adaptive_linear=1, inner_rtol=a.inner_rtol, anderson_m=a.anderson,
                  coarse=a.coarse, skip_final_residual=a.skip_final) if precision == 1 else dict(precision=0)
        if guess != 0:
            kw["out"] = vel_out
        t0 = time.perf_counter()
        r = ctx.glissade_velo_higher_solve(p, bp=bp, beta_external=bext, initial_guess=guess,
                                           stop_rel_err=stop, maxiter_nonlinear=cap, efvs=False, **kw)
        r["seconds_wall"] = time.perf_counter() - t0
        r["seconds"] = r["seconds_device"]
        r["inference_s"] = ctx.learned_guess_last_ms() * 1e-3 if guess == 1 else 0.0
        return r

    def headline():
        """the timed step: learned guess + mixed-precision solve, every step on the device"""
        return solve(1 if use_guess else 0)

    for _ in range(a.warmup):
        headline()
    runs = []
    for s in range(a.steps):
        r = headline()
        runs.append(r)
        log(f"step {s}: {r['seconds']:.3f} s device, {r['seconds_wall']:.3f} s wall (network forward "
            f"{r['inference_s'] * 1e3:.1f} ms), {r['picard_iters']} Picard, {r['fp32_matvecs']} FP32 + "
            f"{r['fp64_matvecs']} FP64 matvecs, err {r['final_rel_err']:.2e}")
    fast_s = float(np.mean([r["seconds"] for r in runs]))
    wall_s = float(np.mean([r["seconds_wall"] for r in runs]))
    stats = runs[-1]

    nog = solve(0) if use_guess else None
    if nog is not None:
        log(f"mixed precision, zero guess: {nog['seconds']:.3f} s, {nog['picard_iters']} Picard, "
            f"err {nog['final_rel_err']:.2e}")
    # the fast path as deployed: no known answer, the reference's Picard test ends the loop
    dep = solve(1 if use_guess else 0, stop=0.0)
    log(f"fast path under the reference's stopping rule: {dep['seconds']:.3f} s, {dep['picard_iters']} "
        f"Picard, converged {dep['converged']}, err {dep['final_rel_err']:.2e}")
    base = solve(0, precision=0)
    log(f"plain FP64: {base['seconds']:.3f} s, {base['picard_iters']} Picard, "
        f"{base['linear_iters']} PCG its, err {base['final_rel_err']:.2e}")

    # kernels: the matrices left by the last solves are current
    solve(1 if use_guess else 0)   # leaves the FP32 matrix of a fast solve
    ms32 = ctx.time_kernel(1, 50)
    ms_asm = ctx.time_kernel(2, 5)   # (also leaves an FP64 matrix: the fast solve stores none)
    ms64f = ctx.time_kernel(5, 50)   # FP64 matvec on the full storage
    ms64 = ctx.time_kernel(0, 50)    # as the FP64 solver runs it: on the symmetric copy
    # algorithmic bytes of one matvec: 108 coefficients + 2 x + 2 y per owned node (448 B
    # in FP32); the FP32 matrix is kept in symmetric storage (14 of 27 slots, 248 B per
    # node with x, y, r), which lowers the bytes moved but not this count
    bytes32 = nodes * (108 * 4 + 2 * 4 + 2 * 4)
    bytes64 = nodes * (108 * 8 + 2 * 8 + 2 * 8)
    gbs32 = bytes32 / (ms32 * 1e-3) / 1e9
    gbs64 = bytes64 / (ms64 * 1e-3) / 1e9
    n_el = int(p.active_cell.sum()) * (p.nz - 1)
    asm_tflops = n_el * FLOP_PER_ELEMENT / (ms_asm * 1e-3) / 1e12
    traffic = measured_traffic(a.workload)

    cpu = cpu_all = None
    if not a.no_cpu_baseline:
        log("cpu baseline (oracle, 1 thread, bounded sample)")
        cpu = cpu_baseline(p, lu, lv, base)
        log("cpu baseline (oracle, all cores, bounded sample)")
        cpu_all = cpu_baseline(p, lu, lv, base, all_cores=True)

    depth = depth_solvers_report(log, with_cpu=not a.no_cpu_baseline)

    line = {
        "metric": "s/velocity solve to 1e-6 err: learned+FP32-PCG vs FP64 PCG; matvec HBM GB/s",
        "value": round(fast_s, 4),
        "unit": "s/solve",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(fast_s * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 inner / fp64 refinement" + (" / bf16 surrogate" if use_guess else ""),
        "data": "synthetic",
        "config": {"workload": desc, "grid": [p.ewn, p.nsn, p.nz],
                   "owned_active_nodes": nodes, "stop_rel_err": STOP_REL_ERR,
                   "floating_cells": int(p.floating_mask.sum()), "A_Pa3yr": float(np.max(p.flwa)),
                   "learned_guess": os.path.relpath(weights, ROOT) if use_guess else None,
                   "entry_point": "cism_glissade_velo_higher_solve (geometry in, velocity out)",
                   "parallelism": "single GPU, whole domain resident"},
        "fast_path": {"seconds": round(fast_s, 4), "anderson_m": a.anderson, "inner_rtol": a.inner_rtol,
                      "coarse_correction": bool(a.coarse), "skip_final_residual": bool(a.skip_final),
                      "picard_iters": stats["picard_iters"],
                      "fp32_matvecs": stats["fp32_matvecs"], "fp64_matvecs": stats["fp64_matvecs"],
                      "rel_err": stats["final_rel_err"], "learned_guess": use_guess,
                      "guess_rejected": stats["guess_rejected"], "fallbacks": stats["fallbacks"],
                      "network_forward_s": round(stats["inference_s"], 4),
                      "seconds_wall_pcie_inclusive": round(wall_s, 4)},
        "fast_path_reference_stopping_rule": {
            "seconds": round(dep["seconds"], 4), "seconds_wall": round(dep["seconds_wall"], 4),
            "picard_iters": dep["picard_iters"], "converged": bool(dep["converged"]),
            "final_L2_norm": dep["final_L2_norm"], "rel_err": dep["final_rel_err"]},
        "fast_zero_guess": None if nog is None else {
            "seconds": round(nog["seconds"], 4), "seconds_wall": round(nog["seconds_wall"], 4),
            "picard_iters": nog["picard_iters"],
            "fp32_matvecs": nog["fp32_matvecs"], "fp64_matvecs": nog["fp64_matvecs"],
            "rel_err": nog["final_rel_err"]},
        "plain_fp64": {"seconds": round(base["seconds"], 4), "seconds_wall": round(base["seconds_wall"], 4),
                       "picard_iters": base["picard_iters"],
                       "pcg_iters": base["linear_iters"], "rel_err": base["final_rel_err"]},
        "speedup_vs_plain_fp64": round(base["seconds"] / fast_s, 3),
        "reference_fp64": {"picard_iters": ref["picard_iters"], "linear_iters": ref["linear_iters"],
                           "linear_solver": ["plain FP64 Chron-Gear", "FP32 Chron-Gear in FP64 "
                                             "refinement"][ref_prec],
                           "linear_tol_fp64_residual": 1e-11, "velocity_change_tol": 1e-10,
                           "seconds": round(ref["seconds"], 1)},
        "matvec": {"fp32_ms": round(ms32, 4), "fp32_GBps": round(gbs32, 1),
                   "fp64_ms": round(ms64, 4), "fp64_GBps": round(gbs64, 1),
                   "fp64_full_storage_ms": round(ms64f, 4),
                   "fp64_full_storage_GBps": round(bytes64 / (ms64f * 1e-3) / 1e9, 1),
                   "bytes_per_node_fp32": 448, "bytes_per_node_fp64": 896,
                   "fp32_stored_bytes_per_node": SYM_SLOTS * 16 + 24,
                   "fp32_stored_GBps": round(nodes * (SYM_SLOTS * 16 + 24) / (ms32 * 1e-3) / 1e9, 1)},
        "assembly": {"ms": round(ms_asm, 3), "TFLOPs": round(asm_tflops, 3),
                     "flop_per_element": FLOP_PER_ELEMENT, "elements": n_el,
                     "frac_fp64_peak": round(asm_tflops / FP64_VALU_PEAK_TFLOPS, 4)},
        "roofline": {"bound": "hbm", "achieved": round(gbs32, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs32 / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_matvec_dots<float,2,MatSym32>",
                     "algorithmic_bytes_per_launch": bytes32},
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
        "solvers_2d": depth,
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


def depth_solvers_report(log, with_cpu=True):
    """The 2D solvers (SSA, L1L2, DIVA) on the GRF sheet: their plain FP64 Picard loops on the
    GPU at 512 x 512 (one-level context), and at 128 x 128 beside the CPU oracle's loop on the
    same problem.  Device seconds from resident inputs, as the headline."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from cism_amd import configs
    from cism_amd import gpu as G
    out = {}
    for n in (512, 128):
        p = configs.grf_sheet(ewn=n, nsn=n, nz=11, dx=1000.0, seed=1)
        lu, lv = O.loads(p)
        with G.Context.for_problem_2d(p) as ctx:
            for name, approx in (("ssa", 1), ("l1l2", 3), ("diva", 4)):
                if approx == 1:
                    r = ctx.velo_higher_solve_2d(p, lu, lv)
                else:
                    r = ctx.velo_higher_solve_depth(p, lu, lv, approx)
                e = {"gpu_seconds": round(r["seconds_device"], 4), "picard_iters": r["picard_iters"],
                     "pcg_iters": r["linear_iters"], "converged": bool(r["converged"])}
                if n == 128 and with_cpu:
                    t0 = time.perf_counter()
                    ro = O.picard_2d(p, lu, lv) if approx == 1 else O.picard_depth(p, lu, lv, approx)
                    e["cpu_oracle_seconds_1_thread"] = round(time.perf_counter() - t0, 3)
                    e["cpu_picard_iters"] = ro["picard_iters"]
                out[f"{name}_{n}x{n}"] = e
                log(f"2D {name} {n}x{n}: {e}")
    return out


def inputs_hash(p):
    """short sha256 of the generated fields: the same value on every host means the same
    problem, bit for bit"""
    import hashlib
    h = hashlib.sha256()
    for a in (p.thck, p.usrf, p.topg, p.beta_internal, p.flwafact):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()[:16]


def measured_traffic(workload_name):
    """HBM bytes per launch of the FP32 matvec from the committed rocprofv3 --pmc summary of
    this workload (FETCH_SIZE + WRITE_SIZE, each scaled by the copy-kernel calibration), or
    None when no such summary exists."""
    path = os.path.join(ROOT, "profiles", f"pmc_k_matvec_dots_fp32_{workload_name}.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return int(float(line.split(":", 1)[1]))
    return None


# FP64 operations of one hexahedral element in the assembly kernels, counted from the code.
# Per quadrature point: Jacobian 9 x 8 x 2 = 144; cofactors 27 + determinant 5 = 32; inverse 9;
# basis derivatives 8 x 3 x 5 = 120; strain rates 8 x 6 x 2 = 96; Glen viscosity 17 + pow (1);
# quadrature factor 3  -> 422, x 8 points = 3376.  Element entries (Kuu, Kvv upper triangles
# 72 x 8 flop, Kuv 64 x 6 flop, per point): (576 + 384) x 8 = 7680.  Element efvs mean 8.
# Total 11064 (the reference forms all 256 entries: ~16.6k; the symmetric half is exact here).
FLOP_PER_ELEMENT = 11064

if __name__ == "__main__":
    sys.exit(main())
