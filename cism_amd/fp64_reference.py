"""FP64 reference velocity: the yardstick of every relative-error statement.

The same problem solved on the GPU with every linear system solved to an FP64
relative residual of 1e-11 (true residual b - A x computed in FP64 from the FP64
matrix), and the Picard loop run until the largest relative velocity change is
below velo_tol.  The linear solves are either the plain FP64 Chron-Gear
(precision 0) or FP32 Chron-Gear inside FP64 iterative refinement
(precision 1); both stop on the same FP64 residual, so both give the Picard
fixed point to FP64 accuracy, and precision 1 gets there faster on large grids.
"""
from __future__ import annotations

import time


def reference_solution(ctx, p, lu, lv, chunk=10, max_picard=400, velo_tol=1e-10, uvel=None,
                       vvel=None, log=None, precision=0, linear_maxiters=2000):
    """Runs the Picard loop in chunks of `chunk` iterations (a Picard iteration depends only
    on the current velocity, so a chunk restarts exactly where the last one stopped) until
    the velocity change is below velo_tol; returns the last solve's result dict with the
    total Picard and PCG counts.  With precision 1 early Picard iterations solve their
    linear systems inexactly (0.05 x the current relative residual, never looser than 1e-2)
    and the last ones to 1e-11.  The context's linear-solver settings are restored to the
    reference's defaults (200 iterations, tolerance 1e-8) afterwards."""
    t0 = time.perf_counter()
    ctx.set_linear_solver(linear_maxiters, 1e-11)
    u = p.uvel if uvel is None else uvel
    v = p.vvel if vvel is None else vvel
    done, lin = 0, 0
    kw = dict(precision=1, adaptive_linear=1, adaptive_gamma=0.05) if precision == 1 else {}
    try:
        while done < max_picard:
            r = ctx.velo_higher_solve_3d(p, lu, lv, uvel=u, vvel=v, maxiter_nonlinear=chunk,
                                         whichresid=0, target_override=velo_tol, **kw)
            u, v = r["uvel"], r["vvel"]
            done += r["picard_iters"]
            lin += r["linear_iters"]
            if log:
                log(f"reference: {done} Picard, {lin} linear its, velocity change "
                    f"{r['final_resid_velo']:.2e}, {time.perf_counter() - t0:.1f} s")
            if r["final_resid_velo"] < velo_tol:
                break
    finally:
        ctx.set_linear_solver(200, 1e-8)
    r["picard_iters"], r["linear_iters"] = done, lin
    r["seconds"] = time.perf_counter() - t0
    return r
