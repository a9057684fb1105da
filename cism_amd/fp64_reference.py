"""FP64 reference velocity: the yardstick of every relative-error statement.

The same problem solved on the GPU by the plain FP64 path (Picard + Chron-Gear,
the reference algorithm) with the linear tolerance tightened to 1e-11 and the
Picard loop run until the largest relative velocity change is below velo_tol.
"""
from __future__ import annotations

import time


def reference_solution(ctx, p, lu, lv, chunk=10, max_picard=300, velo_tol=1e-10, uvel=None,
                       vvel=None, log=None):
    """Runs the Picard loop in chunks of `chunk` iterations (a Picard iteration depends only
    on the current velocity, so a chunk restarts exactly where the last one stopped) until
    the velocity change is below velo_tol; returns the last solve's result dict with the
    total Picard and PCG counts.  The context's linear-solver settings are restored to the
    reference's defaults (200 iterations, tolerance 1e-8) afterwards."""
    t0 = time.perf_counter()
    ctx.set_linear_solver(2000, 1e-11)
    u = p.uvel if uvel is None else uvel
    v = p.vvel if vvel is None else vvel
    done, lin = 0, 0
    try:
        while done < max_picard:
            r = ctx.velo_higher_solve_3d(p, lu, lv, uvel=u, vvel=v, maxiter_nonlinear=chunk,
                                         whichresid=0, target_override=velo_tol)
            u, v = r["uvel"], r["vvel"]
            done += r["picard_iters"]
            lin += r["linear_iters"]
            if log:
                log(f"reference: {done} Picard, {lin} PCG its, velocity change "
                    f"{r['final_resid_velo']:.2e}, {time.perf_counter() - t0:.1f} s")
            if r["final_resid_velo"] < velo_tol:
                break
    finally:
        ctx.set_linear_solver(200, 1e-8)
    r["picard_iters"], r["linear_iters"] = done, lin
    return r
