"""The mixed-precision fast path's options on the GPU: every variant must reach the same
1e-6 error against the FP64 reference (the plain FP64 path is the arbiter), and the
per-column preconditioner scale must act as a scale (a constant one leaves Chron-Gear's
iterates unchanged up to FP32 rounding)."""
import numpy as np
import pytest

from cism_amd import configs, gpu as G
from cism_amd.fp64_reference import reference_solution

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    p = configs.grf_sheet(ewn=64, nsn=64, nz=11, dx=1000.0, seed=3)
    print("[fast] grf_sheet 64x64x11 seed 3", flush=True)
    ctx = G.Context.for_problem(p)
    lu, lv = ctx.loads(p)
    ref = reference_solution(ctx, p, lu, lv)
    ctx.set_reference_velocity(ref["uvel"], ref["vvel"])
    yield p, ctx, lu, lv, ref
    ctx.close()


def fast(ctx, p, lu, lv, **kw):
    return ctx.velo_higher_solve_3d(p, lu, lv, precision=1, adaptive_linear=1, stop_rel_err=1e-6,
                                    inner_rtol=1e-3, maxiter_nonlinear=300, **kw)


def test_reference_agrees_with_plain_fp64(setup):
    p, ctx, lu, lv, ref = setup
    r = ctx.velo_higher_solve_3d(p, lu, lv, precision=0, stop_rel_err=1e-6, maxiter_nonlinear=300)
    print(f"[fast] plain FP64: {r['picard_iters']} Picard, err {r['final_rel_err']:.2e}")
    assert 0 <= r["final_rel_err"] <= 1e-6
    # the refinement-certified reference converges to the same fixed point
    ref1 = reference_solution(ctx, p, lu, lv, precision=1)
    x0 = np.concatenate([ref["uvel"].ravel(), ref["vvel"].ravel()])
    x1 = np.concatenate([ref1["uvel"].ravel(), ref1["vvel"].ravel()])
    ctx.set_reference_velocity(ref["uvel"], ref["vvel"])
    assert np.linalg.norm(x1 - x0) <= 1e-8 * np.linalg.norm(x0)


@pytest.mark.parametrize("m", [0, 2, 4])
def test_anderson_reaches_target(setup, m):
    p, ctx, lu, lv, ref = setup
    r = fast(ctx, p, lu, lv, anderson_m=m)
    print(f"[fast] anderson {m}: {r['picard_iters']} Picard, {r['fp32_matvecs']} FP32 matvecs, "
          f"err {r['final_rel_err']:.2e}")
    assert 0 <= r["final_rel_err"] <= 1e-6


@pytest.mark.parametrize("m", [0, 3])
def test_bench_settings_on_a_sheet_with_inactive_columns(m):
    """the bench's fast-path settings (inner FP32 tolerance 5e-2, Anderson depth 3) on an ice
    sheet whose margin leaves most columns inactive: the Anderson passes cover active columns
    only, and the solve still reaches 1e-6 against the FP64 reference"""
    p = configs.ice_sheet(ewn=36, nsn=42, nz=6, dx=33333.0, seed=7)
    assert (p.active_vertex == 0).sum() > 0
    with G.Context.for_problem(p) as ctx:
        lu, lv = ctx.loads(p)
        ref = reference_solution(ctx, p, lu, lv)
        ctx.set_reference_velocity(ref["uvel"], ref["vvel"])
        r = ctx.velo_higher_solve_3d(p, lu, lv, precision=1, adaptive_linear=1, stop_rel_err=1e-6,
                                     inner_rtol=5e-2, maxiter_nonlinear=300, anderson_m=m)
    print(f"[fast] sheet, anderson {m}: {r['picard_iters']} Picard, {r['fp32_matvecs']} FP32 "
          f"matvecs, err {r['final_rel_err']:.2e}")
    assert 0 <= r["final_rel_err"] <= 1e-6
    # the solve leaves the velocity of inactive owned columns as it found it (zero)
    inact = np.zeros((p.nx - 1, p.ny - 1), dtype=bool)
    inact[2:-2, 2:-2] = p.active_vertex[2:-2, 2:-2] == 0
    assert np.all(r["uvel"][:, inact] == 0.0) and np.all(r["vvel"][:, inact] == 0.0)


def test_constant_precond_scale_is_invisible(setup):
    p, ctx, lu, lv, ref = setup
    base = fast(ctx, p, lu, lv)
    s = np.full((p.nx - 1, p.ny - 1), 4.0, order="F")
    ctx.set_precond_scale(s, s)
    try:
        sc = fast(ctx, p, lu, lv)
    finally:
        ctx.set_precond_scale(None, None)
    print(f"[fast] scale 1: {base['fp32_matvecs']} FP32, scale 4: {sc['fp32_matvecs']} FP32")
    assert sc["final_rel_err"] <= 1e-6
    assert sc["picard_iters"] == base["picard_iters"]
    assert abs(sc["fp32_matvecs"] - base["fp32_matvecs"]) <= 0.05 * base["fp32_matvecs"] + 5
    bad = np.full((p.nx - 1, p.ny - 1), -1.0, order="F")
    with pytest.raises(G.CismError):
        ctx.set_precond_scale(bad, bad)


def test_column_scale_target_shape(setup):
    p, ctx, lu, lv, ref = setup
    ctx.velo_higher_solve_3d(p, lu, lv, precision=0)
    su, sv = ctx.column_scale_target()
    av = p.active_vertex[2:-2, 2:-2] > 0
    for s in (su, sv):
        o = s[2:-2, 2:-2]
        assert np.all(o[av] >= 1 / 16) and np.all(o[av] <= 16)
        assert np.all(o[~av] == 1.0)


def test_stalled_refinement_falls_back_to_fp64(setup):
    """an FP64 linear tolerance beyond the FP32 operator's reach (1e-20): the refinement
    stalls at its rounding floor, that Picard iteration is redone by the FP64 solver (which
    then serves the rest of the solve), and the solve still reaches 1e-6 of the FP64
    reference"""
    p, ctx, lu, lv, ref = setup
    ctx.set_linear_solver(200, 1e-20)
    try:
        r = ctx.velo_higher_solve_3d(p, lu, lv, precision=1, adaptive_linear=0, inner_rtol=5e-2,
                                     stop_rel_err=1e-6, maxiter_nonlinear=300)
    finally:
        ctx.set_linear_solver(200, 1e-8)
    print(f"[fast] unreachable tolerance: {r['picard_iters']} Picard, {r['fallbacks']} fallback(s), "
          f"{r['fp32_matvecs']} FP32 matvecs, err {r['final_rel_err']:.2e}")
    assert r["fallbacks"] == 1
    assert r["converged"] and 0 <= r["final_rel_err"] <= 1e-6


def test_worse_than_zero_guess_is_rejected(setup):
    """an initial guess whose Picard residual exceeds the zero guess's is dropped: the solve
    is then the zero-guess solve, bit for bit"""
    p, ctx, lu, lv, ref = setup
    rng = np.random.default_rng(0)
    bu = np.asfortranarray(1e5 * rng.standard_normal(p.uvel.shape))
    bv = np.asfortranarray(1e5 * rng.standard_normal(p.vvel.shape))
    bad = fast(ctx, p, lu, lv, uvel=bu, vvel=bv, anderson_m=3)
    zero = fast(ctx, p, lu, lv, anderson_m=3)
    print(f"[fast] bad guess: rejected {bad['guess_rejected']}, {bad['picard_iters']} Picard, "
          f"err {bad['final_rel_err']:.2e}")
    assert bad["guess_rejected"] == 1 and zero["guess_rejected"] == 0
    assert 0 <= bad["final_rel_err"] <= 1e-6
    assert bad["picard_iters"] == zero["picard_iters"]
    assert np.array_equal(bad["uvel"], zero["uvel"]) and np.array_equal(bad["vvel"], zero["vvel"])


def test_skipping_the_final_residual_keeps_the_iterates(setup):
    """skip_final_residual only leaves out the FP64 residual that would confirm a linear
    solve its inner estimate already meets: the solve reaches the same 1e-6 with fewer
    FP64 matvecs"""
    p, ctx, lu, lv, ref = setup
    kw = dict(anderson_m=4, coarse=1)
    r0 = fast(ctx, p, lu, lv, skip_final_residual=0, **kw)
    r1 = fast(ctx, p, lu, lv, skip_final_residual=1, **kw)
    print(f"[fast] confirmed: {r0['picard_iters']} Picard, {r0['fp64_matvecs']} FP64; skipped: "
          f"{r1['picard_iters']} Picard, {r1['fp64_matvecs']} FP64, err {r1['final_rel_err']:.2e}")
    for r in (r0, r1):
        assert r["converged"] and 0 <= r["final_rel_err"] <= 1e-6 and r["fallbacks"] == 0
    assert r1["fp64_matvecs"] < r0["fp64_matvecs"]


def test_fp64_residual_from_element_matrices(setup):
    """With skip_final_residual 1 or 2 the mixed path stores no FP64 matrix; a refinement
    step's FP64 residual then comes from the element matrices (the gather's own sums, in
    the stored matrix's operation order).  Mode 2 confirms every step as mode 0 does, so
    the two take the same steps and must agree bit for bit."""
    p, ctx, lu, lv, ref = setup
    kw = dict(inner_rtol=0.3, coarse=1)
    r0 = ctx.velo_higher_solve_3d(p, lu, lv, precision=1, adaptive_linear=1, stop_rel_err=1e-6,
                                  maxiter_nonlinear=300, skip_final_residual=0, **kw)
    r1 = ctx.velo_higher_solve_3d(p, lu, lv, precision=1, adaptive_linear=1, stop_rel_err=1e-6,
                                  maxiter_nonlinear=300, skip_final_residual=2, **kw)
    print(f"[fast] stored FP64 matrix: {r0['picard_iters']} Picard, {r0['fp64_matvecs']} FP64; "
          f"element residuals: {r1['picard_iters']} Picard, {r1['fp64_matvecs']} FP64")
    assert r0["fp64_matvecs"] > r0["picard_iters"]   # refinement residuals did run
    for key in ("picard_iters", "fp32_matvecs", "fp64_matvecs"):
        assert r1[key] == r0[key]
    assert np.array_equal(r1["uvel"], r0["uvel"]) and np.array_equal(r1["vvel"], r0["vvel"])
    # the FP64 matrix is gone: the kernels that need it say so
    with pytest.raises(G.CismError):
        ctx.column_scale_target()
