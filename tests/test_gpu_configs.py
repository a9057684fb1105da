"""The benchmark configurations at their full sizes on the GPU.

  ISMIP-HOM A 160 x 160 x 21 (L = 80 km): one Picard iteration (assembly, Dirichlet,
      halo, symmetry repair, 200 Chron-Gear iterations with the column preconditioner)
      against the oracle; the full Picard solve, whose velocity the oracle's own
      residual then finds converged by the reference's test.
  GRF sheet 512 x 512 x 11: one Picard iteration with 20 Chron-Gear iterations against
      the oracle; the mixed-precision path from zero and from the learned guess (held-out
      seed) to 1e-6 of the FP64 reference, the guess saving Picard iterations.
  Ice sheet 1200 x 1400 x 11: the headline solve (learned guess + mixed precision with the
      coarse correction) and the plain FP64 path, both to 1e-6 of the FP64 reference.
"""
import os

import numpy as np
import pytest

import oracle as O
from cism_amd import configs, gpu as G, surrogate as SG
from cism_amd.fp64_reference import reference_solution
from helpers import rel_l2

pytestmark = pytest.mark.gpu


def _x(r):
    return np.concatenate([r["uvel"].ravel(), r["vvel"].ravel()])


def _oracle_L2(p, u, v, lu, lv):
    """the oracle's pre-solve Picard residual |A(x) x - b| at the velocity (u, v)"""
    st, A, _ = O.assemble(p, u, v)
    assert st == 0
    if p.whichbabc != configs.HO_BABC_NO_SLIP:
        O.basal_local(p, A)
    bu, bv = lu.copy(order="F"), lv.copy(order="F")
    O.dirichlet(p, A, bu, bv, u, v)
    for a in A:
        O.halo(p, a)
    O.halo(p, bu)
    O.halo(p, bv)
    O.check_symmetry(p, A)
    _, _, L2, L2r = O.residual_vector(p, A, bu, bv, u, v)
    return L2, L2r


@pytest.fixture(scope="module")
def c2():
    p = configs.ismip_hom_a(ewn=160, nsn=160, nz=21, L=80000.0)
    print(f"\n[configs] ISMIP-HOM A {p.ewn}x{p.nsn}x{p.nz}", flush=True)
    return p, O.loads(p)


def test_c2_first_picard_iteration_matches_the_oracle(c2):
    p, (lu, lv) = c2
    ro = O.picard(p, lu, lv, maxiter=1)
    with G.Context.for_problem(p) as ctx:
        rg = ctx.velo_higher_solve_3d(p, lu, lv, maxiter_nonlinear=1)
    e = rel_l2(_x(rg), _x(ro))
    print(f"  one Picard iteration: oracle {ro['linear_iters']} PCG, GPU {rg['linear_iters']} PCG, "
          f"rel L2 diff {e:.2e}")
    assert abs(rg["linear_iters"] - ro["linear_iters"]) <= 5
    assert e < 1e-8
    assert np.max(np.abs(rg["efvs"] - ro["efvs"])) <= 1e-12 * np.max(np.abs(ro["efvs"]))


def test_c2_picard_solve_is_converged_by_the_oracles_test(c2):
    p, (lu, lv) = c2
    with G.Context.for_problem(p) as ctx:
        rg = ctx.velo_higher_solve_3d(p, lu, lv)
    L2, L2r = _oracle_L2(p, rg["uvel"], rg["vvel"], lu, lv)
    print(f"  GPU Picard {rg['picard_iters']} iterations, {rg['linear_iters']} PCG, "
          f"{rg['seconds_device']:.3f} s device; oracle residual at its velocity {L2:.2e} ({L2r:.2e} relative)")
    assert rg["converged"] and rg["final_L2_norm"] < 1e-4
    assert L2 < 1e-4


@pytest.fixture(scope="module")
def c3():
    p = configs.grf_sheet(ewn=512, nsn=512, nz=11, dx=1000.0, seed=3101)
    print(f"\n[configs] GRF sheet {p.ewn}x{p.nsn}x{p.nz} seed 3101", flush=True)
    return p, O.loads(p)


def test_c3_first_picard_iteration_matches_the_oracle(c3):
    p, (lu, lv) = c3
    ro = O.picard(p, lu, lv, maxiter=1, linear_maxiters=20)
    with G.Context.for_problem(p) as ctx:
        ctx.set_linear_solver(20, 1e-8)
        rg = ctx.velo_higher_solve_3d(p, lu, lv, maxiter_nonlinear=1)
    e = rel_l2(_x(rg), _x(ro))
    print(f"  one Picard iteration, 20 Chron-Gear iterations: rel L2 diff {e:.2e}")
    assert rg["linear_iters"] == ro["linear_iters"] == 20
    assert e < 1e-10


def test_c3_mixed_precision_and_learned_guess(c3):
    p, (lu, lv) = c3
    w = SG.default_weights("grf")
    with G.Context.for_problem(p) as ctx:
        ref = reference_solution(ctx, p, lu, lv)
        ctx.set_reference_velocity(ref["uvel"], ref["vvel"])
        kw = dict(precision=1, adaptive_linear=1, inner_rtol=5e-2, anderson_m=3, stop_rel_err=1e-6,
                  maxiter_nonlinear=300)
        bp = G.BasalPhysics(configs.HO_BABC_BETA_EXTERNAL)
        bext = p.extra["beta_external"]
        rz = ctx.glissade_velo_higher_solve(p, bp=bp, beta_external=bext, initial_guess=0, **kw)
        runs = [("zero", rz)]
        if os.path.exists(w):
            params, meta = SG.load_weights(w)
            ctx.learned_guess_load(params, meta["nz"], 0, meta["uscale"])
            rl = ctx.glissade_velo_higher_solve(p, bp=bp, beta_external=bext, initial_guess=1, **kw)
            runs.append(("learned", rl))
    for name, r in runs:
        print(f"  mixed precision from {name}: {r['picard_iters']} Picard, {r['fp32_matvecs']} FP32 + "
              f"{r['fp64_matvecs']} FP64 matvecs, {r['seconds_device']:.3f} s device, "
              f"err {r['final_rel_err']:.2e}")
        assert r["converged"] and 0 <= r["final_rel_err"] <= 1e-6 and r["fallbacks"] == 0
        assert rel_l2(_x(r), _x(ref)) <= 1.01e-6
    if len(runs) == 2:
        assert runs[1][1]["guess_rejected"] == 0
        assert runs[1][1]["picard_iters"] < runs[0][1]["picard_iters"]


def test_c5_headline_and_plain_fp64_reach_the_reference():
    p = configs.ice_sheet(ewn=1200, nsn=1400, nz=11, dx=1000.0, seed=5)
    print(f"\n[configs] ice sheet {p.ewn}x{p.nsn}x{p.nz} seed 5", flush=True)
    w = SG.default_weights("sheet")
    with G.Context.for_problem(p) as ctx:
        lu, lv = ctx.loads(p)
        ref = reference_solution(ctx, p, lu, lv, precision=1)
        ctx.set_reference_velocity(ref["uvel"], ref["vvel"])
        bp = G.BasalPhysics(configs.HO_BABC_BETA_EXTERNAL)
        bext = p.extra["beta_external"]
        guess = 0
        if os.path.exists(w):
            params, meta = SG.load_weights(w)
            ctx.learned_guess_load(params, meta["nz"], 1 if meta.get("target") == "log1p" else 0,
                                   meta["uscale"])
            guess = 1
        # the bench's headline settings
        rf = ctx.glissade_velo_higher_solve(p, bp=bp, beta_external=bext, initial_guess=guess,
                                            precision=1, adaptive_linear=1, inner_rtol=1e-2, anderson_m=4,
                                            coarse=1, skip_final_residual=1, stop_rel_err=1e-6,
                                            maxiter_nonlinear=300, efvs=False)
        print(f"  headline: {rf['picard_iters']} Picard, {rf['seconds_device']:.3f} s device, "
              f"{rf['seconds_total']:.3f} s with transfers, err {rf['final_rel_err']:.2e}", flush=True)
        r64 = ctx.glissade_velo_higher_solve(p, bp=bp, beta_external=bext, initial_guess=0, precision=0,
                                             stop_rel_err=1e-6, maxiter_nonlinear=300, efvs=False)
        print(f"  plain FP64: {r64['picard_iters']} Picard, {r64['linear_iters']} PCG, "
              f"{r64['seconds_device']:.3f} s device, err {r64['final_rel_err']:.2e}")
    for r in (rf, r64):
        assert r["converged"] and 0 <= r["final_rel_err"] <= 1e-6 and r["fallbacks"] == 0
    assert rf["guess_rejected"] == 0
    assert 3 * rf["seconds_device"] <= r64["seconds_device"]
