"""The depth-integrated 2D solvers on the GPU, L1L2 (which_ho_approx 3) and DIVA (4),
against the oracle's Picard loops (glissade_velo_higher.F90:1117-1173, 2206-3830,
5208-6356, 7382-7864): same Picard count, 2D and 3D velocity within the FP64 parity bar
(1e-8 relative), the layer viscosities and DIVA's basal traction, on periodic and outflow
grids; the sliding slab against the analytic Glen profile; a no-slip DIVA column
(beta_eff = 1/omega)."""
import math

import numpy as np
import pytest

import oracle as O
from cism_amd import configs, gpu as G
from helpers import rel_l2

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-8


def problems():
    ism = configs.ismip_hom_a(ewn=16, nsn=16, nz=9, L=40000.0)
    ism.umask_dirichlet[:] = 0        # DIVA's no-slip bed is beta_eff = 1/omega, no Dirichlet row
    ism.vmask_dirichlet[:] = 0
    return {
        "slab": configs.sliding_slab(ewn=12, nsn=10, nz=11),
        "grf": configs.grf_sheet(ewn=24, nsn=20, nz=7, seed=4),
        "sheet": configs.ice_sheet(ewn=36, nsn=42, nz=6, dx=33333.0, seed=7),
        "ismip_noslip": ism,
        "grf_outflow": configs.grf_sheet(ewn=24, nsn=20, nz=7, seed=4, periodic=False),
    }


CASES = [("slab", 3), ("slab", 4), ("grf", 3), ("grf", 4), ("sheet", 3), ("sheet", 4), ("ismip_noslip", 4),
         ("grf_outflow", 3), ("grf_outflow", 4)]


@pytest.fixture(scope="module", params=CASES, ids=[f"{n}-{'l1l2' if a == 3 else 'diva'}" for n, a in CASES])
def case(request):
    name, approx = request.param
    p = problems()[name]
    lu, lv = O.loads(p)
    kw = dict(maxiter=60) if approx == 4 else {}
    print(f"[depth] {name} {p.ewn}x{p.nsn}x{p.nz} approx={approx} babc={p.whichbabc}", flush=True)
    ro = O.picard_depth(p, lu, lv, approx, **kw)
    assert ro["status"] == 0
    return name, approx, p, lu, lv, ro


def test_depth_picard_fp64_matches_the_oracle(case):
    name, approx, p, lu, lv, ro = case
    with G.Context.for_problem_2d(p) as ctx:
        rg = ctx.velo_higher_solve_depth(p, lu, lv, approx,
                                         maxiter_nonlinear=60 if approx == 4 else None)
    print(f"  oracle {ro['picard_iters']} Picard {ro['linear_iters']} PCG; GPU {rg['picard_iters']} Picard "
          f"{rg['linear_iters']} PCG; rel L2 2D {rel_l2(rg['uvel_2d'], ro['uvel_2d']):.2e}, "
          f"3D {rel_l2(rg['uvel'], ro['uvel']):.2e}")
    assert rg["picard_iters"] == ro["picard_iters"]
    x2 = lambda r: np.concatenate([r["uvel_2d"].ravel(), r["vvel_2d"].ravel()])   # noqa: E731
    x3 = lambda r: np.concatenate([r["uvel"].ravel(), r["vvel"].ravel()])   # noqa: E731
    assert rel_l2(x2(rg), x2(ro)) < TOL_FP64
    assert rel_l2(x3(rg), x3(ro)) < TOL_FP64
    assert rel_l2(rg["efvs"], ro["efvs"]) < TOL_FP64
    if approx == 4:
        bt = lambda r: np.concatenate([r["btractx"].ravel(), r["btracty"].ravel()])   # noqa: E731
        assert rel_l2(bt(rg), bt(ro)) < TOL_FP64
    assert abs(rg["final_L2_norm"] - ro["L2_norm"]) <= 1e-6 * max(1.0, ro["L2_norm"])


@pytest.mark.parametrize("approx", [3, 4])
def test_depth_slab_glen_profile(approx):
    """u_s - u_b = (A/2) tau^3 H (1 + O(dsigma^2)) on the GPU too"""
    p = configs.sliding_slab(ewn=12, nsn=10, nz=41)
    lu, lv = O.loads(p)
    with G.Context.for_problem_2d(p) as ctx:
        r = ctx.velo_higher_solve_depth(p, lu, lv, approx, maxiter_nonlinear=300 if approx == 4 else 20,
                                        target_override=1e-30)
    H, A, beta = 1000.0, 1.0e-16, 500.0
    tau = 910.0 * 9.81 * H * math.tan(math.radians(0.5))
    u = r["uvel"][:, 5, 5]
    assert u[-1] == pytest.approx(tau / beta, rel=1e-9)
    assert (u[0] - u[-1]) / (0.5 * A * tau ** 3 * H) - 1.0 == pytest.approx(-3.9e-4, abs=2e-5)


def test_depth_inputs_are_checked():
    p = configs.sliding_slab(ewn=12, nsn=10, nz=7)
    lu, lv = O.loads(p)
    with G.Context.for_problem_2d(p) as ctx:
        with pytest.raises(G.CismError):
            ctx.velo_higher_solve_depth(p, lu, lv, 2)
        with pytest.raises(G.CismError):
            ctx.velo_higher_solve_depth(p, lu, lv, 4, anderson_m=3)
    with G.Context.for_problem(p) as ctx3:
        with pytest.raises(ValueError):
            ctx3.velo_higher_solve_depth(p, lu, lv, 3)


@pytest.mark.parametrize("coarse", [0, 1])
@pytest.mark.parametrize("approx", [3, 4])
def test_depth_mixed_precision_reaches_the_fp64_solution(approx, coarse):
    """the fast path's options on the 2D solvers: FP32 Chron-Gear in FP64 refinement with
    inexact Picard (and Anderson for L1L2), with and without the aggregation coarse
    correction, to 1e-6 of the plain FP64 2D solution"""
    p = configs.grf_sheet(ewn=24, nsn=20, nz=7, seed=4)
    lu, lv = O.loads(p)
    with G.Context.for_problem_2d(p) as ctx:
        ref = ctx.velo_higher_solve_depth(p, lu, lv, approx, maxiter_nonlinear=300, target_override=1e-9)
        assert ref["converged"]
        ctx.set_reference_velocity(np.asfortranarray(ref["uvel_2d"][None]), np.asfortranarray(ref["vvel_2d"][None]))
        r = ctx.velo_higher_solve_depth(p, lu, lv, approx, precision=1, adaptive_linear=1, inner_rtol=5e-2,
                                        anderson_m=3 if approx == 3 else 0, stop_rel_err=1e-6,
                                        maxiter_nonlinear=300, coarse=coarse)
    print(f"  approx {approx} coarse {coarse}: FP64 {ref['picard_iters']} Picard; mixed {r['picard_iters']} Picard, "
          f"{r['fp32_matvecs']} FP32 + {r['fp64_matvecs']} FP64 matvecs, err {r['final_rel_err']:.2e}")
    assert r["converged"] and 0 <= r["final_rel_err"] <= 1e-6 and r["fallbacks"] == 0
