"""The 2D (SSA, whichapprox 1) solve on the GPU against the CPU oracle's SSA Picard loop
(glissade_velo_higher.F90:2206-3767 with solve_2d, assemble_stiffness_matrix_2d
5208-5595): same Picard count, velocity within the FP64 parity bar, efvs and the final
residual; the sliding slab against its closed form; mixed precision within 1e-6."""
import math

import numpy as np
import pytest

import oracle as O
from cism_amd import configs, gpu as G
from helpers import rel_l2, rel_max

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-8


def problems():
    return {
        "slab": configs.sliding_slab(),
        "slab_efvs2": configs.sliding_slab(whichefvs=2, ewn=20, nsn=16),
        "grf": configs.grf_sheet(ewn=24, nsn=20, nz=7, seed=4),
        "sheet": configs.ice_sheet(ewn=36, nsn=42, nz=6, dx=33333.0, seed=7),
    }


@pytest.fixture(scope="module", params=["slab", "slab_efvs2", "grf", "sheet"])
def case(request):
    p = problems()[request.param]
    lu, lv = O.loads(p)
    print(f"[ssa] {request.param} {p.ewn}x{p.nsn}x{p.nz} babc={p.whichbabc}", flush=True)
    ro = O.picard_2d(p, lu, lv)
    assert ro["status"] == 0
    return request.param, p, lu, lv, ro


@pytest.mark.parametrize("sparse", [2, 3])
def test_ssa_picard_fp64(case, sparse):
    name, p, lu, lv, _ = case
    ro = O.picard_2d(p, lu, lv, whichsparse=sparse)
    assert ro["status"] == 0
    with G.Context.for_problem_2d(p) as ctx:
        rg = ctx.velo_higher_solve_2d(p, lu, lv, whichsparse=sparse)
    print(f"  oracle picard={ro['picard_iters']} lin={ro['linear_iters']}  "
          f"gpu picard={rg['picard_iters']} lin={rg['linear_iters']}", flush=True)
    assert rg["picard_iters"] == ro["picard_iters"]
    xo = np.concatenate([ro["uvel"].ravel(), ro["vvel"].ravel()])
    xg = np.concatenate([rg["uvel"].ravel(), rg["vvel"].ravel()])
    e = rel_l2(xg, xo)
    print(f"  relative L2 velocity difference {e:.3e}")
    assert e < TOL_FP64
    # every level of the 3D output carries the 2D solution
    assert np.all(rg["uvel"] == rg["uvel"][-1:]) and np.all(rg["vvel"] == rg["vvel"][-1:])
    assert rel_max(rg["efvs"], ro["efvs"]) < 1e-6
    bo = np.concatenate([ro["bu"].ravel(), ro["bv"].ravel()])
    bg = np.concatenate([rg["bu"].ravel(), rg["bv"].ravel()])
    assert rel_max(bg, bo) < 1e-6


@pytest.mark.parametrize("name", ["slab", "slab_efvs2"])
def test_ssa_slab_closed_form(name):
    """the sliding slab has the closed form u = rho g H tan(alpha) / beta"""
    p = problems()[name]
    lu, lv = O.loads(p)
    with G.Context.for_problem_2d(p) as ctx:
        rg = ctx.velo_higher_solve_2d(p, lu, lv)
    ua = 910.0 * 9.81 * 1000.0 * math.tan(math.radians(0.5)) / 500.0
    own = (slice(None), slice(2, p.nx - 2), slice(2, p.ny - 2))
    np.testing.assert_allclose(rg["uvel"][own], ua, rtol=1e-7)
    assert np.abs(rg["vvel"][own]).max() < 1e-7 * ua


def test_ssa_mixed_precision(case):
    """the mixed-precision SSA solve, stopped (as the fast path is) at 1e-6 relative
    error against the FP64 solution, reaches it within the Picard cap"""
    name, p, lu, lv, ro = case
    with G.Context.for_problem_2d(p) as ctx:
        ctx.set_reference_velocity(np.asfortranarray(ro["uvel"][-1:]), np.asfortranarray(ro["vvel"][-1:]))
        rg = ctx.velo_higher_solve_2d(p, lu, lv, precision=1, stop_rel_err=1e-6)
    xo = np.concatenate([ro["uvel"].ravel(), ro["vvel"].ravel()])
    xg = np.concatenate([rg["uvel"].ravel(), rg["vvel"].ravel()])
    e = rel_l2(xg, xo)
    print(f"  mixed: picard={rg['picard_iters']} fp32 matvecs={rg['fp32_matvecs']} err {e:.3e}")
    assert rg["picard_iters"] < p.maxiter_nonlinear
    assert 0 <= rg["final_rel_err"] <= 1e-6      # owned nodes, as the stopping test measures it
    assert e < 2e-6                               # whole arrays, halos included


def test_ssa_context_guards():
    p = configs.grf_sheet(ewn=16, nsn=14, nz=5, seed=3)
    lu, lv = O.loads(p)
    with G.Context.for_problem_2d(p) as ctx:
        with pytest.raises(G.CismError):
            ctx.velo_higher_solve_3d(p, lu, lv)
    with G.Context.for_problem(p) as ctx:
        with pytest.raises(ValueError):
            ctx.velo_higher_solve_2d(p, lu, lv)
