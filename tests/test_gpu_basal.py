"""Grounded fraction and calcbeta on the GPU against the CPU oracle
(glissade_grounding_line.F90:55-178, glissade_basal_traction.F90:67-598), standalone and
inside the device Picard loop (velo_higher:2502-2517)."""
import math

import numpy as np
import pytest

import oracle as O
from cism_amd import configs, gpu as G
from helpers import rel_l2, rel_max

pytestmark = pytest.mark.gpu


def gp(bp):
    """the same basal physics as the library's struct"""
    return G.BasalPhysics(bp.whichbabc, **{n: getattr(bp, n) for n, _ in bp._fields_[2:]})


@pytest.mark.parametrize("whichground", [0, 2])
def test_grounded_fraction(whichground):
    p = configs.ice_sheet(ewn=36, nsn=42, nz=6, dx=33333.0, seed=7)
    st, want = O.grounded_fraction(p.ice_mask, p.floating_mask, p.land_mask, whichground)
    assert st == 0
    with G.Context.for_problem(p) as ctx:
        got = ctx.grounded_fraction(p.ice_mask, p.floating_mask, p.land_mask, whichground)
    assert np.array_equal(got, want)


LAWS = [O.BasalPhysics(0, ho_beta_const=700.0), O.BasalPhysics(4), O.BasalPhysics(5),
        O.BasalPhysics(9, powerlaw_c=1.5e4, powerlaw_m=3.0, beta_powerlaw_umax=300.0),
        O.BasalPhysics(11, powerlaw_c=1.0e4, powerlaw_m=3.0, coulomb_c=0.42),
        O.BasalPhysics(12, powerlaw_c=1.0e4, powerlaw_m=3.0, coulomb_c=0.42),
        O.BasalPhysics(14)]


@pytest.mark.parametrize("bp", LAWS, ids=lambda b: f"babc{b.whichbabc}")
def test_calcbeta(bp):
    p = configs.ice_sheet(ewn=36, nsn=42, nz=6, dx=33333.0, seed=7)
    rng = np.random.default_rng(5)
    s2 = (p.nx - 1, p.ny - 1)
    ub = np.asfortranarray(rng.standard_normal(s2) * 80)
    vb = np.asfortranarray(rng.standard_normal(s2) * 80)
    bext = np.asfortranarray(np.exp(rng.standard_normal(s2) + 6.0))
    N = np.asfortranarray(np.abs(rng.standard_normal(s2)) * 1e6)
    fg = configs.grounded_fraction_vertex(p.ice_mask, p.floating_mask, p.land_mask)
    st, want = O.calcbeta(p, bp, ub, vb, bext, N, fg)
    assert st == 0
    with G.Context.for_problem(p) as ctx:
        got = ctx.calcbeta(gp(bp), ub, vb, bext, N, fg)
    # pow on the device and in libm may differ in the last bit
    assert rel_max(got, want) < 1e-14


def test_calcbeta_rejects_nan():
    p = configs.grf_sheet(ewn=16, nsn=14, nz=5, seed=3)
    s2 = (p.nx - 1, p.ny - 1)
    bext = np.full(s2, 5.0, order="F")
    bext[4, 4] = np.nan
    with G.Context.for_problem(p) as ctx:
        with pytest.raises(G.CismError):
            ctx.calcbeta(G.BasalPhysics(5), np.zeros(s2, order="F"), np.zeros(s2, order="F"), bext,
                         None, np.ones(s2, order="F"))


@pytest.mark.parametrize("law", [9, 12])
def test_picard_with_sliding_law(law):
    """the device Picard loop recomputing beta from the bed velocity every iteration"""
    p = configs.grf_sheet(ewn=24, nsn=20, nz=7, seed=4)
    s2 = (p.nx - 1, p.ny - 1)
    bp = O.BasalPhysics(law, powerlaw_c=2.0e4, powerlaw_m=3.0, coulomb_c=0.3)
    N = np.full(s2, 3.0e6, order="F")
    fg = np.ones(s2, order="F")
    lu, lv = O.loads(p)
    ro = O.picard(p, lu, lv, basal=(bp, np.zeros(s2), N, fg))
    assert ro["status"] == 0
    with G.Context.for_problem(p) as ctx:
        ctx.set_basal_physics(gp(bp), None, N, fg)
        rg = ctx.velo_higher_solve_3d(p, lu, lv)
        ctx.set_basal_physics(None)
    print(f"  law {law}: oracle picard={ro['picard_iters']} gpu picard={rg['picard_iters']}")
    assert rg["picard_iters"] == ro["picard_iters"]
    xo = np.concatenate([ro["uvel"].ravel(), ro["vvel"].ravel()])
    xg = np.concatenate([rg["uvel"].ravel(), rg["vvel"].ravel()])
    assert rel_l2(xg, xo) < 1e-8


def test_ssa_slab_power_law():
    H, th, C, m = 1000.0, 0.5, 2.0e4, 3.0
    p = configs.sliding_slab(H=H, theta_deg=th)
    s2 = (p.nx - 1, p.ny - 1)
    lu, lv = O.load_gravity(p)
    bp = O.BasalPhysics(9, powerlaw_c=C, powerlaw_m=m, beta_grounded_min=0.0)
    with G.Context.for_problem_2d(p) as ctx:
        ctx.set_basal_physics(gp(bp), None, None, np.ones(s2, order="F"))
        rg = ctx.velo_higher_solve_2d(p, lu, lv, maxiter_nonlinear=200)
    ua = (910.0 * 9.81 * H * math.tan(math.radians(th)) / C) ** m
    own = (slice(None), slice(2, p.nx - 2), slice(2, p.ny - 2))
    np.testing.assert_allclose(rg["uvel"][own], ua, rtol=1e-5)
    ro = O.picard_2d(p, lu, lv, basal=(bp, np.zeros(s2), np.zeros(s2), np.ones(s2)), maxiter=200)
    assert rg["picard_iters"] == ro["picard_iters"]
    assert rel_l2(rg["uvel"].ravel(), ro["uvel"].ravel()) < 1e-8
