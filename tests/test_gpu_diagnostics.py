"""Post-solve diagnostics on the GPU against the CPU oracle: internal stress, basal
friction heat flux, basal traction (glissade_velo_higher.F90:3842-3894, 6755-7195)."""
import numpy as np
import pytest

import oracle as O
from cism_amd import configs, gpu as G
from helpers import rel_max

pytestmark = pytest.mark.gpu


def problems():
    return {
        "dome": configs.dome(),
        "grf": configs.grf_sheet(ewn=24, nsn=20, nz=7, seed=4),
        "sheet": configs.ice_sheet(ewn=36, nsn=42, nz=6, dx=33333.0, seed=7),
    }


@pytest.fixture(scope="module", params=["dome", "grf", "sheet"])
def case(request):
    p = problems()[request.param]
    rng = np.random.default_rng(11)
    u = np.asfortranarray(50 * rng.standard_normal(p.uvel.shape))
    v = np.asfortranarray(50 * rng.standard_normal(p.uvel.shape))
    print(f"[diag] {request.param} {p.ewn}x{p.nsn}x{p.nz}", flush=True)
    with G.Context.for_problem(p) as ctx:
        yield request.param, p, ctx, u, v


@pytest.mark.parametrize("approx", [0, 1, 2])
def test_internal_stress(case, approx):
    name, p, ctx, u, v = case
    efvs = np.zeros((p.nz - 1, p.nx, p.ny), order="F")
    got = ctx.compute_internal_stress(p.sigma, p.active_cell, p.xVertex, p.yVertex, p.stagusrf,
                                      p.stagthck, p.flwafact, efvs, p.whichefvs, p.efvs_constant,
                                      approx, u, v)
    st, want = O.internal_stress(p, u, v, whichapprox=approx)
    assert st == 0
    for g, w in zip(got, want):
        assert rel_max(g, w) < 1e-13


@pytest.mark.parametrize("which", [0, 1])
def test_basal_friction_heatflx(case, which):
    name, p, ctx, u, v = case
    beta = np.asfortranarray(np.abs(p.beta_internal) + 100.0)
    got = ctx.compute_basal_friction_heatflx(p.active_cell, p.active_vertex, p.xVertex, p.yVertex,
                                             u[p.nz - 1], v[p.nz - 1], beta, which)
    want = O.basal_friction_heatflx(p, u[p.nz - 1], v[p.nz - 1], beta, which)
    assert rel_max(got, want) < 1e-14


def test_basal_traction(case):
    name, p, ctx, u, v = case
    beta = np.asfortranarray(np.abs(p.beta_internal) + 100.0)
    bx, by = ctx.compute_basal_traction(beta, u, v)
    ox, oy = O.basal_traction(p, u, v, beta)
    assert np.array_equal(bx, ox) and np.array_equal(by, oy)


def test_diagnostics_of_a_solve():
    """the device-resident entry point after a Picard solve equals the standalone ones on
    the solution"""
    p = configs.grf_sheet(ewn=24, nsn=20, nz=7, seed=4)
    with G.Context.for_problem(p) as ctx:
        lu, lv = ctx.loads(p)
        r = ctx.velo_higher_solve_3d(p, lu, lv)
        d = ctx.velo_diagnostics(p.whichapprox, p.whichefvs, p.efvs_constant, 1)
        u, v = r["uvel"], r["vvel"]
    st, T = O.internal_stress(p, u, v)
    for key, w in zip(["tau_xz", "tau_yz", "tau_xx", "tau_yy", "tau_xy", "tau_eff"], T):
        assert rel_max(d[key], w) < 1e-13, key
    bf = O.basal_friction_heatflx(p, u[p.nz - 1], v[p.nz - 1])
    assert rel_max(d["bfricflx"], bf) < 1e-14
    bx, by = O.basal_traction(p, u, v)
    assert np.array_equal(d["btractx"], bx) and np.array_equal(d["btracty"], by)
