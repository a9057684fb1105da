"""CPU pins of the oracle's grounded fraction and calcbeta (oracle/cism_oracle.c:
orc_grounded_fraction, orc_calcbeta, the in-loop calcbeta of orc_picard_solve/_2d):
hand-built masks, the limits of each sliding law, the f_ground weighting and floor,
agreement of the in-loop form with a precomputed beta, and the sliding slab's
power-law closed form."""
import math

import numpy as np
import pytest

import oracle as O
from cism_amd import configs


def masks(rows):
    """cell masks from strings: '.' none, 'G' grounded ice, 'F' floating ice, 'L' bare
    land; rows listed north to south, columns west to east -> (nx, ny) arrays"""
    ny, nx = len(rows), len(rows[0])
    ice, fl, land = (np.zeros((nx, ny), np.int32, order="F") for _ in range(3))
    for jr, row in enumerate(rows):
        j = ny - 1 - jr
        for i, ch in enumerate(row):
            ice[i, j] = ch in "GF"
            fl[i, j] = ch == "F"
            land[i, j] = ch in "GL"
    return ice, fl, land


def test_grounded_fraction_by_hand():
    ice, fl, land = masks(["....",
                           ".FF.",
                           ".GF.",
                           "...."])
    st, fg = O.grounded_fraction(ice, fl, land, 0)
    assert st == 0
    # vertex (i, j) sees cells (i..i+1, j..j+1): the four around the grounded cell (1, 1)
    # are grounded, the other five next to ice touch floating cells only
    assert fg[0, 0] == 1 and fg[1, 1] == 1 and fg[0, 1] == 1 and fg[1, 0] == 1
    assert fg[2, 2] == 0 and fg[2, 1] == 0 and fg[1, 2] == 0
    assert fg[0, 2] == 0 and fg[2, 0] == 0 and fg.sum() == 4
    st, fg_all = O.grounded_fraction(ice, fl, land, 2)
    assert st == 0 and fg_all[2, 2] == 1 and fg_all.sum() == 9
    st, _ = O.grounded_fraction(ice, fl, land, 1)     # the GLP is not supported
    assert st != 0


def test_grounded_fraction_vertex_mask_quirk():
    # a vertex whose only land cell is its south-east cell (i+1, j): the reference's
    # vertex mask misses it, so f_ground stays 0 there
    ice, fl, land = masks(["...",
                           "...",
                           ".L."])
    st, fg = O.grounded_fraction(ice, fl, land, 0)
    assert st == 0
    assert fg[0, 0] == 0.0           # cells (0..1, 0..1): land only at (1, 0) = SE
    assert fg[1, 0] == 1.0           # cells (1..2, 0..1): land at (1, 0) = SW
    assert fg[0, 1] == 0.0 and fg[1, 1] == 0.0


def test_generator_grounded_fraction_is_the_oracles():
    p = configs.ice_sheet(ewn=36, nsn=42, nz=6, dx=33333.0, seed=7)
    st, fg = O.grounded_fraction(p.ice_mask, p.floating_mask, p.land_mask, 0)
    assert st == 0
    assert np.array_equal(fg, configs.grounded_fraction_vertex(p.ice_mask, p.floating_mask,
                                                               p.land_mask))


@pytest.fixture(scope="module")
def grid():
    return configs.grf_sheet(ewn=16, nsn=14, nz=5, seed=3)


def flat(p, val):
    return np.full((p.nx - 1, p.ny - 1), float(val), order="F")


def test_constant_large_simple(grid):
    p = grid
    u = flat(p, 50.0)
    st, b = O.calcbeta(p, O.BasalPhysics(0, ho_beta_const=321.0), u, u)
    assert st == 0 and np.all(b == 321.0)
    st, b = O.calcbeta(p, O.BasalPhysics(4), u, u)
    assert st == 0 and np.all(b == 1.0e10)
    st, b = O.calcbeta(p, O.BasalPhysics(14), u, u)
    assert st == 0
    # the weak strip covers rows 5 .. ny-5 (1-based), the halo rows take periodic images
    assert b[3, 6] == 100.0 and b[3, p.ny - 5 - 1] == 100.0


def test_power_law_and_speed_floor(grid):
    p = grid
    bp = O.BasalPhysics(9, powerlaw_c=2.0e4, powerlaw_m=3.0)
    u, v = flat(p, 30.0), flat(p, 40.0)
    st, b = O.calcbeta(p, bp, u, v)
    s = math.sqrt(50.0 ** 2 + 1e-4)
    assert st == 0 and b[4, 4] == pytest.approx(2.0e4 * s ** (-2.0 / 3.0), rel=1e-14)
    # basal stress beta*|u| = C |u|^(1/m) in the limit of large speed
    assert b[4, 4] * 50.0 == pytest.approx(2.0e4 * 50.0 ** (1.0 / 3.0), rel=1e-7)
    # a stagnant bed sees the 0.01 m/yr floor speed; beta_powerlaw_umax caps the speed
    st, b0 = O.calcbeta(p, bp, flat(p, 0.0), flat(p, 0.0))
    assert b0[4, 4] == pytest.approx(2.0e4 * 0.01 ** (-2.0 / 3.0), rel=1e-14)
    bp.beta_powerlaw_umax = 10.0
    st, bc = O.calcbeta(p, bp, u, v)
    assert bc[4, 4] == pytest.approx(2.0e4 * 10.0 ** (-2.0 / 3.0), rel=1e-14)


def test_coulomb_laws_limits(grid):
    p = grid
    N = flat(p, 2.0e6)
    for law in (11, 12):
        bp = O.BasalPhysics(law, powerlaw_c=1.0e4, powerlaw_m=3.0, coulomb_c=0.5, beta_grounded_min=0.0)
        tc = 0.5 * 2.0e6
        # slow: the power-law stress is the smaller one (Tsai picks it; Schoof tends to it)
        us = flat(p, 1e-3)
        st, b = O.calcbeta(p, bp, us, flat(p, 0.0), effecpress_stag=N)
        s = math.sqrt(1e-6 + 1e-4)
        tp = 1.0e4 * s ** (1.0 / 3.0)
        assert st == 0 and b[4, 4] * s == pytest.approx(tp, rel=1e-3 if law == 11 else 1e-14)
        # fast: the stress tends to the Coulomb limit c N
        uf = flat(p, 1.0e9)
        st, b = O.calcbeta(p, bp, uf, flat(p, 0.0), effecpress_stag=N)
        assert b[4, 4] * 1.0e9 == pytest.approx(tc, rel=1e-2 if law == 11 else 1e-14)


def test_f_ground_weight_floor_and_errors(grid):
    p = grid
    u = flat(p, 10.0)
    fg = flat(p, 1.0)
    fg[3:6, 3:6] = 0.0
    fg[8, 8] = 0.5
    bext = flat(p, 0.3)
    bext[8, 8] = 500.0
    st, b = O.calcbeta(p, O.BasalPhysics(5, beta_grounded_min=1.0), u, u, beta_external=bext, f_ground=fg)
    assert st == 0
    assert np.all(b[3:6, 3:6] == 0.0)            # floating: no traction
    assert b[8, 8] == 250.0                       # weighted by f_ground
    assert b[2, 2] == 1.0                         # floored at beta_grounded_min
    st, _ = O.calcbeta(p, O.BasalPhysics(5), u, u, beta_external=flat(p, -1.0), f_ground=fg)
    assert st != 0                                # every external beta negative
    neg = flat(p, 5.0)
    neg[4, 7] = -2.0
    st, b = O.calcbeta(p, O.BasalPhysics(5, beta_grounded_min=0.0), u, u, beta_external=neg,
                       f_ground=flat(p, 1.0))
    assert st == 0 and b[4, 7] == 0.0             # the floor (0 here) lifts a negative value
    nan = flat(p, 5.0)
    nan[4, 7] = np.nan
    st, _ = O.calcbeta(p, O.BasalPhysics(5), u, u, beta_external=nan, f_ground=flat(p, 1.0))
    assert st != 0                                # NaN fails the beta >= 0 check


def test_in_loop_velocity_independent_law_equals_fixed_beta():
    p = configs.grf_sheet(ewn=16, nsn=14, nz=5, seed=3)
    lu, lv = O.loads(p)
    fg = flat(p, 1.0)
    bp = O.BasalPhysics(5, beta_grounded_min=1.0)
    bext = np.asfortranarray(np.maximum(p.beta_internal, 1.0))
    st, b = O.calcbeta(p, bp, flat(p, 0.0), flat(p, 0.0), beta_external=bext, f_ground=fg)
    p.beta_internal[:] = b
    r0 = O.picard(p, lu, lv)
    r1 = O.picard(p, lu, lv, basal=(bp, bext, flat(p, 0.0), fg))
    assert r0["status"] == r1["status"] == 0
    assert r0["picard_iters"] == r1["picard_iters"]
    assert np.array_equal(r0["uvel"], r1["uvel"]) and np.array_equal(r0["vvel"], r1["vvel"])


def test_ssa_slab_power_law_closed_form():
    """sliding slab with tau_b = C |u|^(1/m): the uniform SSA solution has C u^(1/m) equal
    to the driving stress rho_i g H tan(theta), so u = (rho_i g H tan(theta) / C)^m"""
    H, th, C, m = 1000.0, 0.5, 2.0e4, 3.0
    p = configs.sliding_slab(H=H, theta_deg=th)
    lu, lv = O.load_gravity(p)
    bp = O.BasalPhysics(9, powerlaw_c=C, powerlaw_m=m, beta_grounded_min=0.0)
    r = O.picard_2d(p, lu, lv, basal=(bp, flat(p, 0.0), flat(p, 0.0), flat(p, 1.0)), maxiter=200)
    assert r["status"] == 0
    taud = 910.0 * 9.81 * H * math.tan(math.radians(th))
    ua = (taud / C) ** m
    own = (slice(None), slice(2, p.nx - 2), slice(2, p.ny - 2))
    np.testing.assert_allclose(r["uvel"][own], ua, rtol=1e-5)
