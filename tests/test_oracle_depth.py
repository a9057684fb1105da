"""CPU pins of the oracle's depth-integrated 2D solvers, L1L2 (which_ho_approx 3) and DIVA
(4): the grid operators against exact results on linear fields, the layer viscosities
against Glen's flow law they solve, DIVA's column integrals against closed forms, and
both Picard loops on the sliding slab against the analytic Glen-slab velocity profile,
whose midpoint-rule error must fall as dsigma^2."""
import math

import numpy as np
import pytest

import oracle as O
from cism_amd import configs

RHOI, GRAV = 910.0, 9.81


def _linear_cells(nx, ny, dx, dy, a, b, c):
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dy
    return np.asfortranarray(a + b * x[:, None] + c * y[None, :])


def test_stagger_of_linear_and_constant_fields():
    nx, ny, dx = 9, 7, 1000.0
    f = _linear_cells(nx, ny, dx, dx, 3.0, 2e-3, -5e-4)
    s = O.stagger(f)
    xv = (np.arange(1, nx)) * dx
    yv = (np.arange(1, ny)) * dx
    exact = 3.0 + 2e-3 * xv[:, None] - 5e-4 * yv[None, :]
    assert np.allclose(s, exact, rtol=0, atol=1e-14 * 10)
    # margin 1: a constant survives any mask where a masked cell is near; zero elsewhere
    rng = np.random.default_rng(1)
    mask = (rng.random((nx, ny)) < 0.4).astype(np.int32)
    sm = O.stagger(np.full((nx, ny), 7.25), mask, 1)
    anyc = (mask[:-1, :-1] + mask[1:, :-1] + mask[:-1, 1:] + mask[1:, 1:]) > 0
    assert np.array_equal(sm[anyc], np.full(anyc.sum(), 7.25))
    assert np.all(sm[~anyc] == 0.0)
    # margin 1 over a full mask is margin 0
    assert np.array_equal(O.stagger(f, np.ones((nx, ny), np.int32), 1), s)


def test_gradient_of_a_linear_field_is_exact():
    nx, ny, dx, dy = 10, 8, 1000.0, 1500.0
    f = _linear_cells(nx, ny, dx, dy, 1.0, 3e-3, -2e-3)
    gx, gy = O.gradient(f, dx, dy)
    assert np.allclose(gx, 3e-3, rtol=1e-12) and np.allclose(gy, -2e-3, rtol=1e-12)
    # margin 1: an edge counts only between two masked cells; any counted edge of a
    # linear field gives the exact slope, a vertex with none gives zero
    rng = np.random.default_rng(2)
    m = (rng.random((nx, ny)) < 0.6).astype(np.int32)
    gx, gy = O.gradient(f, dx, dy, m, 1)
    ex = (m[:-1, :] == 1) & (m[1:, :] == 1)          # x edges (i, j) for j = 1..ny
    ey = (m[:, :-1] == 1) & (m[:, 1:] == 1)          # y edges (i, j) for i = 1..nx
    hx = ex[:, :-1] | ex[:, 1:]
    hy = ey[:-1, :] | ey[1:, :]
    assert np.allclose(gx[hx], 3e-3, rtol=1e-12) and np.all(gx[~hx] == 0.0)
    assert np.allclose(gy[hy], -2e-3, rtol=1e-12) and np.all(gy[~hy] == 0.0)


def _qp_tables(p=0):
    r = 1.0 / math.sqrt(3.0)
    xq, yq = ([-r, r, r, -r][p], [-r, -r, r, r][p])
    phi = np.array([(1 - xq) * (1 - yq), (1 + xq) * (1 - yq), (1 + xq) * (1 + yq), (1 - xq) * (1 + yq)]) / 4
    return phi


def _element(dx=1000.0, dy=1000.0, p=0):
    """unit-square-like element: basis derivatives on a dx x dy box at quadrature point p"""
    r = 1.0 / math.sqrt(3.0)
    xq, yq = ([-r, r, r, -r][p], [-r, -r, r, r][p])
    dxr = np.array([-(1 - yq), (1 - yq), (1 + yq), -(1 + yq)]) / 4
    dyr = np.array([-(1 - xq), -(1 + xq), (1 + xq), (1 - xq)]) / 4
    return _qp_tables(p), dxr * 2 / dx, dyr * 2 / dy


def test_l1l2_viscosity_solves_the_flow_law():
    sigma = configs.glide_sigma(11)
    A = np.full(10, 1.0e-16)
    ff = 0.5 * A ** (-1.0 / 3.0)
    phi, ddx, ddy = _element()
    xs = np.array([0.0, 1000.0, 1000.0, 0.0])
    u = 50.0 + 2e-3 * xs            # du/dx = 2e-3 /yr
    v = np.zeros(4)
    h = np.full(4, 1500.0)
    eps = math.sqrt(1e-16 + 4e-6)
    # flat surface: no SIA stress, Glen's law in the horizontal plane alone
    e = O.efvs_l1l2(2, 0.0, sigma, phi, ddx, ddy, u, v, h, np.zeros(4), np.zeros(4), A, ff)
    assert np.allclose(e, 0.5 * A ** (-1 / 3) * eps ** (-2 / 3), rtol=1e-12)
    # sloped surface, both branches of the cubic: A (tau_par^2 + tau_perp^2) tau_par = eps
    for slope in (1e-2, 1e-7):
        ds = np.full(4, -slope)
        e = O.efvs_l1l2(2, 0.0, sigma, phi, ddx, ddy, u, v, h, ds, np.zeros(4), A, ff)
        tperp = RHOI * GRAV * 1500.0 * sigma[1:] * slope
        tpar = np.sqrt(1.0 / (2 * A * e) - tperp ** 2)
        assert np.allclose(A * (tpar ** 2 + tperp ** 2) * tpar, eps, rtol=1e-8)
    assert np.array_equal(O.efvs_l1l2(0, 3.5e6, sigma, phi, ddx, ddy, u, v, h, ds, ds, A, ff),
                          np.full(10, 3.5e6))


def test_diva_viscosity_update_and_fixed_point():
    sigma = configs.glide_sigma(11)
    stag = (sigma[:-1] + sigma[1:]) / 2
    A = 1.0e-16
    ff = np.full(10, 0.5 * A ** (-1.0 / 3.0))
    phi, ddx, ddy = _element()
    xs = np.array([0.0, 1000.0, 1000.0, 0.0])
    u, v = 30.0 + 1e-3 * xs, np.zeros(4)
    eps2 = 1e-16 + 1e-6
    z = np.zeros(4)
    # no basal traction: Glen's law of the horizontal strain rate
    e = O.efvs_diva(2, 0.0, stag, phi, ddx, ddy, u, v, z, z, ff, np.full(10, 1e14))
    assert np.allclose(e, ff * eps2 ** (-1 / 3), rtol=1e-12)
    # a zero efvs starts from the minimum strain rate's viscosity
    tb = np.full(4, 8.0e4)
    e0 = O.efvs_diva(2, 0.0, stag, phi, ddx, ddy, u, v, tb, z, ff, np.zeros(10))
    e1 = ff * (eps2 + 0.25 * (8.0e4 * stag / (ff * 1e-8 ** (-2 / 3))) ** 2) ** (-1 / 3)
    assert np.allclose(e0, e1, rtol=1e-12)
    # iterated to its fixed point, the update satisfies Glen's law with the total
    # effective strain rate: A (2 e eps_e)^3 = eps_e, eps_e^2 = eps^2 + (tau_b sigma / e)^2 / 4
    e = np.full(10, 1e14)
    for _ in range(200):
        e = O.efvs_diva(2, 0.0, stag, phi, ddx, ddy, u, v, tb, z, ff, e)
    epse = np.sqrt(eps2 + 0.25 * (8.0e4 * stag / e) ** 2)
    assert np.allclose(A * (2 * e * epse) ** 3, epse, rtol=1e-10)


def test_diva_integrals_closed_form():
    sigma = configs.glide_sigma(11)
    H, e = 1200.0, 3.0e13
    eq = np.full((10, 4), e)
    ok, om = O.integrals_diva(sigma, H, eq)
    # midpoint rule is exact for the linear integrand: omega_k = H (1 - sigma^2) / (2 efvs)
    assert np.allclose(ok, H * (1 - sigma ** 2) / (2 * e), rtol=1e-12, atol=1e-30)
    # omega is the trapezoid integral of omega_k over sigma
    assert om == pytest.approx(np.trapezoid(ok, sigma), rel=1e-13)
    # a different efvs at each quadrature point: omega_k is the mean of their integrals
    eq2 = np.array([[1.0, 2.0, 4.0, 8.0]] * 10) * e
    ok2, _ = O.integrals_diva(sigma, H, eq2)
    assert np.allclose(ok2, ok * np.mean(1.0 / np.array([1.0, 2.0, 4.0, 8.0])), rtol=1e-12)


@pytest.fixture(scope="module")
def slab_runs():
    out = {}
    for nz in (11, 21, 41):
        p = configs.sliding_slab(ewn=8, nsn=8, nz=nz)
        lu, lv = O.loads(p)
        for approx in (3, 4):
            r = O.picard_depth(p, lu, lv, approx, maxiter=300 if approx == 4 else 20, target=1e-30)
            assert r["status"] == 0
            out[(nz, approx)] = r
    return out


@pytest.mark.parametrize("approx", [3, 4])
def test_slab_converges_to_the_glen_profile(slab_runs, approx):
    """uniform slab on a sliding bed: u_b = tau_b / beta, u_s - u_b = (A/2) tau_b^3 H and
    (DIVA) the depth mean u_b + (2A/5) tau_b^3 H, the discrete profile within O(dsigma^2)"""
    H, A, beta = 1000.0, 1.0e-16, 500.0
    tau = RHOI * GRAV * H * math.tan(math.radians(0.5))
    ub = tau / beta
    errs = []
    for nz in (11, 21, 41):
        r = slab_runs[(nz, approx)]
        u = r["uvel"][:, 5, 5]
        assert np.allclose(r["uvel"][-1], ub, rtol=1e-9)
        assert np.allclose(r["vvel"], 0.0, atol=1e-9)
        errs.append((u[0] - u[-1]) / (0.5 * A * tau ** 3 * H) - 1.0)
        if approx == 4:
            em = (r["uvel_2d"][5, 5] - ub) / (0.4 * A * tau ** 3 * H) - 1.0
            assert abs(em) < 0.01
            # basal traction = driving stress at the fixed point
            assert r["btractx"][5, 5] == pytest.approx(tau, rel=1e-9)
    assert abs(errs[0]) < 0.01
    assert 3.5 < errs[0] / errs[1] < 4.5 and 3.5 < errs[1] / errs[2] < 4.5


def test_l1l2_and_diva_agree_on_the_slab(slab_runs):
    """both integrate the same Glen shear profile with the midpoint rule"""
    for nz in (11, 21, 41):
        a, b = slab_runs[(nz, 3)]["uvel"][:, 5, 5], slab_runs[(nz, 4)]["uvel"][:, 5, 5]
        assert np.allclose(a, b, rtol=1e-9)


def test_l1l2_no_slip_is_zero():
    p = configs.sliding_slab(ewn=8, nsn=8, nz=7)
    p.whichbabc = configs.HO_BABC_NO_SLIP
    p.umask_dirichlet[-1] = 1
    p.vmask_dirichlet[-1] = 1
    lu, lv = O.loads(p)
    r = O.picard_depth(p, lu, lv, 3)
    assert r["status"] == 0 and r["picard_iters"] == 1 and r["L2_norm"] == 0.0
    assert np.all(r["uvel_2d"] == 0.0)
