"""CPU pins of the oracle's post-solve diagnostics: internal stress, basal friction heat
flux and basal traction (glissade_velo_higher.F90:3842-3894, 6755-7195), against closed
forms."""
import numpy as np
import pytest

import oracle as O
from cism_amd import configs

SCYR = 31536000.0


def linear_field(p, a):
    """u = a0 x + a1 y + a2 z + a6, v = a3 x + a4 y + a5 z + a7 at every node (physical z)"""
    x, y = p.xVertex, p.yVertex
    z = p.stagusrf[None] - p.sigma[:, None, None] * p.stagthck[None]
    u = a[0] * x[None] + a[1] * y[None] + a[2] * z + a[6]
    v = a[3] * x[None] + a[4] * y[None] + a[5] * z + a[7]
    return np.asfortranarray(u), np.asfortranarray(v)


@pytest.mark.parametrize("approx", [0, 1, 2])
def test_stress_of_a_linear_field(approx):
    # isoparametric trilinear elements reproduce fields linear in physical (x, y, z)
    # exactly, so every quadrature point sees the same strain rates: closed forms
    p = configs.dome()
    a = np.array([2e-4, -1e-4, 3e-3, 5e-5, 1.5e-4, -2e-3, 10.0, -4.0])
    u, v = linear_field(p, a)
    st, T = O.internal_stress(p, u, v, whichapprox=approx, whichefvs=2)
    assert st == 0
    ux, uy, uz, vx, vy, vz = a[0], a[1], a[2], a[3], a[4], a[5]
    if approx == 0:
        ux = uy = vx = vy = 0.0
        sq = 1e-16 + 0.25 * (uz * uz + vz * vz)
    elif approx == 1:
        uz = vz = 0.0
        sq = 1e-16 + ux * ux + vy * vy + ux * vy + 0.25 * (vx + uy) ** 2
    else:
        sq = 1e-16 + ux * ux + vy * vy + ux * vy + 0.25 * (vx + uy) ** 2 + 0.25 * (uz * uz + vz * vz)
    act = np.zeros((p.nz - 1, p.nx, p.ny), bool)
    h = 2
    ac = p.active_cell.astype(bool)
    rng_ok = np.zeros_like(ac)
    rng_ok[h:p.nx - h + 1, h:p.ny - h + 1] = True      # cells nhalo+1 .. n-nhalo+1 (1-based)
    act[:] = (ac & rng_ok)[None]
    eta = p.flwafact * sq ** (-1.0 / 3.0)
    want = [eta * uz, eta * vz, 2 * eta * ux, 2 * eta * vy, eta * (vx + uy)]
    xz, yz, xx, yy, xy = want
    want.append(np.sqrt(xx * xx + yy * yy + xx * yy + xy * xy + xz * xz + yz * yz))
    for got, w in zip(T, want):
        assert np.all(got[~act] == 0.0)
        scale = np.abs(w[act]).max() + 1e-300
        assert np.abs(got[act] - w[act]).max() <= 1e-9 * scale + 1e-300


def test_stress_constant_viscosity_and_rigid_motion():
    p = configs.dome()
    # rigid translation: no strain, no stress
    u = np.full(p.uvel.shape, 7.0, order="F")
    v = np.full(p.uvel.shape, -3.0, order="F")
    st, T = O.internal_stress(p, u, v, whichefvs=0)
    # roundoff only: efvs (2.3e6) times the rounding of sums of basis derivatives (~1e-17)
    assert st == 0 and all(np.abs(t).max() < 1e-8 for t in T)
    # simple shear u = c z with constant efvs: tau_xz = efvs * c, tau_eff = |tau_xz|
    a = np.array([0, 0, 1e-2, 0, 0, 0, 0, 0])
    u, v = linear_field(p, a)
    st, T = O.internal_stress(p, u, v, whichefvs=0)
    act = T[5] != 0
    assert act.any()
    np.testing.assert_allclose(T[0][act], p.efvs_constant * 1e-2, rtol=1e-9)
    np.testing.assert_allclose(T[5][act], np.abs(T[0][act]), rtol=1e-12)


def test_basal_friction_local_closed_form():
    p = configs.grf_sheet(ewn=16, nsn=12, nz=5, seed=2)
    ub = np.full((p.nx - 1, p.ny - 1), 30.0, order="F")
    vb = np.full((p.nx - 1, p.ny - 1), -40.0, order="F")
    beta = np.full((p.nx - 1, p.ny - 1), 500.0, order="F")
    q = 500.0 * (30.0 ** 2 + 40.0 ** 2) / SCYR
    bf = O.basal_friction_heatflx(p, ub, vb, beta, whichassemble_bfric=1)
    av = p.active_vertex.astype(float)
    nact = av[1:, 1:] + av[:-1, 1:] + av[1:, :-1] + av[:-1, :-1]     # cells 2..n-1 (1-based)
    own = (slice(2, p.nx - 2), slice(2, p.ny - 2))
    np.testing.assert_allclose(bf[own], (nact * q / 4.0)[1:p.nx - 3, 1:p.ny - 3], rtol=1e-14)
    # periodic halo: every halo cell holds its owned image
    nxo, nyo = p.nx - 4, p.ny - 4
    for i in range(p.nx):
        for j in range(p.ny):
            ii = 2 + (i - 2) % nxo
            jj = 2 + (j - 2) % nyo
            assert bf[i, j] == bf[ii, jj]


def test_basal_friction_standard_closed_form():
    # u linear across each cell in x (u = u_c + g xi on the reference square), v = 0, beta
    # constant: the 2x2 Gauss rule is exact and the cell mean of u^2 is u_c^2 + g^2 / 3
    p = configs.grf_sheet(ewn=16, nsn=12, nz=5, seed=2)
    i = np.arange(p.nx - 1)[:, None] * np.ones((1, p.ny - 1))
    ub = np.asfortranarray(5.0 + 2.0 * i)                 # g = 1 per half cell
    vb = np.zeros_like(ub)
    beta = np.full((p.nx - 1, p.ny - 1), 300.0, order="F")
    bf = O.basal_friction_heatflx(p, ub, vb, beta, whichassemble_bfric=0)
    for ic in range(2, p.nx - 2):                          # owned cells (0-based)
        for jc in range(2, p.ny - 2):
            if not p.active_cell[ic, jc]:
                assert bf[ic, jc] == 0.0
                continue
            uc = 0.5 * (ub[ic - 1, jc] + ub[ic, jc])
            want = 300.0 * (uc ** 2 + 1.0 / 3.0) / SCYR
            assert abs(bf[ic, jc] - want) <= 1e-13 * want


def test_basal_traction_uses_the_bed_level():
    p = configs.dome()
    rng = np.random.default_rng(3)
    u = np.asfortranarray(rng.standard_normal(p.uvel.shape))
    v = np.asfortranarray(rng.standard_normal(p.uvel.shape))
    beta = np.asfortranarray(rng.random((p.nx - 1, p.ny - 1)))
    bx, by = O.basal_traction(p, u, v, beta)
    assert np.array_equal(bx, beta * u[p.nz - 1]) and np.array_equal(by, beta * v[p.nz - 1])
