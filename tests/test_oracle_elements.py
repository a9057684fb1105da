"""Pins the oracle's element-level functions against closed forms and the reference's
literal tables (glissade_velo_higher.F90)."""
import math

import numpy as np
import pytest

import oracle as O
from cism_amd import configs

R3 = 1.0 / math.sqrt(3.0)

# literal shift tables of glissade_velo_higher_init (velo_higher:453-486), rows nr = 1..8
ISHIFT = [[0, 1, 1, 0, 0, 1, 1, 0], [-1, 0, 0, -1, -1, 0, 0, -1], [-1, 0, 0, -1, -1, 0, 0, -1],
          [0, 1, 1, 0, 0, 1, 1, 0], [0, 1, 1, 0, 0, 1, 1, 0], [-1, 0, 0, -1, -1, 0, 0, -1],
          [-1, 0, 0, -1, -1, 0, 0, -1], [0, 1, 1, 0, 0, 1, 1, 0]]
JSHIFT = [[0, 0, 1, 1, 0, 0, 1, 1], [0, 0, 1, 1, 0, 0, 1, 1], [-1, -1, 0, 0, -1, -1, 0, 0],
          [-1, -1, 0, 0, -1, -1, 0, 0], [0, 0, 1, 1, 0, 0, 1, 1], [0, 0, 1, 1, 0, 0, 1, 1],
          [-1, -1, 0, 0, -1, -1, 0, 0], [-1, -1, 0, 0, -1, -1, 0, 0]]
KSHIFT = [[0, 0, 0, 0, -1, -1, -1, -1]] * 4 + [[1, 1, 1, 1, 0, 0, 0, 0]] * 4
# reference element node signs (velo_higher:344-396): phi_n = (1+sx x)(1+sy y)(1+sz z)/8
NODE_S = [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1),
          (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)]
QP = [(-R3, -R3, -R3), (R3, -R3, -R3), (R3, R3, -R3), (-R3, R3, -R3),
      (-R3, -R3, R3), (R3, -R3, R3), (R3, R3, R3), (-R3, R3, R3)]


def test_shift_tables_and_indxA():
    t = O.tables()
    assert np.array_equal(t["ishift"], np.array(ISHIFT, dtype=np.int32))
    assert np.array_equal(t["jshift"], np.array(JSHIFT, dtype=np.int32))
    assert np.array_equal(t["kshift"], np.array(KSHIFT, dtype=np.int32))
    ind = t["indxA"]
    assert ind[1, 1, 1] == 14        # diagonal (velo_higher:611-619)
    m = 0
    for k in range(3):
        for j in range(3):
            for i in range(3):
                m += 1
                assert ind[i, j, k] == m


def test_basis_tables_closed_form():
    t = O.tables()
    for p, (x, y, z) in enumerate(QP):
        for n, (sx, sy, sz) in enumerate(NODE_S):
            assert t["phi"][n, p] == pytest.approx((1 + sx * x) * (1 + sy * y) * (1 + sz * z) / 8, abs=1e-16)
            assert t["dxr"][n, p] == pytest.approx(sx * (1 + sy * y) * (1 + sz * z) / 8, abs=1e-16)
            assert t["dyr"][n, p] == pytest.approx(sy * (1 + sx * x) * (1 + sz * z) / 8, abs=1e-16)
            assert t["dzr"][n, p] == pytest.approx(sz * (1 + sx * x) * (1 + sy * y) / 8, abs=1e-16)
        assert t["phi"][:, p].sum() == pytest.approx(1.0, abs=1e-15)        # partition of unity
        for d in ("dxr", "dyr", "dzr"):
            assert abs(t[d][:, p].sum()) < 1e-15


def box_nodes(x0, y0, z_top, dx, dy, dz):
    """Nodes in reference order: 1-4 lower face (z_top-dz), 5-8 upper face."""
    x = np.array([x0, x0 + dx, x0 + dx, x0] * 2, float)
    y = np.array([y0, y0, y0 + dy, y0 + dy] * 2, float)
    z = np.array([z_top - dz] * 4 + [z_top] * 4, float)
    return x, y, z


def test_box_jacobian_and_affine_reproduction():
    dx, dy, dz = 2000.0, 1500.0, 37.5
    x, y, z = box_nodes(1000.0, -300.0, 812.0, dx, dy, dz)
    for p in range(8):
        st, ddx, ddy, ddz, det = O.basis_derivatives(x, y, z, p)
        assert st == 0
        assert det == pytest.approx(dx * dy * dz / 8, rel=1e-14)
        f = 3.0 * x - 2.0 * y + 5.0 * z + 7.0
        assert np.dot(ddx, f) == pytest.approx(3.0, rel=1e-12)
        assert np.dot(ddy, f) == pytest.approx(-2.0, rel=1e-12)
        assert np.dot(ddz, f) == pytest.approx(5.0, rel=1e-12)


def test_sheared_column_affine_reproduction():
    # surface sloping in x and y, thickness varying: the isoparametric map reproduces
    # every affine field exactly on any hexahedron
    rng = np.random.default_rng(2)
    x, y, z = box_nodes(0.0, 0.0, 0.0, 1000.0, 1000.0, 1.0)
    z = np.array([-100.0 + 3e-2 * xi + 1e-2 * yi - 40 * rng.random() for xi, yi in zip(x[:4], y[:4])] +
                 [1.5e-2 * xi - 2e-2 * yi for xi, yi in zip(x[4:], y[4:])])
    for p in range(8):
        st, ddx, ddy, ddz, det = O.basis_derivatives(x, y, z, p)
        assert st == 0 and det > 0
        f = 0.7 * x + 1.3 * y - 2.2 * z
        assert np.dot(ddx, f) == pytest.approx(0.7, rel=1e-11)
        assert np.dot(ddy, f) == pytest.approx(1.3, rel=1e-11)
        assert np.dot(ddz, f) == pytest.approx(-2.2, rel=1e-11)
        assert abs(ddx.sum()) < 1e-15 and abs(ddz.sum()) < 1e-12


def test_singular_jacobian_is_reported():
    x = np.zeros(8)
    y = np.array([0, 0, 1, 1] * 2, float)
    z = np.array([0] * 4 + [1] * 4, float)
    st, *_ = O.basis_derivatives(x, y, z, 0)
    assert st == 2   # ORC_ERR_SINGULAR (velo_higher:6458-6473)


def test_effective_viscosity_linear_field():
    dx, dy, dz = 1000.0, 800.0, 50.0
    x, y, z = box_nodes(0.0, 0.0, 500.0, dx, dy, dz)
    a, b, c = 1.0e-3, 2.0e-1, -4.0e-4   # du/dx, du/dz, dv/dy  (1/yr)
    u, v = a * x + b * z, c * y
    st, ddx, ddy, ddz, det = O.basis_derivatives(x, y, z, 3)
    ff = 0.5 * 1.0e-16 ** (-1.0 / 3.0)
    ev = O.efvs_qp(2, 2, 0.0, ddx, ddy, ddz, u, v, ff)
    sq = 1e-16 + (a * a + c * c + a * c + 0.25 * 0.0) + 0.25 * (b * b)
    assert ev == pytest.approx(ff * sq ** (-1.0 / 3.0), rel=1e-12)
    # SSA drops the vertical shear, SIA keeps only it
    assert O.efvs_qp(2, 1, 0.0, ddx, ddy, ddz, u, v, ff) == pytest.approx(
        ff * (1e-16 + a * a + c * c + a * c) ** (-1.0 / 3.0), rel=1e-12)
    assert O.efvs_qp(2, 0, 0.0, ddx, ddy, ddz, u, v, ff) == pytest.approx(
        ff * (1e-16 + 0.25 * b * b) ** (-1.0 / 3.0), rel=1e-12)
    assert O.efvs_qp(0, 2, 123.0, ddx, ddy, ddz, u, v, ff) == 123.0
    # flow-factor option: typical strain rate vel_scale/len_scale (glimmer_paramets.F90:146-147)
    assert O.efvs_qp(1, 2, 0.0, ddx, ddy, ddz, u, v, ff) == pytest.approx(
        ff * (500.0 / 200.0e3) ** (-2.0 / 3.0), rel=1e-12)


def kron_box_matrices(dx, dy, dz, mu):
    """Exact element matrices of a box with constant viscosity, from 1D linear-element
    stiffness S, mass M and derivative-mass D matrices (node order of the reference)."""
    def S(h):
        return np.array([[1, -1], [-1, 1]]) / h

    def M(h):
        return h / 6 * np.array([[2, 1], [1, 2]])

    D = np.array([[-0.5, -0.5], [0.5, 0.5]])   # D[i][j] = int phi_i' phi_j
    idx = [((s[0] + 1) // 2, (s[1] + 1) // 2, (s[2] + 1) // 2) for s in NODE_S]
    K = {b: np.zeros((8, 8)) for b in ("uu", "uv", "vu", "vv")}
    Sx, Sy, Sz, Mx, My, Mz = S(dx), S(dy), S(dz), M(dx), M(dy), M(dz)
    for a, (ia, ja, ka) in enumerate(idx):
        for b, (ib, jb, kb) in enumerate(idx):
            xx = Sx[ia, ib] * My[ja, jb] * Mz[ka, kb]
            yy = Mx[ia, ib] * Sy[ja, jb] * Mz[ka, kb]
            zz = Mx[ia, ib] * My[ja, jb] * Sz[ka, kb]
            xy = D[ia, ib] * D[jb, ja] * Mz[ka, kb]   # int dphi_a/dx dphi_b/dy
            yx = D[ib, ia] * D[ja, jb] * Mz[ka, kb]   # int dphi_a/dy dphi_b/dx
            K["uu"][a, b] = mu * (4 * xx + yy + zz)
            K["vv"][a, b] = mu * (4 * yy + xx + zz)
            K["uv"][a, b] = mu * (2 * xy + yx)
            K["vu"][a, b] = mu * (2 * yx + xy)
    return K


def test_element_matrix_box_closed_form():
    # one flat slab cell: box elements; constant viscosity
    p = configs.slab(ewn=6, nsn=6, nz=5, dx=1000.0, H=400.0, theta_deg=0.0)
    st, K, ev = O.element_matrix(p, 4, 4, 2)
    assert st == 0
    dz = 400.0 * (p.sigma[2] - p.sigma[1])
    Kx = kron_box_matrices(1000.0, 1000.0, dz, p.efvs_constant / 1e9)
    for name, k in zip(("uu", "uv", "vu", "vv"), K):
        assert np.allclose(k, Kx[name], rtol=1e-12, atol=1e-12 * np.abs(Kx[name]).max())
    assert ev == p.efvs_constant
    # exact symmetries of the computed blocks
    assert np.array_equal(K[0], K[0].T) and np.array_equal(K[3], K[3].T)
    assert np.array_equal(K[2], K[1].T)
