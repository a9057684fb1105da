"""CPU pins of the oracle's 2D (SSA) path: the assembled 9-point operator of a uniform
grid against the Kronecker closed form, translation invariance, and the sliding-slab
solution of the whole Picard loop."""
import math

import numpy as np
import pytest

import oracle as O
from cism_amd import configs

VOL0 = 1.0e9


def slot(iA, jA):
    """0-based slot of the 2D coupling (iA, jA) in the 27-slot layout (kA = 0)"""
    return (iA + 1) + 3 * (jA + 1) + 9


def test_uniform_grid_stencil_closed_form():
    # constant viscosity and thickness on a uniform grid: Q1 elements with exact 2x2 Gauss
    # give Kxx = (dy/dx) k1 (x) m1 and Kyy = (dx/dy) m1 (x) k1, k1 = (-1, 2, -1), m1 =
    # (1, 4, 1)/6, so Auu = eta H (4 Kxx + Kyy) / vol0 and Avv = eta H (Kxx + 4 Kyy) / vol0
    p = configs.sliding_slab(ewn=12, nsn=10, dx=1000.0, whichefvs=0, efvs_constant=2.0e6)
    p.dy = p.dx
    u2 = np.zeros((p.nx - 1, p.ny - 1), order="F")
    st, A, efvs = O.assemble_2d(p, u2, u2)
    assert st == 0
    k1, m1 = np.array([-1.0, 2.0, -1.0]), np.array([1.0, 4.0, 1.0]) / 6.0
    Kxx, Kyy = np.outer(k1, m1), np.outer(m1, k1)          # [iA+1, jA+1]
    c = 2.0e6 * 1000.0 / VOL0
    i, j = 6, 6                                             # an interior vertex (0-based)
    for iA in (-1, 0, 1):
        for jA in (-1, 0, 1):
            m = slot(iA, jA)
            assert A[0][m, 0, i, j] == pytest.approx(c * (4 * Kxx + Kyy)[iA + 1, jA + 1], rel=1e-12, abs=1e-15)
            assert A[3][m, 0, i, j] == pytest.approx(c * (Kxx + 4 * Kyy)[iA + 1, jA + 1], rel=1e-12, abs=1e-15)
    # only the kA = 0 slots are used
    used = [slot(a, b) for a in (-1, 0, 1) for b in (-1, 0, 1)]
    unused = [m for m in range(27) if m not in used]
    assert all(np.all(a[unused] == 0) for a in A)
    # efvs: the element mean at every layer of active cells
    assert np.all(efvs[:, 2:-2, 2:-2] == 2.0e6)


def test_translation_invariance_and_symmetry():
    p = configs.grf_sheet(ewn=16, nsn=14, nz=5, seed=3)
    rng = np.random.default_rng(0)
    u2 = np.asfortranarray(rng.standard_normal((p.nx - 1, p.ny - 1)) * 50)
    v2 = np.asfortranarray(rng.standard_normal((p.nx - 1, p.ny - 1)) * 50)
    st, A, efvs = O.assemble_2d(p, u2, v2)
    assert st == 0
    # a rigid translation has no strain: every interior row of each block sums to zero
    for a in A:
        rows = a[:, 0, 4:-4, 4:-4].sum(axis=0)
        assert np.abs(rows).max() <= 1e-10 * np.abs(a).max()
    # Auu, Avv symmetric; Avu = Auv^T (owned interior vertices)
    for i in range(4, p.nx - 5):
        for j in range(4, p.ny - 5):
            for iA in (-1, 0, 1):
                for jA in (-1, 0, 1):
                    m, mm = slot(iA, jA), slot(-iA, -jA)
                    assert A[0][m, 0, i, j] == pytest.approx(A[0][mm, 0, i + iA, j + jA], rel=1e-12, abs=1e-18)
                    assert A[1][m, 0, i, j] == pytest.approx(A[2][mm, 0, i + iA, j + jA], rel=1e-12, abs=1e-18)


@pytest.mark.parametrize("efvs_kind", [0, 2])
def test_sliding_slab_solution(efvs_kind):
    H, th, beta = 1000.0, 0.5, 500.0
    p = configs.sliding_slab(H=H, theta_deg=th, beta=beta, whichefvs=efvs_kind)
    lu, lv = O.load_gravity(p)
    r = O.picard_2d(p, lu, lv)
    assert r["status"] == 0
    ua = 910.0 * 9.81 * H * math.tan(math.radians(th)) / beta
    own = (slice(None), slice(2, p.nx - 2), slice(2, p.ny - 2))
    np.testing.assert_allclose(r["uvel"][own], ua, rtol=1e-7)
    assert np.abs(r["vvel"][own]).max() < 1e-7 * ua
    # every level carries the 2D solution
    assert np.all(r["uvel"] == r["uvel"][-1:])
