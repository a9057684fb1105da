"""CPU pins of the oracle of the solve's upstream inputs (oracle/upstream_oracle.py):
closed forms on fields whose answers are known."""
import numpy as np

import upstream_oracle as U


def test_masks_by_hand():
    thck = np.array([[0.0, 50.0, 200.0, 1000.0]])
    topg = np.array([[-10.0, 5.0, -500.0, -100.0]])
    ice, fl, oc, la, act = U.get_masks(thck, topg, 0.0, 100.0)
    assert ice.tolist() == [[0, 0, 1, 1]]
    assert oc.tolist() == [[1, 0, 0, 0]]
    # 200 m afloat over 500 m of water (910*200 < 1028*500), 1000 m grounded over 100 m
    assert fl.tolist() == [[0, 0, 1, 0]]
    assert la.tolist() == [[0, 1, 0, 0]]
    assert np.array_equal(act, ice)


def test_stagger_constant_and_masked_mean():
    var = np.full((6, 5), 7.25)
    mask = np.ones((6, 5), np.int32)
    assert np.all(U.stagger_margin1(var, mask) == 7.25)
    var = np.arange(30, dtype=float).reshape(6, 5)
    mask = np.zeros((6, 5), np.int32)
    mask[2, 3] = mask[3, 3] = 1
    st = U.stagger_margin1(var, mask)
    # vertex (2, 2) touches cells (2..3, 2..3): only the two masked cells count
    assert st[2, 2] == (var[2, 3] + var[3, 3]) / 2
    assert st[0, 0] == 0.0


def test_surface_gradient_of_a_plane():
    nx, ny, dx, dy = 12, 10, 1000.0, 2000.0
    i = np.arange(nx)[:, None]
    j = np.arange(ny)[None, :]
    usrf = np.asfortranarray(3000.0 - 0.01 * dx * i + 0.004 * dy * j + 0.0 * i * j)
    thck = np.full((nx, ny), 1500.0)
    act = np.ones((nx, ny), np.int32)
    land = np.ones((nx, ny), np.int32)
    gx, gy = U.surface_gradient(usrf, thck, act, land, dx, dy, 100.0, periodic=False)
    own = (slice(1, nx - 2), slice(1, ny - 2))
    np.testing.assert_allclose(gx[own], -0.01, rtol=1e-12)
    np.testing.assert_allclose(gy[own], 0.004, rtol=1e-12)
    # the slope limit, and no gradient where the upper cell is inactive
    gx, gy = U.surface_gradient(usrf * 200.0, thck, act, land, dx, dy, 100.0, max_slope=1.0,
                                periodic=False)
    assert np.all(np.abs(gx[own]) == 1.0)
    act[:, :] = 0
    gx, gy = U.surface_gradient(usrf, thck, act, land, dx, dy, 100.0, periodic=False)
    assert np.all(gx == 0) and np.all(gy == 0)


def test_periodic_stag_halo_maps():
    a = np.arange(9 * 7, dtype=float).reshape(9, 7)
    h = U.stag_halo(a, True)
    sx, sy = a.shape
    for i in range(sx):
        for j in range(sy):
            ii = i if 2 <= i <= sx - 2 else (i + sx - 3 if i < 2 else 2)
            jj = j if 2 <= j <= sy - 2 else (j + sy - 3 if j < 2 else 2)
            assert h[i, j] == a[ii, jj]


def test_vertex_geometry_invariants():
    nx, ny = 10, 9
    rng = np.random.default_rng(0)
    act = (rng.random((nx, ny)) > 0.5).astype(np.int32)
    xV, yV, ac, av = U.vertex_geometry(nx, ny, 500.0, 700.0, act)
    assert xV[3, 4] == 500.0 * 4 and yV[3, 4] == 700.0 * 5
    assert np.all(ac[:2] == 0) and np.all(ac[:, :2] == 0) and np.all(ac[nx - 1:] == 0)
    for i in range(nx - 1):
        for j in range(ny - 1):
            assert av[i, j] == int(ac[i:i + 2, j:j + 2].any())


def test_flwafact_closed_form():
    flwa = np.full((4, 8, 8), 1e-16)
    flwa[1, 4, 4] = 0.0
    ff = U.flwafact(flwa)
    assert ff[0, 4, 4] == 0.5 * (1e-16) ** (-1.0 / 3.0)
    np.testing.assert_allclose(ff[0, 4, 4], 0.5 * 10 ** (16.0 / 3.0), rtol=1e-14)
    assert ff[1, 4, 4] == 0.0 and ff[0, 0, 0] == 0.0 and ff[0, 7, 7] == 0.0


def test_workload_generator_uses_the_same_inputs():
    """the synthetic workloads' inputs (cism_amd/configs.py) are the oracle's"""
    from cism_amd import configs
    p = configs.ice_sheet(ewn=36, nsn=42, nz=6, dx=33333.0, seed=7)
    ice, fl, oc, la, act = U.get_masks(p.thck, p.topg, 0.0, p.thklim)
    assert np.array_equal(ice, p.ice_mask) and np.array_equal(fl, p.floating_mask)
    ipl = ((act == 1) | (la == 1)).astype(np.int32)
    assert np.array_equal(U.stagger_margin1(p.thck, ipl), p.stagthck)
    gx, gy = U.surface_gradient(p.usrf, p.thck, act, la, p.dx, p.dy, p.thklim)
    assert np.array_equal(gx, p.dusrf_dx) and np.array_equal(gy, p.dusrf_dy)
    xV, yV, ac, av = U.vertex_geometry(p.nx, p.ny, p.dx, p.dy, act)
    assert np.array_equal(ac, p.active_cell) and np.array_equal(av, p.active_vertex)
    # the random generators round their fields (configs.settle) to 1e-3
    np.testing.assert_allclose(p.flwafact, U.flwafact(p.flwa), rtol=0, atol=5e-4 + 1e-12)
