"""CPU-side checks: the C-ABI library loads and exports every declared entry point (no
compute without a GPU); the host-side input preparation restates the reference's
routines (glissade_stagger, surface gradient, masks, vertex geometry, flow factor)."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from cism_amd import configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "cism_gpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cism_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib_path = os.path.join(ROOT, "cism_amd", "libcism_gpu.so")
    if not os.path.exists(lib_path):
        pytest.fail("libcism_gpu.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(lib_path)
    names = declared_symbols()
    assert len(names) >= 19
    for n in names:
        assert hasattr(lib, n), n
    from cism_amd import gpu
    assert set(gpu.EXPORTED) == set(names)


def test_noalloc_entry_point_rejects_bad_sizes_before_any_gpu_call():
    """the no-allocation residual entry point checks its grid against its fixed scratch and
    limits before it touches the GPU: these calls return CISM_ERR_ARG on a host without one"""
    lib_path = os.path.join(ROOT, "cism_amd", "libcism_gpu.so")
    if not os.path.exists(lib_path):
        pytest.fail("libcism_gpu.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(lib_path)
    f = lib.cism_compute_residual_vector_3d_noalloc
    f.argtypes = [ctypes.c_int] * 5 + [ctypes.c_void_p] * 13
    lib.cism_last_error.restype = ctypes.c_char_p
    dummy = [ctypes.c_void_p(8)] * 12
    assert f(40000, 40000, 11, 1, 0, None, *dummy) == 5      # 1.8e10 nodes > 2^30
    assert b"maximum" in lib.cism_last_error()
    assert f(5, 40, 11, 1, 0, None, *dummy) == 5             # nx < 7
    assert f(40, 40, 11, 1, 0, None, *([None] * 12)) == 5   # null arrays


def test_fortran_shim_exports():
    path = os.path.join(ROOT, "cism_amd", "libcism_fortran.so")
    if not os.path.exists(path):
        pytest.skip("Fortran shim not built (amdflang absent)")
    out = os.popen(f"nm -D {path}").read()
    for proc in ("pcg_solver_standard_3d", "pcg_solver_chrongear_3d",
                 "matvec_multiply_structured_3d", "assemble_stiffness_matrix_3d"):
        assert proc in out, proc


def test_glide_sigma():
    s = configs.glide_sigma(10)
    ref = [0, .2533, .4408, .5833, .6943, .7823, .8533, .9115, .9596, 1]
    assert np.allclose(s, ref, atol=5e-5)
    w = configs.stagwbnd(s)
    assert w[0] == 0 and w[-1] == 1 and np.allclose(w[1:-1], (s[:-1] + s[1:]) / 2)


def test_stagger_margin1_loops():
    rng = np.random.default_rng(0)
    v = rng.random((7, 6))
    m = (rng.random((7, 6)) > 0.4).astype(np.int32)
    out = configs.stagger_margin1(v, m)
    for j in range(5):
        for i in range(6):
            sv = (m[i, j + 1] * v[i, j + 1] + m[i + 1, j + 1] * v[i + 1, j + 1]
                  + m[i, j] * v[i, j] + m[i + 1, j] * v[i + 1, j])
            sm = m[i, j + 1] + m[i + 1, j + 1] + m[i, j] + m[i + 1, j]
            assert out[i, j] == (sv / sm if sm > 0 else 0.0)


def test_surface_gradient_loops():
    rng = np.random.default_rng(1)
    nx, ny = 9, 8
    usrf = rng.random((nx, ny)) * 100
    thck = rng.random((nx, ny)) * 50
    active = (rng.random((nx, ny)) > 0.3).astype(np.int32)
    land = (rng.random((nx, ny)) > 0.5).astype(np.int32)
    dx, dy, thklim = 1000.0, 900.0, 20.0
    gx, gy = configs.surface_gradient(usrf, thck, active, land, dx, dy, thklim, ramp=0.0,
                                      max_slope=0.05, periodic=False)
    ex = np.zeros((nx, ny))
    for j in range(ny):
        for i in range(nx - 1):
            iu, il, sf = (i, i + 1, -1.0) if usrf[i, j] > usrf[i + 1, j] else (i + 1, i, 1.0)
            ef = (1.0 if thck[iu, j] > thklim else 0.0) if land[iu, j] == 1 else 1.0
            if active[iu, j] == 1 and (active[il, j] == 1 or land[il, j] == 1):
                ex[i, j] = ef * sf * (usrf[iu, j] - usrf[il, j]) / dx
    ey = np.zeros((nx, ny))
    for j in range(ny - 1):
        for i in range(nx):
            ju, jl, sf = (j, j + 1, -1.0) if usrf[i, j] > usrf[i, j + 1] else (j + 1, j, 1.0)
            ef = (1.0 if thck[i, ju] > thklim else 0.0) if land[i, ju] == 1 else 1.0
            if active[i, ju] == 1 and (active[i, jl] == 1 or land[i, jl] == 1):
                ey[i, j] = ef * sf * (usrf[i, ju] - usrf[i, jl]) / dy
    for j in range(ny - 1):
        for i in range(nx - 1):
            if 0 < i < nx - 2 and 0 < j < ny - 2:   # interior (the outflow halo zeroes the rim)
                gxx = 0.5 * (ex[i, j] + ex[i, j + 1])
                gyy = 0.5 * (ey[i, j] + ey[i + 1, j])
                gxx = min(gxx, 0.05) if gxx > 0 else max(gxx, -0.05)
                gyy = min(gyy, 0.05) if gyy > 0 else max(gyy, -0.05)
                assert gx[i, j] == gxx and gy[i, j] == gyy


def test_vertex_geometry_and_flow_factor():
    p = configs.dome(ewn=11, nsn=11, nz=5, dx=2000.0, R=8000.0)
    # active cells: ice cells in nhalo+1 .. n-nhalo+1; active vertices: their corners
    ice = p.thck > p.thklim
    ac = np.zeros_like(ice)
    ac[2:p.nx - 1, 2:p.ny - 1] = ice[2:p.nx - 1, 2:p.ny - 1]
    assert np.array_equal(p.active_cell.astype(bool), ac)
    av = np.zeros((p.nx - 1, p.ny - 1), bool)
    for i, j in zip(*np.nonzero(ac)):
        av[i - 1:i + 1, j - 1:j + 1] = True
    assert np.array_equal(p.active_vertex.astype(bool), av)
    assert p.xVertex[4, 2] == 2000.0 * 5 and p.yVertex[4, 2] == 2000.0 * 3
    assert p.flwafact[0, 5, 5] == pytest.approx(0.5 * (1e-16) ** (-1 / 3), rel=1e-15)
    assert np.all(p.flwafact[:, 0:2, :] == 0)     # outside nhalo+1 .. n-nhalo+1
    assert np.all(p.umask_dirichlet[-1] == 1) and np.all(p.umask_dirichlet[:-1] == 0)


def test_problem_generators_periodic_consistency():
    for p in (configs.ismip_hom_a(ewn=16, nsn=16, nz=5, L=16000.0),
              configs.grf_sheet(ewn=16, nsn=16, nz=5, dx=2000.0, corr=6000.0)):
        sx, sy = p.nx - 1, p.ny - 1
        for a in (p.stagthck, p.dusrf_dx, p.dusrf_dy, p.beta_internal):
            b = configs.stag_halo(a.copy(order="F"))
            assert np.array_equal(a, b)
