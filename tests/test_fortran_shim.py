"""CPU test of the Fortran drop-in module's argument marshalling: the module and the test
driver are linked against a test double of the C ABI (tests/fortran/mock_abi.c) that
writes documented functions of its inputs, so every Fortran array is seen to reach the
right C argument in the right layout (LOGICAL -> int32, indxA packing, intent(inout))."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = "/opt/rocm/bin/amdflang"
NX, NY, NZ = 9, 8, 4


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if not os.path.exists(FLANG) or not shutil.which("gcc"):
        pytest.skip("amdflang or gcc missing")
    d = tmp_path_factory.mktemp("shim")
    mod = d / "mod"
    mod.mkdir()

    def run(cmd):
        r = subprocess.run([str(c) for c in cmd], capture_output=True, text=True)
        assert r.returncode == 0, r.stdout + r.stderr

    run(["gcc", "-O1", "-shared", "-fPIC", os.path.join(ROOT, "tests", "fortran", "mock_abi.c"),
         "-o", d / "libcism_gpu.so"])
    run([FLANG, "-O2", "-fPIC", "-c", os.path.join(ROOT, "cism_amd", "fortran", "cism_gpu_velo.F90"),
         "-o", d / "m.o", "-module-dir", mod])
    run([FLANG, "-shared", "-fPIC", "-o", d / "libcism_fortran.so", d / "m.o", f"-L{d}", "-lcism_gpu",
         f"-Wl,-rpath,{d}"])
    run([FLANG, "-O2", os.path.join(ROOT, "tests", "fortran", "fortran_driver.F90"), "-o", d / "drv",
         "-I", mod, f"-L{d}", "-lcism_fortran", "-lcism_gpu", f"-Wl,-rpath,{d}"])
    return d


def call(d, header, arrays, out_shapes):
    fin, fout = d / "in.bin", d / "out.bin"
    with open(fin, "wb") as f:
        f.write(np.asarray(header, dtype=np.int32).tobytes())
        for a in arrays:
            f.write(np.asfortranarray(a).tobytes(order="F"))
    r = subprocess.run([str(d / "drv"), str(fin), str(fout)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    data = np.fromfile(fout, dtype=np.float64)
    out, o = [], 0
    for shp in out_shapes:
        n = int(np.prod(shp))
        out.append(data[o:o + n].reshape(shp, order="F"))
        o += n
    assert o == data.size
    return out


rng = np.random.default_rng(7)
V = (NZ, NX - 1, NY - 1)
M = (27,) + V
C = (NZ - 1, NX, NY)


def rnd(shape):
    return np.asfortranarray(rng.standard_normal(shape))


def test_matvec_marshalling(driver):
    av = (rng.random((NX - 1, NY - 1)) > 0.4).astype(np.int32)
    A = [rnd(M) for _ in range(4)]
    xu, xv = rnd(V), rnd(V)
    yu, yv = call(driver, [1, NX, NY, NZ, 0, 2, 2, 0], [av, *A, xu, xv], [V, V])
    np.testing.assert_array_equal(yu, xu + A[0][13] + 1000.0 * av[None])
    np.testing.assert_array_equal(yv, xv + A[3][13] + 2000.0 * av[None])


@pytest.mark.parametrize("mode,niters", [(2, 11), (3, 22)])
def test_pcg_marshalling(driver, mode, niters):
    av = np.ones((NX - 1, NY - 1), np.int32)
    A = [rnd(M) for _ in range(4)]
    bu, bv, xu, xv = rnd(V), rnd(V), rnd(V), rnd(V)
    ou, ov, sc = call(driver, [mode, NX, NY, NZ, 2, 2, 2, 0], [av, *A, bu, bv, xu, xv], [V, V, (4,)])
    np.testing.assert_array_equal(ou, xu + (bu + A[1][0]))
    np.testing.assert_array_equal(ov, xv + (bv + A[2][26]))
    assert sc[0] == 2.25 and sc[1] == niters


def test_assembly_marshalling(driver):
    sigma = np.linspace(0, 1, NZ)
    ac = (rng.random((NX, NY)) > 0.5).astype(np.int32)
    xV, yV, u, v = rnd((NX - 1, NY - 1)), rnd((NX - 1, NY - 1)), rnd(V), rnd(V)
    su, st, ff = rnd((NX - 1, NY - 1)), rnd((NX - 1, NY - 1)), rnd(C)
    ec = 3.5
    outs = call(driver, [4, NX, NY, NZ, 0, 2, 1, 0],
                [sigma, ac, xV, yV, u, v, su, st, ff, np.array([ec, 0, 0, 0])], [C] + [M] * 4)
    np.testing.assert_array_equal(outs[0], 2.0 * ff + ac[None] + sigma[-1])
    m = np.arange(27)[:, None, None, None] / 100.0
    for b in range(4):
        want = u[None] + 10.0 * b + m + v[None] * 2 + st[None, None] + 1 * ec
        np.testing.assert_allclose(outs[1 + b], want, rtol=0, atol=1e-12)


def test_picard_marshalling(driver):
    sigma = np.linspace(0, 1, NZ)
    ac = (rng.random((NX, NY)) > 0.5).astype(np.int32)
    av = (rng.random((NX - 1, NY - 1)) > 0.5).astype(np.int32)
    xV, yV, su, st, beta = (rnd((NX - 1, NY - 1)) for _ in range(5))
    ff = rnd(C)
    um = (rng.random(V) > 0.7).astype(np.int32)
    vm = (rng.random(V) > 0.7).astype(np.int32)
    lu, lv, u0, v0 = rnd(V), rnd(V), rnd(V), rnd(V)
    dx, dy, ec, babc = 2000.0, 2000.0, 7.0, 5.0
    # header: mode, nx, ny, nz, precond, whichapprox, whichefvs, whichsparse
    u, v, sc = call(driver, [5, NX, NY, NZ, 2, 2, 2, 3],
                    [sigma, ac, av, xV, yV, su, st, ff, beta, um, vm, lu, lv, u0, v0,
                     np.array([dx, dy, ec, babc])], [V, V, (4,)])
    np.testing.assert_array_equal(u, lu + um + beta[None] + 2)
    np.testing.assert_array_equal(v, lv + vm + 10.0 * av[None] + dx / 1000.0)
    assert sc[0] == 100 + 5          # maxiter_nonlinear + whichbabc
    assert sc[1] == 300 + 20 + 2     # 100 whichsparse + 10 whichefvs + whichapprox


def test_geometry_solve_marshalling(driver, tmp_path):
    """glissade_velo_higher_solve_gpu: the cell fields, beta_external and the fast path's
    optional arguments reach cism_glissade_velo_higher_solve; the weight file is loaded
    first (mode 6 of the driver: precision from scal(4), adaptive 1, inner_rtol 5e-2,
    Anderson 3, coarse 1, skip_final_residual 1, learned guess from header word 8)"""
    sigma = np.linspace(0, 1, NZ)
    thk, srf, bed = rnd((NX, NY)), rnd((NX, NY)), rnd((NX, NY))
    flw, bext = rnd(C), rnd((NX - 1, NY - 1))
    wfile = tmp_path / "w.bin"
    wfile.write_bytes(b"CISMSG01" + bytes(24))
    fin, fout = driver / "in6.bin", driver / "out6.bin"
    with open(fin, "wb") as f:
        f.write(np.asarray([6, NX, NY, NZ, 2, 2, 2, 1], dtype=np.int32).tobytes())
        for a in (sigma, thk, srf, bed, flw, bext, np.array([1000.0, 0.25, 5.0, 1.0])):
            f.write(np.asfortranarray(a).tobytes(order="F"))
    r = subprocess.run([str(driver / "drv"), str(fin), str(fout), str(wfile)], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    data = np.fromfile(fout, dtype=np.float64)
    n = int(np.prod(V))
    u, v, sc = data[:n].reshape(V, order="F"), data[n:2 * n].reshape(V, order="F"), data[2 * n:]
    cell = (slice(0, NX - 1), slice(0, NY - 1))
    np.testing.assert_array_equal(u, np.broadcast_to((thk[cell] + flw[0][cell])[None] + 1.0 + 10.0, V))
    np.testing.assert_array_equal(v, np.broadcast_to((srf[cell] + bed[cell] + bext)[None] + 3.0, V))
    # picard = whichbabc 5 + 100 adaptive; linear = lateral 1 + 10 local taud; converged;
    # fallbacks = 100 thklim + 1000 coarse + 2000 skip_final_residual
    assert list(sc) == [105.0, 11.0, 1.0, 3025.0]


@pytest.mark.parametrize("approx", [1, 3, 4])
def test_2d_solvers_marshalling(driver, approx):
    """glissade_velo_higher_solve_2d_gpu: SSA goes to cism_velo_higher_solve_2d, L1L2 and
    DIVA to cism_velo_higher_solve_depth, on the module's one-level context; every array
    reaches its argument in Fortran layout and the in/out ones come back"""
    sigma = np.linspace(0, 1, NZ)
    ice = (rng.random((NX, NY)) > 0.3).astype(np.int32)
    ipl = (rng.random((NX, NY)) > 0.3).astype(np.int32)
    ac = (rng.random((NX, NY)) > 0.5).astype(np.int32)
    av = (rng.random((NX - 1, NY - 1)) > 0.5).astype(np.int32)
    S = (NX - 1, NY - 1)
    xV, yV, su, st, dsx, dsy, beta = (rnd(S) for _ in range(7))
    thk = rnd((NX, NY))
    flw, ff = rnd(C), rnd(C)
    um = (rng.random(V) > 0.7).astype(np.int32)
    vm = (rng.random(V) > 0.7).astype(np.int32)
    lu, lv, u0, v0 = rnd(V), rnd(V), rnd(V), rnd(V)
    u2, v2, btx, bty = rnd(S), rnd(S), rnd(S), rnd(S)
    ef0 = rnd(C)
    dx, dy, ec, babc = 2000.0, 2000.0, 7.0, 5.0
    outs = call(driver, [7, NX, NY, NZ, 2, approx, 2, 3],
                [sigma, ice, ipl, ac, av, xV, yV, thk, su, st, dsx, dsy, flw, ff, beta, um, vm, lu, lv,
                 u0, v0, u2, v2, btx, bty, ef0, np.array([dx, dy, ec, babc])],
                [V, V, S, S, S, S, C, (4,)])
    u, v, ou2, ov2, obx, oby, oef, sc = outs
    cell = (slice(0, NX - 1), slice(0, NY - 1))
    if approx == 1:
        np.testing.assert_array_equal(u, lu + um + beta[None] + NZ)
        np.testing.assert_array_equal(v, lv + vm + 10.0 * av[None] + approx)
        np.testing.assert_array_equal(ou2, u[-1])          # the 2D solution out
        np.testing.assert_array_equal(oef, ff + ac[None])
        np.testing.assert_array_equal(obx, btx)
        assert list(sc) == [100 + 5, 300 + 20 + 2, 1.0, 0.0]
    else:
        close = lambda a, b: np.testing.assert_allclose(a, b, rtol=0, atol=1e-12)   # noqa: E731
        close(u, lu + um + (thk[cell] + flw[0][cell] + ice[cell] + 10.0 * ipl[cell])[None])
        close(v, lv + vm + (dsx + 2.0 * dsy + beta + 100.0 * av)[None])
        close(ou2, u2 + btx + xV + st)
        close(ov2, v2 + bty + yV + sigma[-1])
        np.testing.assert_array_equal(obx, 2.0 * btx)
        np.testing.assert_array_equal(oby, 3.0 * bty)
        close(oef, ef0 + ff + ac[None])
        assert list(sc) == [100 + 5 + 1000 * approx, 300 + 20 + 2, 1.0, 0.0]
