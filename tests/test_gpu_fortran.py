"""The Fortran drop-in module (cism_amd/fortran/cism_gpu_velo.F90) called from a Fortran
program with the reference's subroutine signatures, on the GPU, against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from cism_amd import configs
from helpers import assembled_system, rel_l2, rel_max

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "fortran", "fortran_driver")


def run_driver(tmp_path, header, arrays, out_shapes):
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        f.write(np.asarray(header, dtype=np.int32).tobytes())
        for a in arrays:
            f.write(np.asfortranarray(a).tobytes(order="F"))
    print(f"[fortran] mode {header[0]} grid {header[1:4]}", flush=True)
    r = subprocess.run([DRIVER, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    data = np.fromfile(fout, dtype=np.float64)
    out, o = [], 0
    for shp in out_shapes:
        n = int(np.prod(shp))
        out.append(data[o:o + n].reshape(shp, order="F"))
        o += n
    assert o == data.size
    return out


@pytest.fixture(scope="module")
def prob():
    if not os.path.exists(DRIVER):
        pytest.fail("Fortran driver not built (run __graft_entry__.build())")
    return configs.dome()


def test_fortran_matvec_and_pcg(prob, tmp_path):
    p = prob
    A, bu, bv, u, v, _ = assembled_system(p, seed=5)
    vshape = (p.nz, p.nx - 1, p.ny - 1)
    rng = np.random.default_rng(1)
    xu, xv = rng.standard_normal(vshape), rng.standard_normal(vshape)
    xu, xv = np.asfortranarray(xu), np.asfortranarray(xv)
    yu, yv = run_driver(tmp_path, [1, p.nx, p.ny, p.nz, 0, 2, 2, 0],
                        [p.active_vertex.astype(np.int32), *A, xu, xv], [vshape, vshape])
    yo_u, yo_v = O.matvec(p, A, xu, xv)
    assert np.array_equal(yu, yo_u) and np.array_equal(yv, yo_v)
    for mode, cg in ((2, False), (3, True)):
        gu, gv, sc = run_driver(tmp_path, [mode, p.nx, p.ny, p.nz, 2, 2, 2, 0],
                                [p.active_vertex.astype(np.int32), *A, bu, bv, u, v],
                                [vshape, vshape, (4,)])
        ou, ov = u.copy(order="F"), v.copy(order="F")
        st, err, nit = O.pcg(p, A, bu, bv, ou, ov, 2, chrongear=cg)
        assert int(sc[1]) == nit
        assert rel_l2(np.concatenate([gu.ravel(), gv.ravel()]),
                      np.concatenate([ou.ravel(), ov.ravel()])) < 1e-10


def test_fortran_assembly(prob, tmp_path):
    p = prob
    rng = np.random.default_rng(2)
    u = np.asfortranarray(30 * rng.standard_normal(p.uvel.shape))
    v = np.asfortranarray(30 * rng.standard_normal(p.uvel.shape))
    ashape = (27, p.nz, p.nx - 1, p.ny - 1)
    outs = run_driver(tmp_path, [4, p.nx, p.ny, p.nz, 0, p.whichapprox, p.whichefvs, 0],
                      [p.sigma, p.active_cell.astype(np.int32), p.xVertex, p.yVertex, u, v,
                       p.stagusrf, p.stagthck, p.flwafact, np.array([p.efvs_constant, 0, 0, 0])],
                      [(p.nz - 1, p.nx, p.ny)] + [ashape] * 4)
    st, Ao, efo = O.assemble(p, u, v)
    assert rel_max(outs[0], efo) < 1e-13
    for a, b in zip(outs[1:], Ao):
        assert rel_max(a, b) < 1e-13


def test_fortran_picard(prob, tmp_path):
    p = prob
    lu, lv = O.load_gravity(p)
    vshape = (p.nz, p.nx - 1, p.ny - 1)
    gu, gv, sc = run_driver(
        tmp_path, [5, p.nx, p.ny, p.nz, p.whichprecond, p.whichapprox, p.whichefvs, p.whichsparse],
        [p.sigma, p.active_cell.astype(np.int32), p.active_vertex.astype(np.int32), p.xVertex,
         p.yVertex, p.stagusrf, p.stagthck, p.flwafact, p.beta_internal, p.umask_dirichlet,
         p.vmask_dirichlet, lu, lv, p.uvel, p.vvel,
         np.array([p.dx, p.dy, p.efvs_constant, float(p.whichbabc)])],
        [vshape, vshape, (4,)])
    ro = O.picard(p, lu, lv)
    assert int(sc[0]) == ro["picard_iters"]
    e = rel_l2(np.concatenate([gu.ravel(), gv.ravel()]),
               np.concatenate([ro["uvel"].ravel(), ro["vvel"].ravel()]))
    print(f"  Fortran-driven Picard: {int(sc[0])} iterations, rel L2 diff {e:.2e}")
    assert e < 1e-8


def test_fortran_geometry_solve_with_the_learned_guess(tmp_path):
    """glissade_velo_higher_solve_gpu from a Fortran program: geometry in, the fast path
    (precision 1, inexact Picard, Anderson) from the learned guess read from a weight file,
    under the reference's stopping rule.  Same result as the same call from Python, and fewer
    Picard iterations than from zero."""
    from cism_amd import gpu as G, surrogate as SG
    w = SG.default_weights("grf")
    if not os.path.exists(w):
        pytest.skip(f"no trained weights at {w}")
    wbin = tmp_path / "grf.bin"
    SG.export_weights_bin(w, str(wbin))
    p = configs.grf_sheet(ewn=256, nsn=256, nz=11, dx=1000.0, seed=3100)
    vshape = (p.nz, p.nx - 1, p.ny - 1)
    bext = p.extra["beta_external"]
    runs = {}
    for learned in (0, 1):
        fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
        with open(fin, "wb") as f:
            f.write(np.asarray([6, p.nx, p.ny, p.nz, p.whichprecond, p.whichapprox, p.whichefvs, learned],
                               dtype=np.int32).tobytes())
            for a in (p.sigma, p.thck, p.usrf, p.topg, p.flwa, bext,
                      np.array([p.dx, p.thklim, p.efvs_constant, 1.0])):
                f.write(np.asfortranarray(a, dtype=np.float64).tobytes(order="F"))
        print(f"[fortran] mode 6 learned={learned} grid {p.nx}x{p.ny}x{p.nz}", flush=True)
        r = subprocess.run([DRIVER, str(fin), str(fout), str(wbin)], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        data = np.fromfile(fout, dtype=np.float64)
        n = int(np.prod(vshape))
        runs[learned] = (data[:n].reshape(vshape, order="F"), data[n:2 * n].reshape(vshape, order="F"),
                         data[2 * n:])
    for learned, (u, v, sc) in runs.items():
        print(f"  learned={learned}: {int(sc[0])} Picard, converged {int(sc[2])}, fallbacks {int(sc[3])}")
        assert sc[2] == 1.0 and sc[3] == 0.0
    assert runs[1][2][0] < runs[0][2][0]
    params, meta = SG.load_weights(w)
    with G.Context.for_problem(p) as ctx:
        ctx.learned_guess_load(params, meta["nz"], 0, meta["uscale"])
        rp = ctx.glissade_velo_higher_solve(p, bp=G.BasalPhysics(configs.HO_BABC_BETA_EXTERNAL),
                                            beta_external=bext, initial_guess=1, precision=1,
                                            adaptive_linear=1, inner_rtol=5e-2, anderson_m=3,
                                            coarse=1, skip_final_residual=1)
    u, v, sc = runs[1]
    assert rp["picard_iters"] == int(sc[0])
    assert np.array_equal(rp["uvel"], u) and np.array_equal(rp["vvel"], v)


@pytest.mark.parametrize("approx", [1, 3, 4])
def test_fortran_2d_solvers(tmp_path, approx):
    """glissade_velo_higher_solve_2d_gpu (SSA, L1L2, DIVA) from Fortran against the
    oracle's loops on a GRF sheet: same Picard count, 2D and 3D velocity within 1e-8"""
    if not os.path.exists(DRIVER):
        pytest.fail("Fortran driver not built (run __graft_entry__.build())")
    p = configs.grf_sheet(ewn=24, nsn=20, nz=7, seed=4)
    lu, lv = O.loads(p)
    vshape, s2, c3 = (p.nz, p.nx - 1, p.ny - 1), (p.nx - 1, p.ny - 1), (p.nz - 1, p.nx, p.ny)
    z2 = np.zeros(s2, order="F")
    ipl = O.ice_plus_land(p)
    gu, gv, g2u, g2v, gbx, gby, gef, sc = run_driver(
        tmp_path, [7, p.nx, p.ny, p.nz, p.whichprecond, approx, p.whichefvs, p.whichsparse],
        [p.sigma, p.ice_mask.astype(np.int32), ipl, p.active_cell.astype(np.int32),
         p.active_vertex.astype(np.int32), p.xVertex, p.yVertex, p.thck, p.stagusrf, p.stagthck,
         p.dusrf_dx, p.dusrf_dy, p.flwa, p.flwafact, p.beta_internal, p.umask_dirichlet,
         p.vmask_dirichlet, lu, lv, p.uvel, p.vvel, z2, z2, z2, z2, np.zeros(c3, order="F"),
         np.array([p.dx, p.dy, p.efvs_constant, float(p.whichbabc)])],
        [vshape, vshape, s2, s2, s2, s2, c3, (4,)])
    ro = O.picard_2d(p, lu, lv) if approx == 1 else O.picard_depth(p, lu, lv, approx)
    print(f"  Fortran 2D approx {approx}: {int(sc[0])} Picard (oracle {ro['picard_iters']})")
    assert int(sc[0]) == ro["picard_iters"] and sc[2] == 1.0
    x3 = lambda u, v: np.concatenate([u.ravel(), v.ravel()])   # noqa: E731
    assert rel_l2(x3(gu, gv), x3(ro["uvel"], ro["vvel"])) < 1e-8
    if approx != 1:
        assert rel_l2(x3(g2u, g2v), x3(ro["uvel_2d"], ro["vvel_2d"])) < 1e-8
        assert rel_l2(gef, ro["efvs"]) < 1e-8
