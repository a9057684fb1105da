"""cism_glissade_velo_higher_solve (the solve from the ice geometry, every input on the
device) and the device-side learned guess (cism_learned_guess*), against the host-prepared
path that the parity tests pin to the oracle."""
import os

import numpy as np
import pytest

from cism_amd import configs, gpu as G, surrogate as SG
from cism_amd.fp64_reference import reference_solution
from helpers import rel_l2

pytestmark = pytest.mark.gpu


def _bp(p):
    return G.BasalPhysics(configs.HO_BABC_BETA_EXTERNAL)


def _x(r):
    return np.concatenate([r["uvel"].ravel(), r["vvel"].ravel()])


@pytest.mark.parametrize("name", ["sheet", "shelf", "grf", "outflow"])
def test_geometry_solve_equals_the_host_prepared_solve(name):
    """the same Picard loop from the same geometry: inputs, grounded fraction, calcbeta and
    loads computed on the device give the velocities of the solve from host-prepared
    inputs"""
    p = {"sheet": lambda: configs.ice_sheet(ewn=36, nsn=42, nz=6, dx=33333.0, seed=7),
         "shelf": lambda: configs.ice_sheet(ewn=36, nsn=42, nz=6, dx=33333.0, seed=7, shelf=0.02),
         "grf": lambda: configs.grf_sheet(ewn=24, nsn=20, nz=7, dx=2000.0, corr=8000.0, seed=3),
         "outflow": lambda: configs.grf_sheet(ewn=24, nsn=20, nz=7, dx=2000.0, corr=8000.0, seed=3,
                                              periodic=False)}[name]()
    print(f"[velo_solve] {p.name} {p.ewn}x{p.nsn}x{p.nz}", flush=True)
    with G.Context.for_problem(p) as ctx:
        # the generator rounds its flow factor to a millimetre grid; the geometry path takes
        # flwafact = 0.5 A^(-1/3) from flwa itself, so the host-prepared run does too
        d = ctx.glissade_velo_inputs(p.dx, p.dy, p.thklim, 0.0, p.thck, p.usrf, p.topg, p.flwa)
        p.flwafact = d["flwafact"]
        lu, lv = ctx.loads(p)
        rh = ctx.velo_higher_solve_3d(p, lu, lv)
    with G.Context.for_problem(p) as ctx:
        rg = ctx.glissade_velo_higher_solve(p, bp=_bp(p), beta_external=p.extra["beta_external"])
    e = rel_l2(_x(rg), _x(rh))
    print(f"  host-prepared {rh['picard_iters']} Picard, geometry {rg['picard_iters']} Picard, diff {e:.2e}")
    assert rg["picard_iters"] == rh["picard_iters"] and rg["converged"]
    assert e < 1e-12
    assert np.max(np.abs(rg["efvs"] - rh["efvs"])) <= 1e-12 * np.max(np.abs(rh["efvs"]))


@pytest.fixture(scope="module")
def grf256():
    w = SG.default_weights("grf")
    if not os.path.exists(w):
        pytest.skip(f"no trained weights at {w}")
    params, meta = SG.load_weights(w)
    p = configs.grf_sheet(ewn=256, nsn=256, nz=11, dx=1000.0, seed=3100)
    return p, params, meta


def test_device_guess_equals_the_host_predictor(grf256):
    """features, network and decode on the device give the host Predictor's guess (the
    host decodes float32 outputs in float32; features may differ in the last float bit)"""
    p, params, meta = grf256
    pred = SG.Predictor(p, SG.default_weights("grf"))
    hu, hv = pred.predict(p)
    pred.close()
    with G.Context.for_problem(p) as ctx:
        ctx.learned_guess_load(params, meta["nz"], 1 if meta.get("target") == "log1p" else 0,
                               meta["uscale"])
        du, dv = ctx.learned_guess(p)
        ms = ctx.learned_guess_last_ms()
    e = rel_l2(np.concatenate([du.ravel(), dv.ravel()]), np.concatenate([hu.ravel(), hv.ravel()]))
    print(f"[guess] device vs host guess rel L2 {e:.2e}, forward {ms:.2f} ms")
    assert e < 1e-5


def test_geometry_solve_from_the_learned_guess(grf256):
    """the headline path from the geometry: learned guess on the device, mixed precision,
    Anderson; it reaches 1e-6 of the FP64 reference with fewer Picard iterations than from
    zero, and the guess is not rejected"""
    p, params, meta = grf256
    with G.Context.for_problem(p) as ctx:
        lu, lv = ctx.loads(p)
        ref = reference_solution(ctx, p, lu, lv)
        ctx.set_reference_velocity(ref["uvel"], ref["vvel"])
        ctx.learned_guess_load(params, meta["nz"], 1 if meta.get("target") == "log1p" else 0,
                               meta["uscale"])
        kw = dict(precision=1, adaptive_linear=1, inner_rtol=5e-2, anderson_m=3, stop_rel_err=1e-6,
                  maxiter_nonlinear=300)
        bext = p.extra["beta_external"]
        rz = ctx.glissade_velo_higher_solve(p, bp=_bp(p), beta_external=bext, initial_guess=0, **kw)
        rl = ctx.glissade_velo_higher_solve(p, bp=_bp(p), beta_external=bext, initial_guess=1, **kw)
    print(f"[guess] geometry solve: zero {rz['picard_iters']} Picard {rz['seconds_total']:.3f} s, learned "
          f"{rl['picard_iters']} Picard {rl['seconds_total']:.3f} s (device {rl['seconds_device']:.3f} s)")
    for r in (rz, rl):
        assert r["converged"] and 0 <= r["final_rel_err"] <= 1e-6 and r["fallbacks"] == 0
    assert rl["guess_rejected"] == 0
    assert rl["picard_iters"] < rz["picard_iters"]


def test_learned_guess_on_an_outflow_grid():
    """the sheet network on an outflow grid (owned vertices in a zero-feature canvas): the
    guess is accepted and the fast solve from it reaches 1e-6 of the FP64 reference"""
    w = SG.default_weights("sheet")
    if not os.path.exists(w):
        pytest.skip(f"no trained weights at {w}")
    params, meta = SG.load_weights(w)
    p = configs.ice_sheet(ewn=200, nsn=240, nz=11, dx=6000.0, seed=9, periodic=False)
    print(f"[guess] outflow ice sheet {p.ewn}x{p.nsn}x{p.nz}", flush=True)
    with G.Context.for_problem(p) as ctx:
        lu, lv = ctx.loads(p)
        ref = reference_solution(ctx, p, lu, lv)
        ctx.set_reference_velocity(ref["uvel"], ref["vvel"])
        ctx.learned_guess_load(params, meta["nz"], 1 if meta.get("target") == "log1p" else 0,
                               meta["uscale"])
        kw = dict(precision=1, adaptive_linear=1, inner_rtol=5e-2, anderson_m=3, stop_rel_err=1e-6,
                  maxiter_nonlinear=300)
        bext = p.extra["beta_external"]
        rz = ctx.glissade_velo_higher_solve(p, bp=_bp(p), beta_external=bext, initial_guess=0, **kw)
        rl = ctx.glissade_velo_higher_solve(p, bp=_bp(p), beta_external=bext, initial_guess=1, **kw)
    print(f"[guess] outflow: zero {rz['picard_iters']} Picard, learned {rl['picard_iters']} Picard "
          f"(rejected {rl['guess_rejected']}), err {rl['final_rel_err']:.2e}")
    assert rl["guess_rejected"] == 0
    for r in (rz, rl):
        assert r["converged"] and 0 <= r["final_rel_err"] <= 1e-6


def test_persistent_velocity_arrays(grf256):
    """out=(uvel, vvel): the solve writes the caller's arrays in place (a Fortran caller's
    model arrays) with the same solution as into fresh arrays; a zero-guess solve reads
    them as its start; arrays of the wrong shape are refused before any device work"""
    p, params, meta = grf256
    shape = (p.nz, p.nx - 1, p.ny - 1)
    with G.Context.for_problem(p) as ctx:
        ctx.learned_guess_load(params, meta["nz"], 1 if meta.get("target") == "log1p" else 0,
                               meta["uscale"])
        kw = dict(bp=_bp(p), beta_external=p.extra["beta_external"], precision=1, adaptive_linear=1,
                  inner_rtol=1e-2, anderson_m=4, coarse=1, skip_final_residual=1, efvs=False)
        fresh = ctx.glissade_velo_higher_solve(p, initial_guess=1, **kw)
        U, V = np.full(shape, 7.0, order="F"), np.full(shape, -7.0, order="F")
        kept = ctx.glissade_velo_higher_solve(p, initial_guess=1, out=(U, V), **kw)
        assert kept["uvel"] is U and kept["vvel"] is V
        assert np.array_equal(U, fresh["uvel"]) and np.array_equal(V, fresh["vvel"])
        assert kept["picard_iters"] == fresh["picard_iters"]
        # initial_guess 0 reads the arrays: starting from the solution, the first Picard
        # residual is already below the reference's tolerance
        again = ctx.glissade_velo_higher_solve(p, initial_guess=0, out=(U, V), **kw)
        assert again["picard_iters"] <= 2
        with pytest.raises(ValueError):
            ctx.glissade_velo_higher_solve(p, initial_guess=1, out=(U[:-1], V), **kw)
        with pytest.raises(ValueError):
            ctx.glissade_velo_higher_solve(p, initial_guess=1, out=(np.zeros(shape), V), **kw)
