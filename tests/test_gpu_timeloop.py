"""The time-step driver (cism_amd/timeloop.py): the first step's velocity is the direct
solve of the same geometry, a warm-started second step needs fewer Picard iterations
than a cold solve of the same geometry, and the restart file round-trips the state."""
import numpy as np
import pytest

from cism_amd import configs, gpu as G
from cism_amd.timeloop import VelocitySolver
from helpers import rel_l2

pytestmark = pytest.mark.gpu


def sheet():
    p = configs.grf_sheet(ewn=32, nsn=28, nz=7, dx=2000.0, corr=8000.0, seed=6)
    return p, np.asfortranarray(p.beta_internal)


def test_first_step_is_the_direct_solve():
    p, bext = sheet()
    with G.Context.for_problem(p) as ctx:
        lu, lv = ctx.loads(p)
        rd = ctx.velo_higher_solve_3d(p, lu, lv)
    with VelocitySolver(p.ewn, p.nsn, p.nz, p.dx, bext, flwa_value=1.0e-16, thklim=p.thklim,
                        fast=False) as vs:
        r1 = vs.step(p.thck, p.usrf, p.topg)
    assert r1["initial_guess"] == "zero"
    print(f"  direct {rd['picard_iters']} Picard, time step {r1['picard_iters']}")
    # the generator rounds flwafact to 1e-3 (a 2e-9 relative change); the step computes it
    assert abs(r1["picard_iters"] - rd["picard_iters"]) <= 1
    xd = np.concatenate([rd["uvel"].ravel(), rd["vvel"].ravel()])
    x1 = np.concatenate([r1["uvel"].ravel(), r1["vvel"].ravel()])
    assert rel_l2(x1, xd) < 1e-6


def test_warm_start_and_restart(tmp_path):
    p, bext = sheet()
    # step 2: the ice thickens by 2 % in a smooth band
    x = np.linspace(0.0, 2 * np.pi, p.nx)[:, None]
    H2 = np.asfortranarray(p.thck * (1.0 + 0.02 * np.sin(x) ** 2))
    s2 = np.asfortranarray(p.topg + H2)
    with VelocitySolver(p.ewn, p.nsn, p.nz, p.dx, bext, thklim=p.thklim, fast=False) as vs:
        vs.step(p.thck, p.usrf, p.topg)
        path = str(tmp_path / "restart.npz")
        vs.save_restart(path)
        rw = vs.step(H2, s2, p.topg)
        assert rw["initial_guess"] == "previous step"
    with VelocitySolver(p.ewn, p.nsn, p.nz, p.dx, bext, thklim=p.thklim, fast=False) as cold:
        rc = cold.step(H2, s2, p.topg)
    print(f"  warm {rw['picard_iters']} Picard vs cold {rc['picard_iters']}")
    assert rw["picard_iters"] < rc["picard_iters"]
    xw = np.concatenate([rw["uvel"].ravel(), rw["vvel"].ravel()])
    xc = np.concatenate([rc["uvel"].ravel(), rc["vvel"].ravel()])
    assert rel_l2(xw, xc) < 1e-4       # both stop at the reference's Picard tolerance
    # restart: a new solver continues from the saved state exactly as the first did
    with VelocitySolver(p.ewn, p.nsn, p.nz, p.dx, bext, thklim=p.thklim, fast=False) as vr:
        vr.load_restart(path)
        assert vr.steps == 1
        rr = vr.step(H2, s2, p.topg)
    assert rr["picard_iters"] == rw["picard_iters"]
    assert np.array_equal(rr["uvel"], rw["uvel"]) and np.array_equal(rr["vvel"], rw["vvel"])


def test_fast_loop_from_the_learned_guess():
    """fast=True: the first step from the device-computed learned guess, the second warm;
    both reach the reference's Picard tolerance and agree with the FP64 solve of the
    same geometry to it"""
    import os
    from cism_amd import surrogate as SG
    w = SG.default_weights("grf")
    if not os.path.exists(w):
        pytest.skip("no GRF weights")
    p = configs.grf_sheet(ewn=32, nsn=32, nz=11, dx=1000.0, seed=3107)
    bext = np.asfortranarray(p.extra["beta_external"])
    x = np.linspace(0.0, 2 * np.pi, p.nx)[:, None]
    H2 = np.asfortranarray(p.thck * (1.0 + 0.02 * np.sin(x) ** 2))
    s2 = np.asfortranarray(p.topg + H2)
    with VelocitySolver(p.ewn, p.nsn, p.nz, p.dx, bext, thklim=p.thklim, fast=True, weights=w) as vs:
        r1 = vs.step(p.thck, p.usrf, p.topg)
        r2 = vs.step(H2, s2, p.topg)
    with VelocitySolver(p.ewn, p.nsn, p.nz, p.dx, bext, thklim=p.thklim, fast=False) as ref:
        c1 = ref.step(p.thck, p.usrf, p.topg)
        c2 = ref.step(H2, s2, p.topg)
    print(f"  fast: learned {r1['picard_iters']} Picard, warm {r2['picard_iters']}; "
          f"FP64: {c1['picard_iters']}, warm {c2['picard_iters']}")
    assert r1["initial_guess"] == "learned" and r1["guess_rejected"] == 0
    assert r2["initial_guess"] == "previous step"
    for r, c in ((r1, c1), (r2, c2)):
        assert r["converged"] and c["converged"] and r["fallbacks"] == 0
        xr = np.concatenate([r["uvel"].ravel(), r["vvel"].ravel()])
        xc = np.concatenate([c["uvel"].ravel(), c["vvel"].ravel()])
        assert rel_l2(xr, xc) < 1e-4
