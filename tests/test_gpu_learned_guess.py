"""What the trained network gains: the same solves from a zero initial guess and from the
network's prediction, on a held-out problem, each run until its relative L2 velocity error
against the FP64 reference is <= 1e-6 (the plain FP64 path stays the arbiter of the error)."""
import os
import time

import numpy as np
import pytest

from cism_amd import configs, gpu as G, surrogate as SG
from cism_amd.fp64_reference import reference_solution

pytestmark = pytest.mark.gpu

WEIGHTS = SG.default_weights("grf")
N = 256
SEED = 3100          # outside the training (1000..2999) and validation (3000..3007) seeds


def rel_l2(u, v, ru, rv):
    """relative L2 difference over the owned vertices (halo entries are periodic copies)"""
    o = (slice(None), slice(2, -2), slice(2, -2))
    u, v, ru, rv = u[o], v[o], ru[o], rv[o]
    return float(np.sqrt(((u - ru) ** 2).sum() + ((v - rv) ** 2).sum())
                 / np.sqrt((ru ** 2).sum() + (rv ** 2).sum()))


@pytest.fixture(scope="module")
def setup():
    if not os.path.exists(WEIGHTS):
        pytest.skip(f"no trained weights at {WEIGHTS}")
    p = configs.grf_sheet(ewn=N, nsn=N, nz=11, dx=1000.0, seed=SEED)
    print(f"[learned] grf_sheet {N}x{N}x11 seed {SEED}", flush=True)
    ctx = G.Context.for_problem(p)
    lu, lv = ctx.loads(p)
    ref = reference_solution(ctx, p, lu, lv)
    ctx.set_reference_velocity(ref["uvel"], ref["vvel"])
    pred = SG.Predictor(p, WEIGHTS)
    t0 = time.perf_counter()
    gu, gv = pred.predict(p)
    t_inf = time.perf_counter() - t0
    pred.close()
    yield p, ctx, lu, lv, ref, gu, gv, t_inf
    ctx.close()


def test_prediction_is_close(setup):
    p, ctx, lu, lv, ref, gu, gv, t_inf = setup
    e = rel_l2(gu, gv, ref["uvel"], ref["vvel"])
    print(f"[learned] prediction rel L2 error {e:.4f} (inference {t_inf * 1e3:.1f} ms)")
    assert e < 0.25


@pytest.mark.parametrize("precision", [1, 0])
def test_learned_guess_reduces_work(setup, precision):
    p, ctx, lu, lv, ref, gu, gv, t_inf = setup
    kw = dict(precision=precision, stop_rel_err=1e-6)
    if precision == 1:
        kw.update(adaptive_linear=1, inner_rtol=1e-3)
    runs = {}
    for name, (u0, v0) in {"zero": (p.uvel, p.vvel), "learned": (gu, gv)}.items():
        t0 = time.perf_counter()
        r = ctx.velo_higher_solve_3d(p, lu, lv, uvel=u0, vvel=v0, **kw)
        r["seconds"] = time.perf_counter() - t0
        runs[name] = r
        print(f"[learned] precision {precision} {name}: {r['picard_iters']} Picard, "
              f"{r['fp32_matvecs']} FP32 + {r['fp64_matvecs']} FP64 matvecs, "
              f"{r['seconds']:.3f} s, err {r['final_rel_err']:.2e}")
    for r in runs.values():
        assert r["final_rel_err"] <= 1e-6
        assert r["guess_rejected"] == 0 and r["fallbacks"] == 0
        assert rel_l2(r["uvel"], r["vvel"], ref["uvel"], ref["vvel"]) <= 1.01e-6
    z, l = runs["zero"], runs["learned"]
    assert l["picard_iters"] < z["picard_iters"]
    if precision == 1:
        assert l["fp32_matvecs"] + l["fp64_matvecs"] < z["fp32_matvecs"] + z["fp64_matvecs"]
    else:
        assert l["linear_iters"] < z["linear_iters"]
