"""The decision logic of the mixed-precision refinement (cism_amd/csrc/refine_policy.h),
compiled on the host with g++: when a linear solve converges, carries on, or is given up
so that the Picard iteration falls back to the FP64 solver."""
import ctypes as C
import os
import subprocess

import pytest

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cism_amd", "csrc",
                   "refine_policy.h")
SRC = r"""
#include "%s"
extern "C" int run(const double *errs, const int *conv, int n, double tol, int inner_step, int maxiters,
                   int *steps) {
    RefineMonitor m;
    refine_monitor_init(&m);
    for (int t = 0; t < n; ++t) {
        const int v = refine_decide(&m, errs[t], tol, t, t * inner_step, maxiters, conv[t]);
        if (v != REFINE_CONTINUE) { *steps = t; return v; }
    }
    *steps = n;
    return REFINE_CONTINUE;
}
"""
CONTINUE, CONVERGED, STALLED, EXHAUSTED = 0, 1, 2, 3


@pytest.fixture(scope="module")
def policy(tmp_path_factory):
    d = tmp_path_factory.mktemp("refine_policy")
    src, so = d / "p.cpp", d / "p.so"
    src.write_text(SRC % HDR)
    subprocess.run(["g++", "-O2", "-shared", "-fPIC", "-std=c++17", str(src), "-o", str(so)], check=True)
    L = C.CDLL(str(so))
    L.run.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_int, C.c_double, C.c_int, C.c_int,
                      C.POINTER(C.c_int)]

    def run(errs, tol=1e-8, inner_step=10, maxiters=200, conv=None):
        """conv[t]: whether the inner solve before residual t reached its tolerance"""
        conv = [1] * len(errs) if conv is None else conv
        a = (C.c_double * len(errs))(*errs)
        c = (C.c_int * len(errs))(*conv)
        steps = C.c_int()
        v = L.run(a, c, len(errs), tol, inner_step, maxiters, C.byref(steps))
        return v, steps.value
    return run


def test_healthy_refinement_converges(policy):
    errs = [1e-1 * 0.05 ** t for t in range(10)]        # inner tolerance 5e-2 per step
    v, n = policy(errs)
    assert v == CONVERGED and errs[n] < 1e-8 and errs[n - 1] >= 1e-8


def test_converging_on_the_last_permitted_step_is_not_a_stall(policy):
    # four flat residuals (a full window) are a stall; a fall below tol on the fourth is not
    assert policy([1.0] * 4 + [1e-9], inner_step=0) == (STALLED, 3)
    assert policy([1.0] * 3 + [1e-9], inner_step=0) == (CONVERGED, 3)
    # convergence on outer step 50, where the cap would end the refinement
    errs = [0.75 ** t for t in range(50)] + [1e-9]     # 0.42 per window: steady progress
    assert policy(errs, inner_step=1) == (CONVERGED, 50)


def test_rounding_floor_is_a_stall(policy):
    # an FP64 tolerance below what the FP32 operator can resolve: the residual drops to
    # its floor and stays there; judged over a window of three outer steps
    errs = [1e-2, 1e-6, 1e-10, 3e-14, 2.9e-14, 3.1e-14, 3e-14, 2.8e-14]
    v, n = policy(errs, tol=1e-20)
    assert v == STALLED and n == 6


def test_slow_but_steady_progress_continues(policy):
    errs = [0.7 ** t for t in range(12)]                # 0.343 per window of three: not a stall
    v, n = policy(errs, tol=1e-12, inner_step=1)
    assert v == CONTINUE and n == 12


def test_non_finite_residual_is_a_stall(policy):
    assert policy([1e-2, float("nan")]) == (STALLED, 1)
    assert policy([1e-2, float("inf")]) == (STALLED, 1)


def test_budgets_end_unconverged(policy):
    # 50 outer steps
    v, n = policy([0.75 ** t for t in range(80)], tol=1e-30, inner_step=1)
    assert (v, n) == (EXHAUSTED, 50)
    # 4 x maxiters inner iterations: reached at outer step 8 with 100 per step, maxiters 200
    v, n = policy([0.5 ** t for t in range(20)], tol=1e-30, inner_step=100, maxiters=200)
    assert (v, n) == (EXHAUSTED, 8)


def test_slow_inner_solves_are_not_a_precision_floor(policy):
    # inner solves ending on their iteration count with little progress: the window
    # restarts after each of them, and the refinement runs to its budget instead
    errs = [1.0, 0.95, 0.9, 0.86, 0.82, 0.78, 0.74, 0.7, 0.66]
    conv = [1] + [0] * (len(errs) - 1)
    assert policy(errs, tol=1e-8, inner_step=200, maxiters=200, conv=conv) == (EXHAUSTED, 4)
    # the same residuals after converged inner solves are a stall
    assert policy(errs, tol=1e-8, inner_step=1, maxiters=200) == (STALLED, 3)
