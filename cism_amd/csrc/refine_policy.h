// refine_policy.h -- when the mixed-precision iterative refinement is given up.
//
// Host-only decision logic (no HIP), shared by solver.hip and the CPU unit test
// (tests/test_refine_policy.py compiles this header with g++).  The refinement
// x += |r| e, with e from an FP32 Chron-Gear solve of A32 e = r/|r|, is expected to
// cut the FP64 relative residual by roughly the inner tolerance per outer step.  It
// is declared stalled, and the Picard iteration falls back to the plain FP64 solver,
// when
//   - the residual is not finite (the FP32 solve diverged or met a NaN),
//   - the outer-step cap or the inner-iteration budget is spent, or
//   - the last REFINE_WINDOW outer steps together did not cut the residual by
//     REFINE_MIN_GAIN (the FP32 operator no longer resolves the FP64 residual, for
//     example below its rounding floor when the tolerance is out of FP32's reach).
// Convergence (err < tol) is tested first, so a solve that converges on its last
// permitted step is not a stall.
#pragma once

#include <math.h>

#define REFINE_MAX_OUTER 50      // outer (FP64 residual) steps per linear solve
#define REFINE_INNER_BUDGET 4    // inner FP32 iterations: at most this many x maxiters
#define REFINE_WINDOW 3          // outer steps over which progress is judged
#define REFINE_MIN_GAIN 0.5      // the residual must at least halve over the window

enum RefineVerdict { REFINE_CONTINUE = 0, REFINE_CONVERGED = 1, REFINE_STALLED = 2 };

struct RefineMonitor {
    double hist[REFINE_WINDOW + 1];   // the last REFINE_WINDOW + 1 relative residuals
    int n;                            // residuals seen so far
};

inline void refine_monitor_init(RefineMonitor *m) { m->n = 0; }

// err: the FP64 relative residual after `outer` completed outer steps (0 before the
// first inner solve) and `inner_total` FP32 iterations; returns the verdict.
inline int refine_decide(RefineMonitor *m, double err, double tol, int outer, int inner_total,
                         int maxiters) {
    if (!(err == err) || isinf(err)) return REFINE_STALLED;
    if (err < tol) return REFINE_CONVERGED;
    if (outer >= REFINE_MAX_OUTER || inner_total >= REFINE_INNER_BUDGET * maxiters) return REFINE_STALLED;
    // shift the window
    if (m->n <= REFINE_WINDOW) {
        m->hist[m->n] = err;
    } else {
        for (int t = 0; t < REFINE_WINDOW; ++t) m->hist[t] = m->hist[t + 1];
        m->hist[REFINE_WINDOW] = err;
    }
    m->n++;
    if (m->n > REFINE_WINDOW) {
        const double before = m->hist[0], now = m->hist[REFINE_WINDOW];
        if (now > REFINE_MIN_GAIN * before) return REFINE_STALLED;
    }
    return REFINE_CONTINUE;
}
