// refine_policy.h -- when the mixed-precision iterative refinement is given up.
//
// Host-only decision logic (no HIP), shared by solver.hip and the CPU unit test
// (tests/test_refine_policy.py compiles this header with g++).  The refinement
// x += |r| e, with e from an FP32 Chron-Gear solve of A32 e = r/|r|, is expected to
// cut the FP64 relative residual by roughly the inner tolerance per outer step.
//   - A non-finite residual (the FP32 solve diverged or met a NaN) is a stall.
//   - So is a refinement whose inner solves keep reaching their tolerance while the
//     FP64 residual does not fall: REFINE_WINDOW such outer steps in a row that together
//     do not cut it by REFINE_MIN_GAIN.  The FP32 operator then no longer resolves the
//     FP64 residual (below its rounding floor, when the tolerance is out of FP32's
//     reach), and the Picard iteration falls back to the plain FP64 solver.
//   - An inner solve that ends on its iteration count restarts that window: slow
//     progress on a hard system is not a precision floor.
//   - When the outer-step cap or the inner-iteration budget is spent, the refinement
//     ends unconverged (as the reference's PCG ends at maxiters, not an error).
// Convergence (err < tol) is tested first, so a solve that converges on its last
// permitted step is neither.
#pragma once

#include <math.h>

#define REFINE_MAX_OUTER 50      // outer (FP64 residual) steps per linear solve
#define REFINE_INNER_BUDGET 4    // inner FP32 iterations: at most this many x maxiters
#define REFINE_WINDOW 3          // outer steps over which progress is judged
#define REFINE_MIN_GAIN 0.5      // the residual must at least halve over the window

enum RefineVerdict {
    REFINE_CONTINUE = 0,
    REFINE_CONVERGED = 1,
    REFINE_STALLED = 2,     // give up: fall back to the FP64 solver
    REFINE_EXHAUSTED = 3    // budget spent: end unconverged
};

struct RefineMonitor {
    double hist[REFINE_WINDOW + 1];   // the last residuals after converged inner solves
    int n;                            // residuals in the window
};

inline void refine_monitor_init(RefineMonitor *m) { m->n = 0; }

// err: the FP64 relative residual after `outer` completed outer steps (0 before the
// first inner solve) and `inner_total` FP32 iterations; inner_converged: whether the
// inner solve of the last outer step reached its tolerance (ignored for outer = 0).
inline int refine_decide(RefineMonitor *m, double err, double tol, int outer, int inner_total,
                         int maxiters, int inner_converged) {
    if (!(err == err) || isinf(err)) return REFINE_STALLED;
    if (err < tol) return REFINE_CONVERGED;
    if (outer >= REFINE_MAX_OUTER || inner_total >= REFINE_INNER_BUDGET * maxiters) return REFINE_EXHAUSTED;
    if (outer > 0 && !inner_converged) m->n = 0;   // the window restarts from this residual
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
