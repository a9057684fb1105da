"""The velocity part of a time step, with the device context kept across steps.

In the reference, every time step calls glissade_diagnostic_variable_solve
(libglissade/glissade.F90:1710-2505), which recomputes the masks, the grounded
fraction and beta from the current geometry and then calls the velocity solver from
the previous step's velocity.  `VelocitySolver` does the same on one GPU:

  - one `gpu.Context` for the whole run (no per-step allocation or upload of the
    solver's buffers);
  - per step, the solve's inputs from the geometry on the device
    (cism_glissade_velo_inputs: masks, staggered thickness and surface, surface
    gradient, vertex geometry, flow factor), the grounded fraction
    (cism_glissade_grounded_fraction), beta from the external field (calcbeta,
    which_ho_babc = 5) and the gravity and lateral loads;
  - the Picard loop starts from the previous step's velocity; on the first step from
    the learned initial guess when weights are given, else from zero;
  - `save_restart` / `load_restart` keep uvel, vvel and efvs (the restart velocity
    fields of glide_vars.def) in an .npz file.

Geometry evolution (transport, calving) is the caller's; this class takes the new
thickness, surface and bed of each step.
"""
from __future__ import annotations

import os

import numpy as np

from . import configs
from . import gpu as G


class VelocitySolver:
    def __init__(self, ewn, nsn, nz, dx, beta_external, flwa_value=1.0e-16, thklim=100.0, eus=0.0,
                 periodic=True, fast=True, weights=None, max_slope=1.0, device=0):
        """beta_external: (nx-1, ny-1) vertex field (Pa yr/m); fast: the mixed-precision
        path with inexact Picard and Anderson(4), else the plain FP64 reference path."""
        self.ewn, self.nsn, self.nz = ewn, nsn, nz
        self.nx, self.ny = ewn + 2 * configs.NHALO, nsn + 2 * configs.NHALO
        self.dx = float(dx)
        self.thklim, self.eus, self.max_slope = thklim, eus, max_slope
        self.periodic = periodic
        self.fast = fast
        self.beta_external = np.asfortranarray(beta_external, dtype=np.float64)
        self.flwa = np.full((nz - 1, self.nx, self.ny), float(flwa_value), order="F")
        self.sigma = configs.glide_sigma(nz)
        self.ctx = G.Context(self.nx, self.ny, nz, periodic, device)
        self.weights = weights
        self.uvel = self.vvel = self.efvs = None
        self.steps = 0

    def close(self):
        self.ctx.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def problem(self, thck, usrf, topg):
        """The step's Problem, every input computed on the device from the geometry."""
        c = self.ctx
        f = lambda a: np.asfortranarray(a, dtype=np.float64)   # noqa: E731
        o = c.glissade_velo_inputs(self.dx, self.dx, self.thklim, self.eus, f(thck), f(usrf), f(topg),
                                   self.flwa, max_slope=self.max_slope)
        fg = c.grounded_fraction(o["ice_mask"], o["floating_mask"], o["land_mask"], 0)
        s2 = (self.nx - 1, self.ny - 1)
        beta = c.calcbeta(G.BasalPhysics(configs.HO_BABC_BETA_EXTERNAL), np.zeros(s2, order="F"),
                          np.zeros(s2, order="F"), self.beta_external, None, fg)
        z3 = lambda: np.zeros((self.nz, self.nx - 1, self.ny - 1), order="F")   # noqa: E731
        mask = lambda: np.zeros((self.nz, self.nx - 1, self.ny - 1), dtype=np.int32, order="F")  # noqa: E731
        return configs.Problem(
            name="time_step", ewn=self.ewn, nsn=self.nsn, nz=self.nz, dx=self.dx, dy=self.dx,
            periodic=self.periodic, sigma=self.sigma, stagwbndsigma=configs.stagwbnd(self.sigma),
            thck=f(thck), usrf=f(usrf), topg=f(topg), thklim=self.thklim,
            active_cell=o["active_cell"], active_vertex=o["active_vertex"], xVertex=o["xVertex"],
            yVertex=o["yVertex"], stagusrf=o["stagusrf"], stagthck=o["stagthck"],
            dusrf_dx=o["dusrf_dx"], dusrf_dy=o["dusrf_dy"], flwa=self.flwa, flwafact=o["flwafact"],
            beta_internal=beta, umask_dirichlet=mask(), vmask_dirichlet=mask(), uvel=z3(), vvel=z3(),
            whichbabc=configs.HO_BABC_BETA_EXTERNAL, ice_mask=o["ice_mask"],
            floating_mask=o["floating_mask"], ocean_mask=o["ocean_mask"], land_mask=o["land_mask"])

    def step(self, thck, usrf, topg, stop_rel_err=0.0):
        """Solves the velocity of the new geometry; returns the solve's result dict."""
        p = self.problem(thck, usrf, topg)
        lu, lv = self.ctx.loads(p)
        u0, v0 = self.uvel, self.vvel
        guess = "previous step"
        if u0 is None:
            guess = "zero"
            if self.weights is not None and os.path.exists(self.weights):
                from . import surrogate as SG
                pred = SG.Predictor(p, self.weights)
                try:
                    u0, v0 = pred.predict(p)
                finally:
                    pred.close()
                guess = "learned"
        kw = dict(precision=1, adaptive_linear=1, inner_rtol=5e-2, anderson_m=3) if self.fast else {}
        r = self.ctx.velo_higher_solve_3d(p, lu, lv, uvel=u0, vvel=v0, stop_rel_err=stop_rel_err, **kw)
        self.uvel, self.vvel, self.efvs = r["uvel"], r["vvel"], r["efvs"]
        self.steps += 1
        r["initial_guess"] = guess
        return r

    def save_restart(self, path):
        np.savez(path, uvel=self.uvel, vvel=self.vvel, efvs=self.efvs, steps=np.int64(self.steps))

    def load_restart(self, path):
        with np.load(path, allow_pickle=False) as z:
            self.uvel = np.asfortranarray(z["uvel"])
            self.vvel = np.asfortranarray(z["vvel"])
            self.efvs = np.asfortranarray(z["efvs"])
            self.steps = int(z["steps"])
