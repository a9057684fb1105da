"""The velocity part of a time step, with the device context kept across steps.

In the reference, every time step calls glissade_diagnostic_variable_solve
(libglissade/glissade.F90:1710-2505), which recomputes the masks, the grounded
fraction and beta from the current geometry and then calls the velocity solver from
the previous step's velocity.  `VelocitySolver` does the same on one GPU:

  - one `gpu.Context` for the whole run (no per-step allocation or upload of the
    solver's buffers);
  - per step, one call of cism_glissade_velo_higher_solve with the new thickness,
    surface and bed: the masks, staggered thickness and surface, surface gradient,
    vertex geometry, flow factor, grounded fraction, beta (calcbeta,
    which_ho_babc = 5) and the gravity and lateral loads are computed on the device,
    so the geometry's derived fields never round-trip through the host;
  - the Picard loop starts from the previous step's velocity, which stays resident
    on the device; on the first step from the learned initial guess (computed on the
    device) when weights are given, else from zero;
  - `save_restart` / `load_restart` keep uvel, vvel and efvs (the restart velocity
    fields of glide_vars.def) in an .npz file.

Geometry evolution (transport, calving) is the caller's; this class takes the new
thickness, surface and bed of each step.
"""
from __future__ import annotations

import os
from types import SimpleNamespace

import numpy as np

from . import configs
from . import gpu as G


class VelocitySolver:
    def __init__(self, ewn, nsn, nz, dx, beta_external, flwa_value=1.0e-16, thklim=100.0, eus=0.0,
                 periodic=True, fast=True, weights=None, max_slope=1.0, device=0):
        """beta_external: (nx-1, ny-1) vertex field (Pa yr/m); fast: the mixed-precision
        path with inexact Picard, Anderson(4) and the coarse correction of the FP32 inner
        solves (the bench's settings), else the plain FP64 reference path."""
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
        self._resident = False   # the context holds the last step's velocity
        self.steps = 0

    def close(self):
        self.ctx.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _geometry(self, thck, usrf, topg):
        """The step's inputs to cism_glissade_velo_higher_solve: the geometry, the flow
        factor and the solver options; the masks, staggered fields, grounded fraction,
        beta and loads are computed from them on the device."""
        f = lambda a: np.asfortranarray(a, dtype=np.float64)   # noqa: E731
        z3 = np.zeros((self.nz, self.nx - 1, self.ny - 1), order="F")
        return SimpleNamespace(
            nx=self.nx, ny=self.ny, nz=self.nz, dx=self.dx, dy=self.dx, thklim=self.thklim,
            sigma=self.sigma, thck=f(thck), usrf=f(usrf), topg=f(topg), flwa=self.flwa,
            uvel=z3, vvel=z3, whichbabc=configs.HO_BABC_BETA_EXTERNAL,
            whichefvs=configs.HO_EFVS_NONLINEAR, whichapprox=configs.HO_APPROX_BP,
            whichresid=configs.HO_RESID_L2NORM, whichsparse=configs.HO_SPARSE_PCG_CHRONGEAR,
            whichprecond=configs.HO_PRECOND_SIA, whichassemble_taud=1, whichassemble_beta=1,
            efvs_constant=2336041.42829, maxiter_nonlinear=100)

    def _load_weights(self):
        if self.weights is None or not os.path.exists(self.weights):
            return False
        from . import surrogate as SG
        params, meta = SG.load_weights(self.weights)
        self.ctx.learned_guess_load(params, meta["nz"], 1 if meta.get("target") == "log1p" else 0,
                                    meta["uscale"])
        return True

    def step(self, thck, usrf, topg, stop_rel_err=0.0):
        """Solves the velocity of the new geometry; returns the solve's result dict.  The
        velocity stays on the device between steps: the next step starts from it without
        a transfer."""
        g = self._geometry(thck, usrf, topg)
        u0 = v0 = None
        if self.uvel is None:
            guess, mode = ("learned", 1) if self._load_weights() else ("zero", 0)
        elif self._resident:
            guess, mode = "previous step", 2
        else:                              # from a restart file
            guess, mode, u0, v0 = "previous step", 0, self.uvel, self.vvel
        kw = dict(precision=1, adaptive_linear=1, inner_rtol=1e-2, anderson_m=4, coarse=1,
                  skip_final_residual=1) if self.fast else {}
        r = self.ctx.glissade_velo_higher_solve(
            g, bp=G.BasalPhysics(configs.HO_BABC_BETA_EXTERNAL), beta_external=self.beta_external,
            initial_guess=mode, uvel=u0, vvel=v0, lateral_bc=1, max_slope=self.max_slope, eus=self.eus,
            stop_rel_err=stop_rel_err, **kw)
        self.uvel, self.vvel, self.efvs = r["uvel"], r["vvel"], r["efvs"]
        self._resident = True
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
        self._resident = False
