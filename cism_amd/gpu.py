"""ctypes binding of libcism_gpu.so (the C ABI declared in include/cism_gpu.h).

The Python names and argument order follow the reference's Fortran
subroutines (glissade_velo_higher_pcg.F90, glissade_velo_higher.F90), with the
context first.  Arrays are numpy arrays in the reference's Fortran layouts;
arguments the reference declares intent(inout)/intent(out) are modified in
place, as in the reference.

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libcism_gpu.so")

CISM_OK = 0
STATUS_NAMES = {1: "CISM_ERR_HIP", 2: "CISM_ERR_SINGULAR", 3: "CISM_ERR_ASYMMETRIC",
                4: "CISM_ERR_NAN", 5: "CISM_ERR_ARG", 6: "CISM_ERR_NOMEM", 7: "CISM_ERR_UNSUPPORTED"}

EXPORTED = [
    "cism_ctx_create", "cism_ctx_destroy", "cism_last_error", "cism_ctx_set_linear_solver",
    "cism_matvec_multiply_structured_3d", "cism_pcg_solver_standard_3d",
    "cism_pcg_solver_chrongear_3d", "cism_assemble_stiffness_matrix_3d", "cism_basal_sliding_bc_3d",
    "cism_dirichlet_boundary_conditions_3d", "cism_check_symmetry_assembled_matrix_3d",
    "cism_count_nonzeros_3d", "cism_compute_residual_vector_3d", "cism_compute_residual_velocity_3d",
    "cism_load_vector_gravity", "cism_staggered_parallel_halo", "cism_solve_opts_default",
    "cism_velo_higher_solve_3d", "cism_velo_higher_solve_2d", "cism_velo_higher_solve_depth", "cism_set_reference_velocity", "cism_time_kernel",
    "cism_coarse_build", "cism_coarse_level", "cism_coarse_correction",
    "cism_load_vector_lateral_bc",
    "cism_set_precond_scale", "cism_column_scale_target", "cism_set_fp64_symmetric",
    "cism_compute_internal_stress", "cism_compute_basal_friction_heatflx",
    "cism_compute_basal_traction", "cism_velo_diagnostics", "cism_glissade_velo_inputs",
    "cism_glissade_grounded_fraction", "cism_calcbeta", "cism_set_basal_physics",
    "cism_surrogate_create", "cism_surrogate_destroy", "cism_surrogate_num_params",
    "cism_surrogate_get_params", "cism_surrogate_set_params", "cism_surrogate_set_channel_weights",
    "cism_surrogate_forward", "cism_surrogate_loss_grad", "cism_surrogate_train_step",
    "cism_surrogate_last_forward_ms", "cism_glissade_velo_higher_solve", "cism_learned_guess_load",
    "cism_learned_guess_load_file", "cism_learned_guess", "cism_learned_guess_last_ms",
    "cism_compute_residual_vector_3d_noalloc", "cism_basal_sliding_bc",
]


class CismError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class SolveOpts(C.Structure):
    _fields_ = [("whichbabc", C.c_int), ("whichefvs", C.c_int), ("whichresid", C.c_int),
                ("whichsparse", C.c_int), ("whichprecond", C.c_int), ("whichapprox", C.c_int),
                ("maxiter_nonlinear", C.c_int), ("precision", C.c_int),
                ("efvs_constant", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
                ("target_override", C.c_double), ("inner_rtol", C.c_double),
                ("adaptive_linear", C.c_int), ("adaptive_gamma", C.c_double),
                ("stop_rel_err", C.c_double), ("anderson_m", C.c_int), ("initial_guess", C.c_int),
                ("whichassemble_beta", C.c_int), ("coarse", C.c_int), ("skip_final_residual", C.c_int)]


class VeloSetup(C.Structure):
    """cism_velo_setup: what cism_glissade_velo_higher_solve needs besides the geometry"""
    _fields_ = [("dx", C.c_double), ("dy", C.c_double), ("thklim", C.c_double), ("eus", C.c_double),
                ("max_slope", C.c_double), ("whichground", C.c_int), ("whichassemble_taud", C.c_int),
                ("lateral_bc", C.c_int), ("initial_guess", C.c_int), ("reserved", C.c_int * 4)]


class BasalPhysics(C.Structure):
    """cism_basal_physics: the basal_physics fields calcbeta reads, reference defaults"""
    _fields_ = [("whichbabc", C.c_int), ("reserved", C.c_int)] + \
               [(n, C.c_double) for n in ("ho_beta_const", "ho_beta_large", "powerlaw_c",
                                          "powerlaw_m", "coulomb_c", "beta_grounded_min",
                                          "beta_powerlaw_umax")]

    def __init__(self, whichbabc=5, **kw):
        d = dict(ho_beta_const=1000.0, ho_beta_large=1.0e10, powerlaw_c=1.0e4, powerlaw_m=3.0,
                 coulomb_c=0.42, beta_grounded_min=1.0, beta_powerlaw_umax=0.0)
        d.update(kw)
        super().__init__(whichbabc=whichbabc, **d)


class SolveStats(C.Structure):
    _fields_ = [("picard_iters", C.c_int), ("linear_iters", C.c_int), ("fp64_matvecs", C.c_int),
                ("fp32_matvecs", C.c_int), ("final_L2_norm", C.c_double),
                ("final_resid_velo", C.c_double), ("seconds_assembly", C.c_double),
                ("seconds_linear", C.c_double), ("seconds_total", C.c_double),
                ("status", C.c_int), ("refinement_steps", C.c_int),
                ("final_rel_err", C.c_double), ("seconds_device", C.c_double),
                ("converged", C.c_int), ("fallbacks", C.c_int), ("guess_rejected", C.c_int),
                ("unconverged_linear", C.c_int)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_ if n != "reserved"}


_lib = None
D = C.POINTER(C.c_double)
I32 = C.POINTER(C.c_int32)
I64 = C.POINTER(C.c_int64)
VP = C.c_void_p


def lib():
    """Loads libcism_gpu.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m cism_amd.build` "
                               "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name in EXPORTED:
            getattr(L, name)   # every declared symbol must be exported
        L.cism_last_error.restype = C.c_char_p
        L.cism_ctx_create.argtypes = [C.c_int] * 6 + [C.POINTER(VP)]
        L.cism_ctx_destroy.argtypes = [VP]
        L.cism_ctx_set_linear_solver.argtypes = [VP, C.c_int, C.c_double]
        L.cism_matvec_multiply_structured_3d.argtypes = [VP, I32, I32, D, D, D, D, D, D, D, D]
        pcg = [VP, I32, I32, D, D, D, D, D, D, D, D, C.c_int, D, C.POINTER(C.c_int)]
        L.cism_pcg_solver_standard_3d.argtypes = pcg
        L.cism_pcg_solver_chrongear_3d.argtypes = pcg
        L.cism_assemble_stiffness_matrix_3d.argtypes = [VP, D, I32, D, D, D, D, D, D, D, C.c_int, D,
                                                        C.c_int, C.c_double, D, D, D, D]
        L.cism_basal_sliding_bc_3d.argtypes = [VP, C.c_double, C.c_double, I32, D, C.c_int, D, D]
        L.cism_basal_sliding_bc.argtypes = [VP, C.c_double, C.c_double, I32, I32, D, D, D, C.c_int, D, D]
        L.cism_dirichlet_boundary_conditions_3d.argtypes = [VP, I32, I32, I32, D, D, D, D, D, D,
                                                            D, D]
        L.cism_check_symmetry_assembled_matrix_3d.argtypes = [VP, I32, D, D, D, D, I64]
        L.cism_count_nonzeros_3d.argtypes = [VP, D, D, D, D, I32, I64]
        L.cism_compute_residual_vector_3d.argtypes = [VP, I32, D, D, D, D, D, D, D, D, D, D, D, D]
        L.cism_compute_residual_velocity_3d.argtypes = [VP, C.c_int, D, D, D, D, D]
        L.cism_load_vector_gravity.argtypes = [VP, D, D, C.c_double, C.c_double, I32, I32, D, D,
                                               D, D, D, D, C.c_int, D, D]
        L.cism_staggered_parallel_halo.argtypes = [VP, C.c_int, D]
        L.cism_load_vector_lateral_bc.argtypes = [VP, D, I32, I32, I32, I32, D, D, D, D, D, D]
        L.cism_solve_opts_default.argtypes = [C.POINTER(SolveOpts)]
        L.cism_solve_opts_default.restype = None
        L.cism_velo_higher_solve_3d.argtypes = [VP, C.POINTER(SolveOpts), D, I32, I32, D, D, D, D,
                                                D, D, I32, I32, D, D, D, D, D, D, D, D, D,
                                                C.POINTER(SolveStats)]
        L.cism_velo_higher_solve_2d.argtypes = [VP, C.POINTER(SolveOpts), C.c_int, D, I32, I32, D, D,
                                                D, D, D, D, I32, I32, D, D, D, D, D, D, D, D, D,
                                                C.POINTER(SolveStats)]
        L.cism_velo_higher_solve_depth.argtypes = [VP, C.POINTER(SolveOpts), C.c_int, D, I32, I32, I32, I32,
                                                   D, D, D, D, D, D, D, D, D, I32, I32, D, D, D, D, D, D,
                                                   D, D, D, D, D, D, D, C.POINTER(SolveStats)]
        L.cism_set_reference_velocity.argtypes = [VP, D, D]
        L.cism_time_kernel.argtypes = [VP, C.c_int, C.c_int, D]
        L.cism_coarse_build.argtypes = [VP, I32, I32, I32, D, D, D, D, C.POINTER(C.c_int)]
        L.cism_coarse_level.argtypes = [VP, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), D]
        L.cism_coarse_correction.argtypes = [VP, D, D, D, D]
        L.cism_set_precond_scale.argtypes = [VP, D, D]
        L.cism_set_fp64_symmetric.argtypes = [VP, C.c_int]
        L.cism_compute_internal_stress.argtypes = [VP, D, I32, D, D, D, D, D, D, C.c_int, C.c_double,
                                                   C.c_int, D, D, D, D, D, D, D, D]
        L.cism_compute_basal_friction_heatflx.argtypes = [VP, I32, I32, D, D, D, D, D, C.c_int, D]
        L.cism_compute_basal_traction.argtypes = [VP, D, D, D, D, D]
        L.cism_velo_diagnostics.argtypes = [VP, C.c_int, C.c_int, C.c_double, C.c_int] + [D] * 9
        L.cism_glissade_velo_inputs.argtypes = [VP] + [C.c_double] * 5 + [D] * 4 + [I32] * 6 + [D] * 7
        L.cism_column_scale_target.argtypes = [VP, D, D]
        L.cism_glissade_grounded_fraction.argtypes = [VP, I32, I32, I32, C.c_int, D]
        L.cism_calcbeta.argtypes = [VP, C.POINTER(BasalPhysics), D, D, D, D, D, D]
        L.cism_set_basal_physics.argtypes = [VP, C.POINTER(BasalPhysics), D, D, D]
        L.cism_glissade_velo_higher_solve.argtypes = [VP, C.POINTER(SolveOpts), C.POINTER(VeloSetup), D, D, D,
                                                      D, D, C.POINTER(BasalPhysics), D, D, D, D, D,
                                                      C.POINTER(SolveStats)]
        L.cism_learned_guess_load.argtypes = [VP, C.POINTER(C.c_float), C.c_longlong, C.c_int, C.c_int,
                                              C.c_double]
        L.cism_learned_guess_load_file.argtypes = [VP, C.c_char_p]
        L.cism_learned_guess.argtypes = [VP, D, D, D, D, D, I32, D, D]
        L.cism_learned_guess_last_ms.argtypes = [VP]
        L.cism_learned_guess_last_ms.restype = C.c_double
        _lib = L
    return _lib


def _check(st):
    if st != CISM_OK:
        raise CismError(st, lib().cism_last_error().decode())


def dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim <= 1), "need F-ordered float64"
    return a.ctypes.data_as(D)


def ip(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and (a.flags.f_contiguous or a.ndim <= 1), "need F-ordered int32"
    return a.ctypes.data_as(I32)


def standard_indxA():
    ind = np.zeros((3, 3, 3), dtype=np.int32, order="F")
    m = 0
    for k in range(3):
        for j in range(3):
            for i in range(3):
                m += 1
                ind[i, j, k] = m
    return ind


class Context:
    """Owns one device context (grid + device buffers), like glissade_velo_higher_init."""

    def __init__(self, nx, ny, nz, periodic=True, device=0, nhalo=2):
        self.nx, self.ny, self.nz = nx, ny, nz
        h = VP()
        _check(lib().cism_ctx_create(nx, ny, nz, nhalo, 1 if periodic else 0, device, C.byref(h)))
        self.h = h
        self.indxA = standard_indxA()

    @classmethod
    def for_problem(cls, p, device=0):
        return cls(p.nx, p.ny, p.nz, p.periodic, device)

    @classmethod
    def for_problem_2d(cls, p, device=0):
        """a one-level context for the SSA solve of problem p"""
        return cls(p.nx, p.ny, 1, p.periodic, device)

    def close(self):
        if self.h:
            lib().cism_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_linear_solver(self, maxiters=200, tolerance=1e-8):
        _check(lib().cism_ctx_set_linear_solver(self.h, maxiters, tolerance))

    # --- glissade_velo_higher_pcg ------------------------------------------------
    def matvec_multiply_structured_3d(self, active_vertex, Auu, Auv, Avu, Avv, xu, xv, yu, yv):
        _check(lib().cism_matvec_multiply_structured_3d(
            self.h, ip(self.indxA), ip(active_vertex), dp(Auu), dp(Auv), dp(Avu), dp(Avv),
            dp(xu), dp(xv), dp(yu), dp(yv)))

    def _pcg(self, fn, active_vertex, Auu, Auv, Avu, Avv, bu, bv, xu, xv, precond):
        err, nit = C.c_double(), C.c_int()
        _check(fn(self.h, ip(self.indxA), ip(active_vertex), dp(Auu), dp(Auv), dp(Avu), dp(Avv),
                  dp(bu), dp(bv), dp(xu), dp(xv), precond, C.byref(err), C.byref(nit)))
        return err.value, nit.value

    def pcg_solver_standard_3d(self, *a):
        """(active_vertex, Auu, Auv, Avu, Avv, bu, bv, xu, xv, precond) -> (err, niters)"""
        return self._pcg(lib().cism_pcg_solver_standard_3d, *a)

    def pcg_solver_chrongear_3d(self, *a):
        return self._pcg(lib().cism_pcg_solver_chrongear_3d, *a)

    # --- glissade_velo_higher --------------------------------------------------
    def assemble_stiffness_matrix_3d(self, sigma, active_cell, xVertex, yVertex, uvel, vvel,
                                     stagusrf, stagthck, flwafact, whichapprox, efvs, whichefvs,
                                     efvs_constant, Auu, Auv, Avu, Avv):
        _check(lib().cism_assemble_stiffness_matrix_3d(
            self.h, dp(sigma), ip(active_cell), dp(xVertex), dp(yVertex), dp(uvel), dp(vvel),
            dp(stagusrf), dp(stagthck), dp(flwafact), whichapprox, dp(efvs), whichefvs,
            efvs_constant, dp(Auu), dp(Auv), dp(Avu), dp(Avv)))

    def basal_sliding_bc_3d(self, dx, dy, active_vertex, beta, whichassemble_beta, Auu, Avv):
        _check(lib().cism_basal_sliding_bc_3d(self.h, dx, dy, ip(active_vertex), dp(beta),
                                              whichassemble_beta, dp(Auu), dp(Avv)))

    def basal_sliding_bc(self, dx, dy, active_cell, active_vertex, beta, xVertex, yVertex,
                         whichassemble_beta, Auu, Avv):
        """basal_sliding_bc with the reference's arguments (0 standard, 1 local assembly)"""
        _check(lib().cism_basal_sliding_bc(self.h, dx, dy, ip(active_cell), ip(active_vertex), dp(beta),
                                           dp(xVertex), dp(yVertex), whichassemble_beta, dp(Auu), dp(Avv)))

    def dirichlet_boundary_conditions_3d(self, active_vertex, umask, vmask, uvel, vvel,
                                         Auu, Auv, Avu, Avv, bu, bv):
        _check(lib().cism_dirichlet_boundary_conditions_3d(
            self.h, ip(active_vertex), ip(umask), ip(vmask), dp(uvel), dp(vvel), dp(Auu), dp(Auv),
            dp(Avu), dp(Avv), dp(bu), dp(bv)))

    def check_symmetry_assembled_matrix_3d(self, active_vertex, Auu, Auv, Avu, Avv):
        nw = C.c_int64()
        _check(lib().cism_check_symmetry_assembled_matrix_3d(
            self.h, ip(active_vertex), dp(Auu), dp(Auv), dp(Avu), dp(Avv), C.byref(nw)))
        return nw.value

    def count_nonzeros_3d(self, Auu, Auv, Avu, Avv, active_vertex):
        n = C.c_int64()
        _check(lib().cism_count_nonzeros_3d(self.h, dp(Auu), dp(Auv), dp(Avu), dp(Avv),
                                            ip(active_vertex), C.byref(n)))
        return n.value

    def compute_residual_vector_3d(self, active_vertex, Auu, Auv, Avu, Avv, bu, bv, uvel, vvel,
                                   resid_u, resid_v):
        L2, L2r = C.c_double(), C.c_double()
        _check(lib().cism_compute_residual_vector_3d(
            self.h, ip(active_vertex), dp(Auu), dp(Auv), dp(Avu), dp(Avv), dp(bu), dp(bv),
            dp(uvel), dp(vvel), dp(resid_u), dp(resid_v), C.byref(L2), C.byref(L2r)))
        return L2.value, L2r.value

    def compute_residual_velocity_3d(self, whichresid, uvel, vvel, usav, vsav):
        r = C.c_double()
        _check(lib().cism_compute_residual_velocity_3d(self.h, whichresid, dp(uvel), dp(vvel),
                                                       dp(usav), dp(vsav), C.byref(r)))
        return r.value

    def load_vector_gravity(self, sigma, stagwbndsigma, dx, dy, active_cell, active_vertex,
                            xVertex, yVertex, stagusrf, stagthck, dusrf_dx, dusrf_dy,
                            whichassemble_taud, loadu, loadv):
        _check(lib().cism_load_vector_gravity(
            self.h, dp(sigma), dp(stagwbndsigma), dx, dy, ip(active_cell), ip(active_vertex),
            dp(xVertex), dp(yVertex), dp(stagusrf), dp(stagthck), dp(dusrf_dx), dp(dusrf_dy),
            whichassemble_taud, dp(loadu), dp(loadv)))

    def set_reference_velocity(self, uref, vref):
        _check(lib().cism_set_reference_velocity(self.h, dp(uref), dp(vref)))

    # --- grounded fraction and basal traction ----------------------------------------
    def grounded_fraction(self, ice_mask, floating_mask, land_mask, whichground=0):
        fg = np.zeros((self.nx - 1, self.ny - 1), order="F")
        ms = [np.asfortranarray(m, dtype=np.int32) for m in (ice_mask, floating_mask, land_mask)]
        _check(lib().cism_glissade_grounded_fraction(self.h, *[ip(m) for m in ms], whichground, dp(fg)))
        return fg

    @staticmethod
    def _opt(a):
        return None if a is None else np.asfortranarray(a, dtype=np.float64)

    def calcbeta(self, bp, ubas, vbas, beta_external=None, effecpress_stag=None, f_ground=None):
        s2 = (self.nx - 1, self.ny - 1)
        fg = np.ones(s2, order="F") if f_ground is None else f_ground
        arrs = [self._opt(a) for a in (ubas, vbas, beta_external, effecpress_stag, fg)]
        beta = np.zeros(s2, order="F")
        _check(lib().cism_calcbeta(self.h, C.byref(bp), *[None if a is None else dp(a) for a in arrs],
                                   dp(beta)))
        return beta

    def set_basal_physics(self, bp, beta_external=None, effecpress_stag=None, f_ground=None):
        """bp None switches the in-loop calcbeta off"""
        if bp is None:
            _check(lib().cism_set_basal_physics(self.h, None, None, None, None))
            return
        arrs = [self._opt(a) for a in (beta_external, effecpress_stag, f_ground)]
        _check(lib().cism_set_basal_physics(self.h, C.byref(bp),
                                            *[None if a is None else dp(a) for a in arrs]))

    # --- the solve's inputs from the geometry (velo_higher:1449-1939) -------------
    def glissade_velo_inputs(self, dx, dy, thklim, eus, thck, usrf, topg, flwa, max_slope=1.0):
        """dict of masks, vertex geometry, staggered fields, surface gradient, flwafact"""
        nx, ny, nz = self.nx, self.ny, self.nz
        ci = lambda: np.zeros((nx, ny), dtype=np.int32, order="F")
        vi = lambda: np.zeros((nx - 1, ny - 1), dtype=np.int32, order="F")
        vd = lambda: np.zeros((nx - 1, ny - 1), order="F")
        o = dict(ice_mask=ci(), floating_mask=ci(), ocean_mask=ci(), land_mask=ci(),
                 active_cell=ci(), active_vertex=vi(), xVertex=vd(), yVertex=vd(), stagusrf=vd(),
                 stagthck=vd(), dusrf_dx=vd(), dusrf_dy=vd(),
                 flwafact=np.zeros((nz - 1, nx, ny), order="F"))
        f = lambda a: np.asfortranarray(a, dtype=np.float64)
        _check(lib().cism_glissade_velo_inputs(
            self.h, dx, dy, thklim, eus, max_slope, dp(f(thck)), dp(f(usrf)), dp(f(topg)), dp(f(flwa)),
            *[ip(o[k]) for k in ("ice_mask", "floating_mask", "ocean_mask", "land_mask",
                                 "active_cell", "active_vertex")],
            *[dp(o[k]) for k in ("xVertex", "yVertex", "stagusrf", "stagthck", "dusrf_dx",
                                 "dusrf_dy", "flwafact")]))
        return o

    # --- post-solve diagnostics (velo_higher:3842-3894, 6755-7195) ----------------
    def compute_internal_stress(self, sigma, active_cell, xVertex, yVertex, stagusrf, stagthck,
                                flwafact, efvs, whichefvs, efvs_constant, whichapprox, uvel, vvel):
        """(tau_xz, tau_yz, tau_xx, tau_yy, tau_xy, tau_eff), each (nz-1, nx, ny)"""
        outs = [np.zeros((self.nz - 1, self.nx, self.ny), order="F") for _ in range(6)]
        _check(lib().cism_compute_internal_stress(
            self.h, dp(sigma), ip(active_cell), dp(xVertex), dp(yVertex), dp(stagusrf),
            dp(stagthck), dp(flwafact), dp(efvs), whichefvs, efvs_constant, whichapprox, dp(uvel),
            dp(vvel), *[dp(o) for o in outs]))
        return tuple(outs)

    def compute_basal_friction_heatflx(self, active_cell, active_vertex, xVertex, yVertex, ubas,
                                       vbas, beta, whichassemble_bfric=1):
        out = np.zeros((self.nx, self.ny), order="F")
        _check(lib().cism_compute_basal_friction_heatflx(
            self.h, ip(active_cell), ip(active_vertex), dp(xVertex), dp(yVertex),
            dp(np.asfortranarray(ubas)), dp(np.asfortranarray(vbas)), dp(beta), whichassemble_bfric,
            dp(out)))
        return out

    def compute_basal_traction(self, beta, uvel, vvel):
        bx = np.zeros((self.nx - 1, self.ny - 1), order="F")
        by = np.zeros((self.nx - 1, self.ny - 1), order="F")
        _check(lib().cism_compute_basal_traction(self.h, dp(beta), dp(uvel), dp(vvel), dp(bx), dp(by)))
        return bx, by

    def velo_diagnostics(self, whichapprox=2, whichefvs=2, efvs_constant=2336041.42829,
                         whichassemble_bfric=1):
        """stress components, basal friction heat flux and basal traction of the last solve"""
        tau = [np.zeros((self.nz - 1, self.nx, self.ny), order="F") for _ in range(6)]
        bf = np.zeros((self.nx, self.ny), order="F")
        bx = np.zeros((self.nx - 1, self.ny - 1), order="F")
        by = np.zeros((self.nx - 1, self.ny - 1), order="F")
        _check(lib().cism_velo_diagnostics(self.h, whichapprox, whichefvs, efvs_constant,
                                           whichassemble_bfric, *[dp(t) for t in tau], dp(bf),
                                           dp(bx), dp(by)))
        return dict(tau_xz=tau[0], tau_yz=tau[1], tau_xx=tau[2], tau_yy=tau[3], tau_xy=tau[4],
                    tau_eff=tau[5], bfricflx=bf, btractx=bx, btracty=by)

    def set_fp64_symmetric(self, enable):
        """1 (default): the FP64 solver reads the exactly symmetric FP64 matrix from its
        symmetric copy (the same iterates); 0: always from the full storage."""
        _check(lib().cism_set_fp64_symmetric(self.h, int(enable)))

    def set_precond_scale(self, su, sv):
        """Per-column scale of the mixed-precision SIA preconditioner (vertex fields); None
        clears it."""
        if su is None or sv is None:
            _check(lib().cism_set_precond_scale(self.h, None, None))
        else:
            su = np.asfortranarray(su, dtype=np.float64)
            sv = np.asfortranarray(sv, dtype=np.float64)
            _check(lib().cism_set_precond_scale(self.h, dp(su), dp(sv)))

    def column_scale_target(self):
        """(su, sv) vertex fields from the last solve's matrix, rhs and velocity."""
        su = np.ones((self.nx - 1, self.ny - 1), order="F")
        sv = np.ones((self.nx - 1, self.ny - 1), order="F")
        _check(lib().cism_column_scale_target(self.h, dp(su), dp(sv)))
        return su, sv

    def time_kernel(self, which, reps=20):
        """Mean ms per launch (HIP events on the context stream): 0 FP64 matvec as the
        FP64 solver runs it, 1 FP32 matvec, 2 assembly, 3 1 GiB copy (counter
        calibration), 4 FP32 fused update, 5 FP64 matvec on the full storage, 6 FP64
        symmetric matvec in the other launch order."""
        ms = C.c_double()
        _check(lib().cism_time_kernel(self.h, which, reps, C.byref(ms)))
        return ms.value

    # --- coarse-correction diagnostics -----------------------------------------
    def coarse_build(self, active_vertex, umask, vmask, Auu, Auv, Avu, Avv):
        """symmetry pass (FP32 matrix) and coarse hierarchy from a matrix in the reference's
        layout; returns the level count"""
        n = C.c_int()
        _check(lib().cism_coarse_build(self.h, ip(active_vertex), ip(umask), ip(vmask), dp(Auu), dp(Auv),
                                       dp(Avu), dp(Avv), C.byref(n)))
        return n.value

    def coarse_level(self, level):
        """(ax, ay, blocks): the FP64 Galerkin stencil of a level, blocks[aj, ai, d, b] with
        d = (di+1) + 3 (dj+1) and b = uu, uv, vu, vv"""
        ax, ay = C.c_int(), C.c_int()
        _check(lib().cism_coarse_level(self.h, level, C.byref(ax), C.byref(ay), None))
        blk = np.zeros(36 * ax.value * ay.value)
        _check(lib().cism_coarse_level(self.h, level, C.byref(ax), C.byref(ay), dp(blk)))
        return ax.value, ay.value, blk.reshape(ay.value, ax.value, 9, 4)

    def coarse_correction(self, ru, rv):
        """z = P V(P^T r) of the current hierarchy for r in the reference's layout"""
        ru, rv = np.asfortranarray(ru, dtype=np.float64), np.asfortranarray(rv, dtype=np.float64)
        zu, zv = np.zeros_like(ru, order="F"), np.zeros_like(rv, order="F")
        _check(lib().cism_coarse_correction(self.h, dp(ru), dp(rv), dp(zu), dp(zv)))
        return zu, zv

    def load_vector_lateral_bc(self, sigma, land_mask, ocean_mask, calving_front_mask, active_cell,
                               xVertex, yVertex, stagusrf, stagthck, loadu, loadv):
        _check(lib().cism_load_vector_lateral_bc(
            self.h, dp(sigma), ip(land_mask), ip(ocean_mask), ip(calving_front_mask),
            ip(active_cell), dp(xVertex), dp(yVertex), dp(stagusrf), dp(stagthck), dp(loadu),
            dp(loadv)))

    def loads(self, p):
        """Gravity plus lateral ice-ocean load, computed on the device (velo_higher:2033-2111)."""
        shape = (p.nz, p.nx - 1, p.ny - 1)
        lu, lv = np.zeros(shape, order="F"), np.zeros(shape, order="F")
        self.load_vector_gravity(p.sigma, p.stagwbndsigma, p.dx, p.dy, p.active_cell,
                                 p.active_vertex, p.xVertex, p.yVertex, p.stagusrf, p.stagthck,
                                 p.dusrf_dx, p.dusrf_dy, p.whichassemble_taud, lu, lv)
        if p.ocean_mask.any():
            cf = np.zeros_like(p.ocean_mask)
            self.load_vector_lateral_bc(p.sigma, p.land_mask, p.ocean_mask, cf, p.active_cell,
                                        p.xVertex, p.yVertex, p.stagusrf, p.stagthck, lu, lv)
        return lu, lv

    def staggered_parallel_halo(self, a):
        lead = int(np.prod(a.shape[:-2]))
        _check(lib().cism_staggered_parallel_halo(self.h, lead, dp(a)))

    # --- device-resident Picard solve ------------------------------------------
    def _opts(self, p, precision=0, maxiter_nonlinear=None, whichsparse=None, whichprecond=None,
              whichresid=None, target_override=0.0, inner_rtol=1e-3, adaptive_linear=0,
              adaptive_gamma=0.05, stop_rel_err=0.0, anderson_m=0, initial_guess=0, coarse=0,
              skip_final_residual=0):
        """the Picard loop's options (cism_solve_opts) from problem p; coarse=1 adds the
        aggregation coarse correction to the SIA preconditioner of the FP32 inner solves;
        skip_final_residual=1 lets a linear solve end on its inner solve's estimate"""
        o = SolveOpts()
        lib().cism_solve_opts_default(C.byref(o))
        o.whichbabc = p.whichbabc
        o.whichefvs = p.whichefvs
        o.whichresid = p.whichresid if whichresid is None else whichresid
        o.whichsparse = p.whichsparse if whichsparse is None else whichsparse
        o.whichprecond = p.whichprecond if whichprecond is None else whichprecond
        o.whichapprox = p.whichapprox
        o.maxiter_nonlinear = p.maxiter_nonlinear if maxiter_nonlinear is None else maxiter_nonlinear
        o.precision = precision
        o.efvs_constant = p.efvs_constant
        o.dx, o.dy = p.dx, p.dy
        o.target_override = target_override
        o.inner_rtol = inner_rtol
        o.adaptive_linear = adaptive_linear
        o.adaptive_gamma = adaptive_gamma
        o.stop_rel_err = stop_rel_err
        o.anderson_m = anderson_m
        o.whichassemble_beta = p.whichassemble_beta
        o.initial_guess = initial_guess
        o.coarse = coarse
        o.skip_final_residual = skip_final_residual
        return o

    def glissade_velo_higher_solve(self, p, bp=None, beta_external=None, effecpress_stag=None,
                                   initial_guess=0, uvel=None, vvel=None, lateral_bc=None,
                                   whichground=0, max_slope=1.0, eus=0.0, efvs=True, out=None, **kw):
        """glissade_velo_higher_solve from the geometry of p (thck, usrf, topg, flwa) with every
        input computed on the device; bp (BasalPhysics) is required unless p is no-slip.
        initial_guess: 0 uvel/vvel (zero by default), 1 the learned guess, 2 the context's
        last velocity.  Options of the Picard loop as in velo_higher_solve_3d.
        out: (uvel, vvel), F-ordered float64 arrays of shape (nz, nx-1, ny-1) that the solve
        uses in place as its in/out velocity (read only with initial_guess 0), as a Fortran
        caller passes its model arrays; without it uvel/vvel (default p.uvel/p.vvel) are
        copied for initial_guess 0, and new arrays receive the solution otherwise."""
        o = self._opts(p, **kw)
        su = VeloSetup(dx=p.dx, dy=p.dy, thklim=p.thklim, eus=eus, max_slope=max_slope,
                       whichground=whichground, whichassemble_taud=p.whichassemble_taud,
                       lateral_bc=int(p.ocean_mask.any()) if lateral_bc is None else int(lateral_bc),
                       initial_guess=initial_guess)
        shape = (p.nz, p.nx - 1, p.ny - 1)
        if out is not None:
            u, v = out
            for a in (u, v):
                if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.shape == shape
                        and a.flags.f_contiguous and a.flags.writeable):
                    raise ValueError(f"out: writeable F-ordered float64 arrays of shape {shape}")
        elif initial_guess == 0:
            u = (p.uvel if uvel is None else uvel).copy(order="F")
            v = (p.vvel if vvel is None else vvel).copy(order="F")
        else:   # the input velocity is not read
            u, v = np.empty(shape, order="F"), np.empty(shape, order="F")
        ef = np.zeros((p.nz - 1, p.nx, p.ny), order="F") if efvs else None
        f = lambda a: None if a is None else np.asfortranarray(a, dtype=np.float64)   # noqa: E731
        bext, N = f(beta_external), f(effecpress_stag)
        st = SolveStats()
        _check(lib().cism_glissade_velo_higher_solve(
            self.h, C.byref(o), C.byref(su), dp(p.sigma), dp(f(p.thck)), dp(f(p.usrf)), dp(f(p.topg)),
            dp(f(p.flwa)), C.byref(bp) if bp is not None else None, dp(bext), dp(N), dp(u), dp(v),
            dp(ef), C.byref(st)))
        return dict(uvel=u, vvel=v, efvs=ef, **st.as_dict())

    def learned_guess_load(self, params, nz, encoding, uscale):
        w = np.ascontiguousarray(params, dtype=np.float32)
        _check(lib().cism_learned_guess_load(self.h, w.ctypes.data_as(C.POINTER(C.c_float)), w.size, nz,
                                             encoding, uscale))

    def learned_guess_load_file(self, path):
        _check(lib().cism_learned_guess_load_file(self.h, os.fsencode(path)))

    def learned_guess(self, p=None, download=True):
        """the learned guess into the context's velocity from p's fields (None: the
        context's resident fields); returns (uvel, vvel) when download"""
        shape = (self.nz, self.nx - 1, self.ny - 1)
        u = np.zeros(shape, order="F") if download else None
        v = np.zeros(shape, order="F") if download else None
        if p is None:
            args = [None] * 5 + [None]
        else:
            args = [dp(p.stagthck), dp(p.dusrf_dx), dp(p.dusrf_dy), dp(p.beta_internal), dp(p.flwafact),
                    ip(p.active_vertex)]
        _check(lib().cism_learned_guess(self.h, *args, dp(u), dp(v)))
        return u, v

    def learned_guess_last_ms(self):
        return float(lib().cism_learned_guess_last_ms(self.h))

    def velo_higher_solve_3d(self, p, loadu, loadv, uvel=None, vvel=None, **kw):
        """the device Picard loop from prepared inputs; options as in _opts.  initial_guess
        2: start from the context's velocity (the last solve's or the learned guess),
        uvel/vvel then only receive the solution"""
        o = self._opts(p, **kw)
        u = (p.uvel if uvel is None else uvel).copy(order="F")
        v = (p.vvel if vvel is None else vvel).copy(order="F")
        shape = (p.nz, p.nx - 1, p.ny - 1)
        efvs = np.zeros((p.nz - 1, p.nx, p.ny), order="F")
        bu, bv, ru, rv = (np.zeros(shape, order="F") for _ in range(4))
        st = SolveStats()
        _check(lib().cism_velo_higher_solve_3d(
            self.h, C.byref(o), dp(p.sigma), ip(p.active_cell), ip(p.active_vertex),
            dp(p.xVertex), dp(p.yVertex), dp(p.stagusrf), dp(p.stagthck), dp(p.flwafact),
            dp(p.beta_internal), ip(p.umask_dirichlet), ip(p.vmask_dirichlet), dp(loadu),
            dp(loadv), dp(u), dp(v), dp(efvs), dp(bu), dp(bv), dp(ru), dp(rv), C.byref(st)))
        return dict(uvel=u, vvel=v, efvs=efvs, bu=bu, bv=bv, resid_u=ru, resid_v=rv,
                    **st.as_dict())

    def velo_higher_solve_depth(self, p, loadu, loadv, approx, uvel=None, vvel=None, uvel_2d=None,
                                vvel_2d=None, btractx=None, btracty=None, efvs=None, **kw):
        """The L1L2 (approx 3) or DIVA (approx 4) branch on a one-level context: 3D inputs of
        p in; the 3D velocity, the 2D solution (uvel_2d/vvel_2d), DIVA's basal traction and
        the layer viscosities out.  DIVA starts from uvel_2d/vvel_2d, btractx/btracty and
        efvs (zero by default).  Options of the Picard loop as in velo_higher_solve_3d."""
        if self.nz != 1:
            raise ValueError("velo_higher_solve_depth needs a one-level context (Context.for_problem_2d)")
        o = self._opts(p, **kw)
        o.whichapprox = approx
        f = lambda a: np.asfortranarray(a, dtype=np.float64)   # noqa: E731
        s2 = (p.nx - 1, p.ny - 1)
        z2 = lambda a: np.zeros(s2, order="F") if a is None else f(a).copy(order="F")   # noqa: E731
        u = (p.uvel if uvel is None else uvel).copy(order="F")
        v = (p.vvel if vvel is None else vvel).copy(order="F")
        u2, v2, bx, by = z2(uvel_2d), z2(vvel_2d), z2(btractx), z2(btracty)
        ef = np.zeros((p.nz - 1, p.nx, p.ny), order="F") if efvs is None else f(efvs).copy(order="F")
        ipl = np.asfortranarray((p.ice_mask != 0) | (p.land_mask != 0), dtype=np.int32)
        bu, bv, ru, rv = (np.zeros(s2, order="F") for _ in range(4))
        st = SolveStats()
        _check(lib().cism_velo_higher_solve_depth(
            self.h, C.byref(o), p.nz, dp(p.sigma), ip(p.ice_mask), ip(ipl), ip(p.active_cell),
            ip(p.active_vertex), dp(p.xVertex), dp(p.yVertex), dp(f(p.thck)), dp(p.stagthck), dp(p.dusrf_dx),
            dp(p.dusrf_dy), dp(f(p.flwa)), dp(p.flwafact), dp(p.beta_internal), ip(p.umask_dirichlet),
            ip(p.vmask_dirichlet), dp(loadu), dp(loadv), dp(u), dp(v), dp(u2), dp(v2), dp(bx), dp(by), dp(ef),
            dp(bu), dp(bv), dp(ru), dp(rv), C.byref(st)))
        return dict(uvel=u, vvel=v, uvel_2d=u2, vvel_2d=v2, btractx=bx, btracty=by, efvs=ef, bu=bu, bv=bv,
                    resid_u=ru, resid_v=rv, **st.as_dict())

    def velo_higher_solve_2d(self, p, loadu, loadv, uvel=None, vvel=None, **kw):
        """The SSA branch (whichapprox 1) on a one-level context: 3D inputs of p in, the
        2D solution broadcast to every level of uvel/vvel out; bu/bv/resid are 2D.
        stop_rel_err > 0 stops at that relative error against the reference velocity
        (set_reference_velocity with (1, nx-1, ny-1) arrays)."""
        if self.nz != 1:
            raise ValueError("velo_higher_solve_2d needs a one-level context (Context.for_problem_2d)")
        o = self._opts(p, **kw)
        o.whichapprox = 1
        u = (p.uvel if uvel is None else uvel).copy(order="F")
        v = (p.vvel if vvel is None else vvel).copy(order="F")
        s2 = (p.nx - 1, p.ny - 1)
        efvs = np.zeros((p.nz - 1, p.nx, p.ny), order="F")
        bu, bv, ru, rv = (np.zeros(s2, order="F") for _ in range(4))
        st = SolveStats()
        _check(lib().cism_velo_higher_solve_2d(
            self.h, C.byref(o), p.nz, dp(p.sigma), ip(p.active_cell), ip(p.active_vertex),
            dp(p.xVertex), dp(p.yVertex), dp(p.stagusrf), dp(p.stagthck), dp(p.flwafact),
            dp(p.beta_internal), ip(p.umask_dirichlet), ip(p.vmask_dirichlet), dp(loadu),
            dp(loadv), dp(u), dp(v), dp(efvs), dp(bu), dp(bv), dp(ru), dp(rv), C.byref(st)))
        return dict(uvel=u, vvel=v, efvs=efvs, bu=bu, bv=bv, resid_u=ru, resid_v=rv,
                    **st.as_dict())
