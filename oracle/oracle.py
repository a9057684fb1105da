"""ctypes binding of the CPU oracle (liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  Arrays are numpy
arrays in the reference's Fortran layouts (see cism_oracle.h).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liborc.so")
# the same restatement with its hot loops on all host cores (OpenMP), for the all-core
# CPU baseline only: it sums in another order and is never the checker
LIB_OMP = os.path.join(HERE, "liborc_omp.so")

_libs = {}
_all_cores = False


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def set_all_cores(on: bool):
    """True: the functions below run the OpenMP build (timing only); False: the serial
    oracle."""
    global _all_cores
    _all_cores = bool(on)


def lib():
    path = LIB_OMP if _all_cores else LIB
    if path not in _libs:
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        _declare(L)
        _libs[path] = L
    return _libs[path]


class Grid(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("nx", "ny", "nz", "nhalo", "ilo", "ihi", "jlo", "jhi",
                                       "periodic")]


class BasalPhysics(C.Structure):
    """the basal_physics fields calcbeta reads (glide_types.F90:1308-1370), reference defaults"""
    _fields_ = [("whichbabc", C.c_int), ("reserved", C.c_int)] + \
               [(n, C.c_double) for n in ("ho_beta_const", "ho_beta_large", "powerlaw_c",
                                          "powerlaw_m", "coulomb_c", "beta_grounded_min",
                                          "beta_powerlaw_umax")]

    def __init__(self, whichbabc=5, **kw):
        d = dict(ho_beta_const=1000.0, ho_beta_large=1.0e10, powerlaw_c=1.0e4, powerlaw_m=3.0,
                 coulomb_c=0.42, beta_grounded_min=1.0, beta_powerlaw_umax=0.0)
        d.update(kw)
        super().__init__(whichbabc=whichbabc, **d)


class PicardOpts(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("whichbabc", "whichefvs", "whichresid", "whichsparse",
                                       "whichprecond", "whichapprox", "maxiter_nonlinear",
                                       "linear_maxiters")] + \
               [(n, C.c_double) for n in ("efvs_constant", "linear_tolerance", "dx", "dy")] + \
               [(n, C.c_int) for n in ("picard_iters", "total_linear_iters", "status")] + \
               [(n, C.c_double) for n in ("final_L2_norm", "final_resid_velo")] + \
               [("bp", C.POINTER(BasalPhysics))] + \
               [(n, C.POINTER(C.c_double)) for n in ("beta_external", "effecpress_stag", "f_ground")] + \
               [("target_override", C.c_double), ("L2_history", C.POINTER(C.c_double)),
                ("L2_history_cap", C.c_int), ("converged", C.c_int),
                ("assemble_beta_standard", C.c_int)]


class DepthFields(C.Structure):
    _fields_ = [("thck", C.POINTER(C.c_double)), ("dusrf_dx", C.POINTER(C.c_double)),
                ("dusrf_dy", C.POINTER(C.c_double)), ("flwa", C.POINTER(C.c_double)),
                ("ice_mask", C.POINTER(C.c_int32)), ("ice_plus_land_mask", C.POINTER(C.c_int32)),
                ("uvel_2d", C.POINTER(C.c_double)), ("vvel_2d", C.POINTER(C.c_double)),
                ("btractx", C.POINTER(C.c_double)), ("btracty", C.POINTER(C.c_double))]


def _set_basal(o, basal, keep):
    """basal = (BasalPhysics, beta_external, effecpress_stag, f_ground) or None"""
    if basal is None:
        return
    bp, bext, N, fg = basal
    arrs = [np.asfortranarray(a, dtype=np.float64) for a in (bext, N, fg)]
    keep.extend([bp] + arrs)
    o.bp = C.pointer(bp)
    o.beta_external, o.effecpress_stag, o.f_ground = (dp(a) for a in arrs)


D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int32)
G = C.POINTER(Grid)


def _declare(L):
    L.orc_grid_init.argtypes = [G, C.c_int, C.c_int, C.c_int, C.c_int]
    L.orc_tables.argtypes = [D, D, D, D, C.POINTER(C.c_int), C.POINTER(C.c_int),
                             C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.orc_basis_derivatives.argtypes = [D, D, D, C.c_int, D, D, D, D]
    L.orc_efvs_qp.restype = C.c_double
    L.orc_efvs_qp.argtypes = [C.c_int, C.c_int, C.c_double, D, D, D, D, D, C.c_double]
    L.orc_element_matrix.argtypes = [G, C.c_int, C.c_int, C.c_int, D, D, D, D, D, D, D, D,
                                     C.c_int, C.c_int, C.c_double, D, D, D, D, D]
    L.orc_assemble_3d.argtypes = [G, D, I, D, D, D, D, D, D, D, C.c_int, D, C.c_int,
                                  C.c_double, D, D, D, D]
    L.orc_basal_sliding_local.argtypes = [G, C.c_double, C.c_double, I, D, D, D]
    L.orc_basal_sliding_standard.argtypes = [G, I, D, D, D, D, D]
    L.orc_dirichlet_3d.argtypes = [G, I, I, I, D, D, D, D, D, D, D, D]
    L.orc_halo.argtypes = [G, C.c_int, D]
    L.orc_check_symmetry_3d.argtypes = [G, I, D, D, D, D, D]
    L.orc_count_nonzeros_3d.restype = C.c_long
    L.orc_count_nonzeros_3d.argtypes = [G, I, D, D, D, D]
    L.orc_residual_vector_3d.argtypes = [G, I, D, D, D, D, D, D, D, D, D, D, D, D]
    L.orc_residual_velocity_3d.restype = C.c_double
    L.orc_residual_velocity_3d.argtypes = [G, C.c_int, D, D, D, D]
    L.orc_load_gravity_local.argtypes = [G, D, C.c_double, C.c_double, I, D, D, D, D, D]
    L.orc_load_gravity_standard.argtypes = [G, D, I, D, D, D, D, D, D, D, D]
    L.orc_matvec_3d.argtypes = [G, I, D, D, D, D, D, D, D, D]
    L.orc_load_lateral_bc.argtypes = [G, D, I, I, I, D, D, D, D, D, D]
    L.orc_tridiag.argtypes = [C.c_int, D, D, D, D, D]
    L.orc_cell_halo.argtypes = [G, D]
    L.orc_vertical_average.argtypes = [C.c_int, C.c_int, C.c_int, D, D, D]
    L.orc_assemble_2d_ssa.argtypes = [G, C.c_int, D, I, D, D, D, D, D, D, C.c_int, C.c_double, D,
                                      D, D, D, D]
    L.orc_picard_solve_2d.argtypes = [G, C.POINTER(PicardOpts), D, I, I, D, D, D, D, D, D, I, I,
                                      D, D, D, D, D, D, D, D, D]
    L.orc_internal_stress.argtypes = [G, D, I, D, D, D, D, D, D, C.c_int, C.c_double, C.c_int, D, D,
                                      D, D, D, D, D, D]
    L.orc_basal_friction_heatflx.argtypes = [G, I, I, D, D, D, C.c_int, D]
    L.orc_basal_traction.argtypes = [G, D, D, D, D, D]
    L.orc_grounded_fraction.argtypes = [C.c_int, C.c_int, I, I, I, C.c_int, D]
    L.orc_calcbeta.argtypes = [G, C.POINTER(BasalPhysics), D, D, D, D, D, D]
    pcg = [G, I, D, D, D, D, D, D, D, D, C.c_int, C.c_int, C.c_double, D, C.POINTER(C.c_int)]
    L.orc_pcg_standard_3d.argtypes = pcg
    L.orc_pcg_chrongear_3d.argtypes = pcg
    L.orc_picard_solve.argtypes = [G, C.POINTER(PicardOpts), D, I, I, D, D, D, D, D, D, I, I,
                                   D, D, D, D, D, D, D, D, D]
    L.orc_precond_apply.argtypes = [G, I, C.c_int, D, D, D, D, D, D]
    L.orc_stagger.argtypes = [C.c_int, C.c_int, D, D, I, C.c_int]
    L.orc_gradient.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, D, D, D, I, C.c_int]
    L.orc_efvs_l1l2.argtypes = [C.c_int, C.c_double, C.c_int, D, D, D, D, D, D, D, D, D, D, D, D]
    L.orc_efvs_diva.argtypes = [C.c_int, C.c_double, C.c_int, D, D, D, D, D, D, D, D, D, D]
    L.orc_integrals_diva.argtypes = [C.c_int, D, C.c_double, D, D, D]
    L.orc_velocity_3d_l1l2.argtypes = [G, D, C.c_double, C.c_double, I, I, I, I, I, D, D, D, D, D, D,
                                       D, D, D, D]
    L.orc_velocity_3d_diva.argtypes = [G, I, D, D, D, D, D, D, D, D]
    L.orc_picard_solve_depth.argtypes = [G, C.POINTER(PicardOpts), C.POINTER(DepthFields), D, I, I, D,
                                         D, D, D, D, I, I, D, D, D, D, D, D, D, D, D]


def dp(a):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim <= 1), a.flags
    return a.ctypes.data_as(D)


def ip(a):
    assert a.dtype == np.int32 and (a.flags.f_contiguous or a.ndim <= 1)
    return a.ctypes.data_as(I)


def grid(nx, ny, nz, periodic=True):
    g = Grid()
    lib().orc_grid_init(C.byref(g), nx, ny, nz, 1 if periodic else 0)
    return g


def grid_of(p):
    return grid(p.nx, p.ny, p.nz, p.periodic)


def tables():
    phi, dxr, dyr, dzr = (np.zeros((8, 8), order="F") for _ in range(4))
    ish, jsh, ksh = (np.zeros((8, 8), dtype=np.int32, order="F") for _ in range(3))
    ind = np.zeros((3, 3, 3), dtype=np.int32, order="F")
    P = C.POINTER(C.c_int)
    lib().orc_tables(dp(phi), dp(dxr), dp(dyr), dp(dzr), ish.ctypes.data_as(P),
                     jsh.ctypes.data_as(P), ksh.ctypes.data_as(P), ind.ctypes.data_as(P))
    return dict(phi=phi, dxr=dxr, dyr=dyr, dzr=dzr, ishift=ish, jshift=jsh, kshift=ksh, indxA=ind)


def basis_derivatives(x, y, z, p):
    dx, dy, dz = np.zeros(8), np.zeros(8), np.zeros(8)
    det = C.c_double()
    st = lib().orc_basis_derivatives(dp(np.ascontiguousarray(x, float)), dp(np.ascontiguousarray(y, float)),
                                     dp(np.ascontiguousarray(z, float)), p, dp(dx), dp(dy), dp(dz),
                                     C.byref(det))
    return st, dx, dy, dz, det.value


def efvs_qp(whichefvs, whichapprox, efvs_constant, dphx, dphy, dphz, u, v, flwafact):
    a = [np.ascontiguousarray(t, float) for t in (dphx, dphy, dphz, u, v)]
    return lib().orc_efvs_qp(whichefvs, whichapprox, efvs_constant, *[dp(t) for t in a], flwafact)


def mat_zeros(p):
    return [np.zeros((27, p.nz, p.nx - 1, p.ny - 1), order="F") for _ in range(4)]


def vec_zeros(p):
    return np.zeros((p.nz, p.nx - 1, p.ny - 1), order="F")


def element_matrix(p, i, j, k, uvel=None, vvel=None):
    uvel = p.uvel if uvel is None else uvel
    vvel = p.vvel if vvel is None else vvel
    K = [np.zeros((8, 8), order="F") for _ in range(4)]
    ev = C.c_double()
    g = grid_of(p)
    st = lib().orc_element_matrix(C.byref(g), i, j, k, dp(p.sigma), dp(p.xVertex), dp(p.yVertex),
                                  dp(uvel), dp(vvel), dp(p.stagusrf), dp(p.stagthck),
                                  dp(p.flwafact), p.whichapprox, p.whichefvs, p.efvs_constant,
                                  *[dp(k_) for k_ in K], C.byref(ev))
    return st, K, ev.value


def assemble(p, uvel=None, vvel=None):
    uvel = p.uvel if uvel is None else uvel
    vvel = p.vvel if vvel is None else vvel
    A = mat_zeros(p)
    efvs = np.zeros((p.nz - 1, p.nx, p.ny), order="F")
    g = grid_of(p)
    st = lib().orc_assemble_3d(C.byref(g), dp(p.sigma), ip(p.active_cell), dp(p.xVertex),
                               dp(p.yVertex), dp(uvel), dp(vvel), dp(p.stagusrf), dp(p.stagthck),
                               dp(p.flwafact), p.whichapprox, dp(efvs), p.whichefvs,
                               p.efvs_constant, *[dp(a) for a in A])
    return st, A, efvs


def basal_local(p, A):
    g = grid_of(p)
    lib().orc_basal_sliding_local(C.byref(g), p.dx, p.dy, ip(p.active_vertex), dp(p.beta_internal),
                                  dp(A[0]), dp(A[3]))


def basal_standard(p, A, beta=None):
    """basal friction by the standard finite-element assembly, added into A[0], A[3]"""
    g = grid_of(p)
    b = p.beta_internal if beta is None else beta
    lib().orc_basal_sliding_standard(C.byref(g), ip(p.active_cell), dp(p.xVertex), dp(p.yVertex), dp(b),
                                     dp(A[0]), dp(A[3]))


def dirichlet(p, A, bu, bv, uvel=None, vvel=None):
    uvel = p.uvel if uvel is None else uvel
    vvel = p.vvel if vvel is None else vvel
    g = grid_of(p)
    lib().orc_dirichlet_3d(C.byref(g), ip(p.active_vertex), ip(p.umask_dirichlet),
                           ip(p.vmask_dirichlet), dp(uvel), dp(vvel), *[dp(a) for a in A],
                           dp(bu), dp(bv))


def halo(p, a):
    g = grid_of(p)
    lead = int(np.prod(a.shape[:-2])) if a.ndim > 2 else 1
    lib().orc_halo(C.byref(g), lead, dp(a))


def check_symmetry(p, A):
    g = grid_of(p)
    md = C.c_double()
    nw = lib().orc_check_symmetry_3d(C.byref(g), ip(p.active_vertex), *[dp(a) for a in A],
                                     C.byref(md))
    return nw, md.value


def count_nonzeros(p, A):
    g = grid_of(p)
    return lib().orc_count_nonzeros_3d(C.byref(g), ip(p.active_vertex), *[dp(a) for a in A])


def residual_vector(p, A, bu, bv, uvel, vvel):
    g = grid_of(p)
    ru, rv = vec_zeros(p), vec_zeros(p)
    L2, L2r = C.c_double(), C.c_double()
    lib().orc_residual_vector_3d(C.byref(g), ip(p.active_vertex), *[dp(a) for a in A], dp(bu),
                                 dp(bv), dp(uvel), dp(vvel), dp(ru), dp(rv), C.byref(L2),
                                 C.byref(L2r))
    return ru, rv, L2.value, L2r.value


def residual_velocity(p, whichresid, u, v, us, vs):
    g = grid_of(p)
    return lib().orc_residual_velocity_3d(C.byref(g), whichresid, dp(u), dp(v), dp(us), dp(vs))


def load_gravity(p, standard=False):
    g = grid_of(p)
    lu, lv = vec_zeros(p), vec_zeros(p)
    if standard:
        lib().orc_load_gravity_standard(C.byref(g), dp(p.sigma), ip(p.active_cell), dp(p.xVertex),
                                        dp(p.yVertex), dp(p.stagusrf), dp(p.stagthck),
                                        dp(p.dusrf_dx), dp(p.dusrf_dy), dp(lu), dp(lv))
    else:
        lib().orc_load_gravity_local(C.byref(g), dp(p.stagwbndsigma), p.dx, p.dy,
                                     ip(p.active_vertex), dp(p.stagthck), dp(p.dusrf_dx),
                                     dp(p.dusrf_dy), dp(lu), dp(lv))
    return lu, lv


def load_lateral(p, loadu, loadv):
    """Adds the ice-ocean face pressure to loadu, loadv in place."""
    g = grid_of(p)
    cf = np.zeros((p.nx, p.ny), dtype=np.int32, order="F")
    lib().orc_load_lateral_bc(C.byref(g), dp(p.sigma), ip(p.ocean_mask), ip(cf), ip(p.active_cell),
                              dp(p.xVertex), dp(p.yVertex), dp(p.stagusrf), dp(p.stagthck),
                              dp(loadu), dp(loadv))


def loads(p):
    """Gravity (per p.whichassemble_taud) plus the lateral ice-ocean load, as the reference
    assembles them before the Picard loop (velo_higher:2033-2111)."""
    lu, lv = load_gravity(p, standard=(p.whichassemble_taud == 0))
    load_lateral(p, lu, lv)
    return lu, lv


def matvec(p, A, xu, xv, active_vertex=None):
    g = grid_of(p)
    av = p.active_vertex if active_vertex is None else active_vertex
    yu, yv = vec_zeros(p), vec_zeros(p)
    lib().orc_matvec_3d(C.byref(g), ip(av), *[dp(a) for a in A], dp(xu), dp(xv), dp(yu), dp(yv))
    return yu, yv


def tridiag(sb, d, sp, rhs):
    n = len(d)
    out = np.zeros(n)
    lib().orc_tridiag(n, *[dp(np.ascontiguousarray(t, float)) for t in (sb, d, sp, rhs)], dp(out))
    return out


def pcg(p, A, bu, bv, xu, xv, precond, chrongear=True, maxiters=200, tolerance=1e-8):
    """Runs the oracle PCG in place on xu, xv; returns (status, err, niters)."""
    g = grid_of(p)
    err, nit = C.c_double(), C.c_int()
    f = lib().orc_pcg_chrongear_3d if chrongear else lib().orc_pcg_standard_3d
    st = f(C.byref(g), ip(p.active_vertex), *[dp(a) for a in A], dp(bu), dp(bv), dp(xu), dp(xv),
           precond, maxiters, tolerance, C.byref(err), C.byref(nit))
    return st, err.value, nit.value


def precond_apply(p, A, precond, ru, rv, zu=None, zv=None, active_vertex=None):
    """z = M^-1 r with the PCG preconditioner built from A (0 none, 1 diagonal, 2 column
    tridiagonal); zu, zv start as given (zeros by default)."""
    g = grid_of(p)
    av = p.active_vertex if active_vertex is None else active_vertex
    zu = vec_zeros(p) if zu is None else zu
    zv = vec_zeros(p) if zv is None else zv
    lib().orc_precond_apply(C.byref(g), ip(av), precond, dp(A[0]), dp(A[3]), dp(ru), dp(rv),
                            dp(zu), dp(zv))
    return zu, zv


def picard(p, loadu, loadv, maxiter=None, linear_maxiters=200, linear_tolerance=1e-8,
           whichsparse=None, whichprecond=None, whichresid=None, basal=None, target=None,
           uvel=None, vvel=None, assemble_beta=None):
    """Runs the oracle Picard loop from p.uvel/p.vvel (copied; or from uvel/vvel). Returns a
    result dict with the pre-solve residual norm of every Picard iteration (L2_history).
    basal: (BasalPhysics, beta_external, effecpress_stag, f_ground) to recompute beta with
    calcbeta every Picard iteration, or None for the fixed p.beta_internal.  target > 0
    replaces the Picard target (1e-4 for the L2 norm)."""
    g = grid_of(p)
    o = PicardOpts()
    keep = []
    _set_basal(o, basal, keep)
    cap = p.maxiter_nonlinear if maxiter is None else maxiter
    hist = np.zeros(max(cap, 1))
    o.L2_history = dp(hist)
    o.L2_history_cap = cap
    o.target_override = 0.0 if target is None else float(target)
    wab = p.whichassemble_beta if assemble_beta is None else assemble_beta
    o.assemble_beta_standard = 1 if wab == 0 else 0
    o.whichbabc = p.whichbabc
    o.whichefvs = p.whichefvs
    o.whichresid = p.whichresid if whichresid is None else whichresid
    o.whichsparse = p.whichsparse if whichsparse is None else whichsparse
    o.whichprecond = p.whichprecond if whichprecond is None else whichprecond
    o.whichapprox = p.whichapprox
    o.maxiter_nonlinear = p.maxiter_nonlinear if maxiter is None else maxiter
    o.linear_maxiters = linear_maxiters
    o.efvs_constant = p.efvs_constant
    o.linear_tolerance = linear_tolerance
    o.dx, o.dy = p.dx, p.dy
    u = (p.uvel if uvel is None else uvel).copy(order="F")
    v = (p.vvel if vvel is None else vvel).copy(order="F")
    efvs = np.zeros((p.nz - 1, p.nx, p.ny), order="F")
    bu, bv, ru, rv = vec_zeros(p), vec_zeros(p), vec_zeros(p), vec_zeros(p)
    st = lib().orc_picard_solve(C.byref(g), C.byref(o), dp(p.sigma), ip(p.active_cell),
                                ip(p.active_vertex), dp(p.xVertex), dp(p.yVertex),
                                dp(p.stagusrf), dp(p.stagthck), dp(p.flwafact),
                                dp(p.beta_internal), ip(p.umask_dirichlet),
                                ip(p.vmask_dirichlet), dp(loadu), dp(loadv), dp(u), dp(v),
                                dp(efvs), dp(bu), dp(bv), dp(ru), dp(rv))
    return dict(status=st, uvel=u, vvel=v, efvs=efvs, bu=bu, bv=bv, resid_u=ru, resid_v=rv,
                picard_iters=o.picard_iters, linear_iters=o.total_linear_iters,
                L2_norm=o.final_L2_norm, resid_velo=o.final_resid_velo,
                L2_history=hist[:min(o.picard_iters, cap)].copy(), converged=bool(o.converged))


def grounded_fraction(ice, floating, land, whichground=0):
    """glissade_grounded_fraction: f_ground (nx-1, ny-1) from cell masks (nx, ny)"""
    nx, ny = ice.shape
    fg = np.zeros((nx - 1, ny - 1), order="F")
    ms = [np.asfortranarray(m, dtype=np.int32) for m in (ice, floating, land)]
    st = lib().orc_grounded_fraction(nx, ny, *[ip(m) for m in ms], whichground, dp(fg))
    return st, fg


def calcbeta(p, bp, ubas, vbas, beta_external=None, effecpress_stag=None, f_ground=None):
    """calcbeta on p's grid: beta (nx-1, ny-1) with its staggered halo"""
    s2 = (p.nx - 1, p.ny - 1)
    z = np.zeros(s2, order="F")
    arrs = [np.asfortranarray(a if a is not None else z, dtype=np.float64)
            for a in (ubas, vbas, beta_external, effecpress_stag,
                      f_ground if f_ground is not None else np.ones(s2))]
    beta = np.zeros(s2, order="F")
    st = lib().orc_calcbeta(C.byref(grid_of(p)), C.byref(bp), *[dp(a) for a in arrs], dp(beta))
    return st, beta


# --- 2D (SSA) solve on a one-level grid ------------------------------------------
def grid2d_of(p):
    return grid(p.nx, p.ny, 1, p.periodic)


def mat2d_zeros(p):
    return [np.zeros((27, 1, p.nx - 1, p.ny - 1), order="F") for _ in range(4)]


def assemble_2d(p, u2, v2, whichefvs=None):
    """SSA matrix on the one-level grid (27, 1, nx-1, ny-1) (kA = 0 slots) and efvs"""
    g = grid2d_of(p)
    A = mat2d_zeros(p)
    efvs = np.zeros((p.nz - 1, p.nx, p.ny), order="F")
    st = lib().orc_assemble_2d_ssa(C.byref(g), p.nz, dp(p.sigma), ip(p.active_cell),
                                   dp(p.xVertex), dp(p.yVertex), dp(np.asfortranarray(u2)),
                                   dp(np.asfortranarray(v2)), dp(p.stagthck), dp(p.flwafact),
                                   p.whichefvs if whichefvs is None else whichefvs,
                                   p.efvs_constant, dp(efvs), *[dp(a) for a in A])
    return st, A, efvs


def picard_2d(p, loadu, loadv, maxiter=None, linear_maxiters=200, linear_tolerance=1e-8,
              whichsparse=None, whichprecond=None, whichresid=None, basal=None):
    """The SSA branch of the Picard loop from p.uvel/p.vvel (copied); 3D velocity out"""
    g = grid_of(p)
    o = PicardOpts()
    keep = []
    _set_basal(o, basal, keep)
    o.whichbabc = p.whichbabc
    o.whichefvs = p.whichefvs
    o.whichresid = p.whichresid if whichresid is None else whichresid
    o.whichsparse = p.whichsparse if whichsparse is None else whichsparse
    o.whichprecond = p.whichprecond if whichprecond is None else whichprecond
    o.whichapprox = 1
    o.maxiter_nonlinear = p.maxiter_nonlinear if maxiter is None else maxiter
    o.linear_maxiters = linear_maxiters
    o.efvs_constant = p.efvs_constant
    o.linear_tolerance = linear_tolerance
    o.dx, o.dy = p.dx, p.dy
    u, v = p.uvel.copy(order="F"), p.vvel.copy(order="F")
    efvs = np.zeros((p.nz - 1, p.nx, p.ny), order="F")
    s2 = (p.nx - 1, p.ny - 1)
    bu, bv, ru, rv = (np.zeros(s2, order="F") for _ in range(4))
    st = lib().orc_picard_solve_2d(C.byref(g), C.byref(o), dp(p.sigma), ip(p.active_cell),
                                   ip(p.active_vertex), dp(p.xVertex), dp(p.yVertex),
                                   dp(p.stagusrf), dp(p.stagthck), dp(p.flwafact),
                                   dp(p.beta_internal), ip(p.umask_dirichlet),
                                   ip(p.vmask_dirichlet), dp(loadu), dp(loadv), dp(u), dp(v),
                                   dp(efvs), dp(bu), dp(bv), dp(ru), dp(rv))
    return dict(status=st, uvel=u, vvel=v, efvs=efvs, bu=bu, bv=bv, resid_u=ru, resid_v=rv,
                picard_iters=o.picard_iters, linear_iters=o.total_linear_iters,
                L2_norm=o.final_L2_norm, resid_velo=o.final_resid_velo)


# --- post-solve diagnostics (velo_higher:3842-3894, 6755-7195) -----------------
def cell_halo(p, a):
    g = grid_of(p)
    lib().orc_cell_halo(C.byref(g), dp(a))


def internal_stress(p, uvel, vvel, efvs=None, whichapprox=None, whichefvs=None):
    """(tau_xz, tau_yz, tau_xx, tau_yy, tau_xy, tau_eff), each (nz-1, nx, ny); status"""
    g = grid_of(p)
    efvs = np.zeros((p.nz - 1, p.nx, p.ny), order="F") if efvs is None else efvs
    outs = [np.zeros((p.nz - 1, p.nx, p.ny), order="F") for _ in range(6)]
    st = lib().orc_internal_stress(
        C.byref(g), dp(p.sigma), ip(p.active_cell), dp(p.xVertex), dp(p.yVertex), dp(p.stagusrf),
        dp(p.stagthck), dp(p.flwafact), dp(efvs),
        p.whichefvs if whichefvs is None else whichefvs, p.efvs_constant,
        p.whichapprox if whichapprox is None else whichapprox, dp(uvel), dp(vvel),
        *[dp(o) for o in outs])
    return st, outs


def basal_friction_heatflx(p, ubas, vbas, beta=None, whichassemble_bfric=1):
    g = grid_of(p)
    out = np.zeros((p.nx, p.ny), order="F")
    beta = p.beta_internal if beta is None else beta
    lib().orc_basal_friction_heatflx(C.byref(g), ip(p.active_cell), ip(p.active_vertex),
                                     dp(np.asfortranarray(ubas)), dp(np.asfortranarray(vbas)),
                                     dp(beta), whichassemble_bfric, dp(out))
    return out


def basal_traction(p, uvel, vvel, beta=None):
    g = grid_of(p)
    beta = p.beta_internal if beta is None else beta
    bx = np.zeros((p.nx - 1, p.ny - 1), order="F")
    by = np.zeros((p.nx - 1, p.ny - 1), order="F")
    lib().orc_basal_traction(C.byref(g), dp(beta), dp(uvel), dp(vvel), dp(bx), dp(by))
    return bx, by


# --- depth-integrated 2D solvers: L1L2 and DIVA (velo_higher:5208-6356, 7382-7864) ----
def stagger(var, mask=None, margin=0):
    """glissade_stagger of a cell field (nx, ny) to the vertices (nx-1, ny-1)"""
    var = np.asfortranarray(var, dtype=np.float64)
    nx, ny = var.shape
    m = np.zeros((nx, ny), np.int32, order="F") if mask is None else np.asfortranarray(mask, dtype=np.int32)
    out = np.zeros((nx - 1, ny - 1), order="F")
    lib().orc_stagger(nx, ny, dp(var), dp(out), ip(m), margin)
    return out


def gradient(f, dx, dy, mask=None, margin=0):
    """glissade_gradient of a cell field: (df_dx, df_dy) at the vertices"""
    f = np.asfortranarray(f, dtype=np.float64)
    nx, ny = f.shape
    m = np.zeros((nx, ny), np.int32, order="F") if mask is None else np.asfortranarray(mask, dtype=np.int32)
    gx, gy = np.zeros((nx - 1, ny - 1), order="F"), np.zeros((nx - 1, ny - 1), order="F")
    lib().orc_gradient(nx, ny, dx, dy, dp(f), dp(gx), dp(gy), ip(m), margin)
    return gx, gy


def _v(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def efvs_l1l2(whichefvs, efvs_constant, sigma, phi, dphi_dx, dphi_dy, u, v, h, dsdx, dsdy, flwa, flwafact):
    """compute_effective_viscosity_L1L2 at one quadrature point: the nz-1 layer values"""
    sigma = _v(sigma)
    out = np.zeros(len(sigma) - 1)
    a = [_v(x) for x in (phi, dphi_dx, dphi_dy, u, v, h, dsdx, dsdy, flwa, flwafact)]
    lib().orc_efvs_l1l2(whichefvs, efvs_constant, len(sigma), dp(sigma), *[dp(x) for x in a], dp(out))
    return out


def efvs_diva(whichefvs, efvs_constant, stagsigma, phi, dphi_dx, dphi_dy, u, v, bx, by, flwafact, efvs_old):
    """compute_effective_viscosity_diva (cubic = .false.) at one quadrature point"""
    stagsigma = _v(stagsigma)
    out = _v(efvs_old).copy()
    a = [_v(x) for x in (phi, dphi_dx, dphi_dy, u, v, bx, by, flwafact)]
    lib().orc_efvs_diva(whichefvs, efvs_constant, len(stagsigma) + 1, dp(stagsigma), *[dp(x) for x in a],
                        dp(out))
    return out


def integrals_diva(sigma, thck, efvs_qp):
    """compute_integrals_diva: (omega_k (nz), omega); efvs_qp is (nz-1, 4)"""
    sigma = _v(sigma)
    e = np.asfortranarray(efvs_qp, dtype=np.float64)
    ok = np.zeros(len(sigma))
    om = np.zeros(1)
    lib().orc_integrals_diva(len(sigma), dp(sigma), float(thck), dp(e), dp(ok), dp(om))
    return ok, float(om[0])


def velocity_3d_diva(p, stag_omega_k, stag_omega, btx, bty, u2, v2, uvel=None, vvel=None):
    g = grid_of(p)
    u = np.zeros((p.nz, p.nx - 1, p.ny - 1), order="F") if uvel is None else uvel.copy(order="F")
    v = np.zeros((p.nz, p.nx - 1, p.ny - 1), order="F") if vvel is None else vvel.copy(order="F")
    f = lambda a: np.asfortranarray(a, dtype=np.float64)   # noqa: E731
    lib().orc_velocity_3d_diva(C.byref(g), ip(p.active_vertex), dp(f(stag_omega_k)), dp(f(stag_omega)),
                               dp(f(btx)), dp(f(bty)), dp(f(u2)), dp(f(v2)), dp(u), dp(v))
    return u, v


def velocity_3d_l1l2(p, ubas, vbas, efvs, umask2=None, vmask2=None):
    """compute_3d_velocity_L1L2 from the basal velocity (nx-1, ny-1) and the layer
    viscosities (nz-1, nx, ny); no halo update; (status, uvel, vvel)"""
    g = grid_of(p)
    u = np.zeros((p.nz, p.nx - 1, p.ny - 1), order="F")
    v = np.zeros((p.nz, p.nx - 1, p.ny - 1), order="F")
    u[-1], v[-1] = ubas, vbas
    z = np.zeros((p.nx - 1, p.ny - 1), np.int32, order="F")
    um = z if umask2 is None else np.asfortranarray(umask2, dtype=np.int32)
    vm = z if vmask2 is None else np.asfortranarray(vmask2, dtype=np.int32)
    st = lib().orc_velocity_3d_l1l2(C.byref(g), dp(p.sigma), p.dx, p.dy, ip(p.ice_mask), ip(p.active_cell),
                                    ip(p.active_vertex), ip(um), ip(vm), dp(p.xVertex), dp(p.yVertex),
                                    dp(p.thck), dp(p.stagthck), dp(p.dusrf_dx), dp(p.dusrf_dy), dp(p.flwa),
                                    dp(np.asfortranarray(efvs)), dp(u), dp(v))
    return st, u, v


def ice_plus_land(p):
    return np.asfortranarray((p.ice_mask != 0) | (p.land_mask != 0), dtype=np.int32)


def picard_depth(p, loadu, loadv, approx, maxiter=None, linear_maxiters=200, linear_tolerance=1e-8,
                 whichsparse=None, whichprecond=None, whichresid=None, basal=None, uvel_2d=None,
                 vvel_2d=None, btractx=None, btracty=None, efvs=None, target=0.0, assemble_beta=1):
    """The L1L2 (approx 3) or DIVA (approx 4) branch of the Picard loop from p.uvel/p.vvel
    (L1L2: their bed level; DIVA: uvel_2d/vvel_2d, btractx/btracty and efvs, zero by
    default)"""
    g = grid_of(p)
    o = PicardOpts()
    keep = []
    _set_basal(o, basal, keep)
    o.whichbabc = p.whichbabc
    o.whichefvs = p.whichefvs
    o.whichresid = p.whichresid if whichresid is None else whichresid
    o.whichsparse = p.whichsparse if whichsparse is None else whichsparse
    o.whichprecond = p.whichprecond if whichprecond is None else whichprecond
    o.whichapprox = approx
    o.maxiter_nonlinear = p.maxiter_nonlinear if maxiter is None else maxiter
    o.linear_maxiters = linear_maxiters
    o.efvs_constant = p.efvs_constant
    o.linear_tolerance = linear_tolerance
    o.dx, o.dy = p.dx, p.dy
    o.target_override = target
    o.assemble_beta_standard = 1 if assemble_beta == 0 else 0
    s2 = (p.nx - 1, p.ny - 1)
    z2 = lambda a: np.zeros(s2, order="F") if a is None else np.asfortranarray(a, dtype=np.float64).copy(order="F")  # noqa: E731
    u2, v2, bx, by = z2(uvel_2d), z2(vvel_2d), z2(btractx), z2(btracty)
    ipl = ice_plus_land(p)
    df = DepthFields(thck=dp(p.thck), dusrf_dx=dp(p.dusrf_dx), dusrf_dy=dp(p.dusrf_dy), flwa=dp(p.flwa),
                     ice_mask=ip(p.ice_mask), ice_plus_land_mask=ip(ipl), uvel_2d=dp(u2), vvel_2d=dp(v2),
                     btractx=dp(bx), btracty=dp(by))
    u, v = p.uvel.copy(order="F"), p.vvel.copy(order="F")
    ef = np.zeros((p.nz - 1, p.nx, p.ny), order="F") if efvs is None else efvs.copy(order="F")
    bu, bv, ru, rv = (np.zeros(s2, order="F") for _ in range(4))
    st = lib().orc_picard_solve_depth(C.byref(g), C.byref(o), C.byref(df), dp(p.sigma), ip(p.active_cell),
                                      ip(p.active_vertex), dp(p.xVertex), dp(p.yVertex), dp(p.stagthck),
                                      dp(p.flwafact), dp(p.beta_internal), ip(p.umask_dirichlet),
                                      ip(p.vmask_dirichlet), dp(loadu), dp(loadv), dp(u), dp(v), dp(ef),
                                      dp(bu), dp(bv), dp(ru), dp(rv))
    return dict(status=st, uvel=u, vvel=v, uvel_2d=u2, vvel_2d=v2, btractx=bx, btracty=by, efvs=ef, bu=bu,
                bv=bv, resid_u=ru, resid_v=rv, picard_iters=o.picard_iters, linear_iters=o.total_linear_iters,
                L2_norm=o.final_L2_norm, resid_velo=o.final_resid_velo, converged=o.converged)
