"""Synthetic inputs for the Blatter-Pattyn velocity solve.

Builds, on the host and in the reference's Fortran array layouts, every array
that ``glissade_velo_higher_solve`` prepares before its Picard loop
(libglissade/glissade_velo_higher.F90:1095-2160): sigma levels, masks,
staggered thickness and surface, surface gradient, vertex geometry, the
Glen flow factor, basal traction, Dirichlet masks and the gravitational load.

Restated reference routines (host-side input preparation):
  glide_calc_sigma            libglide/glide_setup.F90:375-382, 283-290
  stagwbndsigma               libglide/glide_setup.F90:355-363
  glissade_get_masks          libglissade/glissade_masks.F90:55-210 (no subgrid calving front)
  glissade_stagger            libglissade/glissade_grid_operators.F90:60-148 (stagger_margin 1)
  surface gradient            libglissade/glissade_grid_operators.F90:734-1168
                              (HO_GRADIENT_CENTERED, HO_GRADIENT_MARGIN_HYBRID, max_slope)
  get_vertex_geometry         libglissade/glissade_velo_higher.F90:4163-4313
  flwafact                    libglissade/glissade_velo_higher.F90:1930-1939
  Dirichlet masks             libglissade/glissade_velo_higher.F90:1269-1324

Arrays are numpy arrays in Fortran order with the reference's shapes:
cell fields (nx, ny), vertex fields (nx-1, ny-1), node fields (nz, nx-1, ny-1),
layer fields (nz-1, nx, ny).  ``nx = ewn + 4`` (two halo cells each side).
Periodic halos of cell fields are filled as parallel_halo does.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

RHOI = 910.0
RHOO = 1028.0
GRAV = 9.81
SCYR = 31536000.0
GN = 3
NHALO = 2

# option enums (libglide/glide_types.F90:198-283)
HO_EFVS_CONSTANT, HO_EFVS_FLOWFACT, HO_EFVS_NONLINEAR = 0, 1, 2
HO_BABC_BETA_LARGE, HO_BABC_BETA_EXTERNAL, HO_BABC_NO_SLIP = 4, 5, 6
HO_RESID_MAXU, HO_RESID_MAXU_NO_UBAS, HO_RESID_MEANU, HO_RESID_L2NORM, HO_RESID_L2NORM_RELATIVE = 0, 1, 2, 3, 4
HO_SPARSE_PCG_STANDARD, HO_SPARSE_PCG_CHRONGEAR = 2, 3
HO_APPROX_SIA, HO_APPROX_SSA, HO_APPROX_BP = 0, 1, 2
HO_PRECOND_NONE, HO_PRECOND_DIAG, HO_PRECOND_SIA = 0, 1, 2


def fzeros(shape, dtype=np.float64):
    return np.zeros(shape, dtype=dtype, order="F")


def glide_sigma(nz: int) -> np.ndarray:
    """Glide default sigma levels (glide_setup.F90:283-290, 375-382)."""
    n = 2.0
    s = np.empty(nz)
    for up in range(1, nz + 1):
        x = float(up - 1) / float(nz - 1)
        s[up - 1] = (1 - (x + 1) ** (-n)) / (1 - 2 ** (-n))
    return s


def stagwbnd(sigma: np.ndarray) -> np.ndarray:
    """stagwbndsigma(0:nz): 0, layer midpoints, 1 (glide_setup.F90:355-363)."""
    nz = len(sigma)
    w = np.empty(nz + 1)
    w[0] = 0.0
    w[1:nz] = (sigma[:-1] + sigma[1:]) / 2.0
    w[nz] = 1.0
    return w


def pad_periodic(a: np.ndarray) -> np.ndarray:
    """(ewn, nsn) owned cell field -> (ewn+4, nsn+4) with periodic halos."""
    return np.asfortranarray(np.pad(a, NHALO, mode="wrap"))


def stag_halo(a: np.ndarray, periodic: bool = True) -> np.ndarray:
    """staggered_parallel_halo on the last two axes (parallel_slap.F90:2807-2925)."""
    a = np.asfortranarray(a)
    sx, sy = a.shape[-2], a.shape[-1]
    if periodic:
        a[..., sx - 1, 2:sy - 1] = a[..., 2, 2:sy - 1]
        a[..., 0:2, 2:sy - 1] = a[..., sx - 3:sx - 1, 2:sy - 1]
        a[..., :, sy - 1] = a[..., :, 2]
        a[..., :, 0:2] = a[..., :, sy - 3:sy - 1]
    else:
        a[..., 0, :] = 0
        a[..., sx - 1, :] = 0
        a[..., :, 0] = 0
        a[..., :, sy - 1] = 0
    return a


def get_masks(thck, topg, eus, thklim):
    """glissade_get_masks without a subgrid calving front (glissade_masks.F90:130-210)."""
    ice = (thck > thklim).astype(np.int32)
    ocean = ((topg < eus) & (ice == 0)).astype(np.int32)
    floating = ((topg - eus < (-RHOI / RHOO) * thck) & (ice == 1)).astype(np.int32)
    land = (topg >= eus).astype(np.int32)
    active = ice.copy()
    return (np.asfortranarray(ice), np.asfortranarray(floating), np.asfortranarray(ocean),
            np.asfortranarray(land), np.asfortranarray(active))


def stagger_margin1(var, mask):
    """glissade_stagger with stagger_margin_in = 1 (glissade_grid_operators.F90:118-146)."""
    m = mask.astype(np.float64)
    a = m[:-1, 1:] * var[:-1, 1:]
    b = m[1:, 1:] * var[1:, 1:]
    c = m[:-1, :-1] * var[:-1, :-1]
    d = m[1:, :-1] * var[1:, :-1]
    sumvar = ((a + b) + c) + d
    summask = (mask[:-1, 1:] + mask[1:, 1:] + mask[:-1, :-1] + mask[1:, :-1]).astype(np.float64)
    out = np.zeros_like(sumvar)
    pos = summask > 0
    out[pos] = sumvar[pos] / summask[pos]
    return np.asfortranarray(out)


def _edge_factor(thck_u, land_u, thklim, ramp):
    if ramp > 0:
        f = np.minimum(1.0, (thck_u - thklim) / ramp)
        f = np.maximum(f, 0.0)
    else:
        # (H - thklim)/0 is +Inf (-> 1) for H > thklim and -Inf (-> 0) below; the
        # reference's 0/0 case at H == thklim exactly is taken as 0
        f = np.where(thck_u > thklim, 1.0, 0.0)
    return np.where(land_u == 1, f, 1.0)


def surface_gradient(usrf, thck, active, land, dx, dy, thklim, ramp=0.0, max_slope=1.0,
                     periodic=True):
    """HO_GRADIENT_CENTERED with HO_GRADIENT_MARGIN_HYBRID (glissade_grid_operators.F90:830-990)."""
    nx, ny = usrf.shape
    ex = np.zeros((nx, ny))
    up_is_left = usrf[:-1, :] > usrf[1:, :]
    s_u = np.where(up_is_left, usrf[:-1, :], usrf[1:, :])
    s_l = np.where(up_is_left, usrf[1:, :], usrf[:-1, :])
    sign = np.where(up_is_left, -1.0, 1.0)
    th_u = np.where(up_is_left, thck[:-1, :], thck[1:, :])
    la_u = np.where(up_is_left, land[:-1, :], land[1:, :])
    ac_u = np.where(up_is_left, active[:-1, :], active[1:, :])
    ac_l = np.where(up_is_left, active[1:, :], active[:-1, :])
    la_l = np.where(up_is_left, land[1:, :], land[:-1, :])
    ef = _edge_factor(th_u, la_u, thklim, ramp)
    ok = (ac_u == 1) & ((ac_l == 1) | (la_l == 1))
    ex[:-1, :] = np.where(ok, ef * sign * (s_u - s_l) / dx, 0.0)

    ey = np.zeros((nx, ny))
    up_is_low = usrf[:, :-1] > usrf[:, 1:]
    s_u = np.where(up_is_low, usrf[:, :-1], usrf[:, 1:])
    s_l = np.where(up_is_low, usrf[:, 1:], usrf[:, :-1])
    sign = np.where(up_is_low, -1.0, 1.0)
    th_u = np.where(up_is_low, thck[:, :-1], thck[:, 1:])
    la_u = np.where(up_is_low, land[:, :-1], land[:, 1:])
    ac_u = np.where(up_is_low, active[:, :-1], active[:, 1:])
    ac_l = np.where(up_is_low, active[:, 1:], active[:, :-1])
    la_l = np.where(up_is_low, land[:, 1:], land[:, :-1])
    ef = _edge_factor(th_u, la_u, thklim, ramp)
    ok = (ac_u == 1) & ((ac_l == 1) | (la_l == 1))
    ey[:, :-1] = np.where(ok, ef * sign * (s_u - s_l) / dy, 0.0)

    dsdx = 0.5 * (ex[:-1, :-1] + ex[:-1, 1:])
    dsdy = 0.5 * (ey[:-1, :-1] + ey[1:, :-1])
    dsdx = np.where(dsdx > 0.0, np.minimum(dsdx, max_slope), np.maximum(dsdx, -max_slope))
    dsdy = np.where(dsdy > 0.0, np.minimum(dsdy, max_slope), np.maximum(dsdy, -max_slope))
    return stag_halo(np.asfortranarray(dsdx), periodic), stag_halo(np.asfortranarray(dsdy), periodic)


def vertex_geometry(nx, ny, dx, dy, active_ice_mask):
    """get_vertex_geometry (velo_higher:4163-4313): coordinates, active cells and vertices."""
    iv = np.arange(1, nx)[:, None]
    jv = np.arange(1, ny)[None, :]
    xV = np.asfortranarray(np.broadcast_to(dx * iv, (nx - 1, ny - 1)).astype(np.float64))
    yV = np.asfortranarray(np.broadcast_to(dy * jv, (nx - 1, ny - 1)).astype(np.float64))
    ac = np.zeros((nx, ny), dtype=np.int32, order="F")
    lo = NHALO  # 0-based index of cell nhalo+1
    ac[lo:nx - NHALO + 1, lo:ny - NHALO + 1] = active_ice_mask[lo:nx - NHALO + 1, lo:ny - NHALO + 1] == 1
    av = np.zeros((nx - 1, ny - 1), dtype=np.int32, order="F")
    ci, cj = np.nonzero(ac)
    for di in (-1, 0):
        for dj in (-1, 0):
            av[ci + di, cj + dj] = 1
    return xV, yV, ac, av


def flwafact_of(flwa, nz, nx, ny):
    """0.5 * A^(-1/n) on cells nhalo+1 .. n-nhalo+1 (velo_higher:1930-1939)."""
    ff = fzeros((nz - 1, nx, ny))
    sl = (slice(None), slice(NHALO, nx - NHALO + 1), slice(NHALO, ny - NHALO + 1))
    f = flwa[sl]
    ff[sl] = np.where(f > 0.0, 0.5 * np.power(np.where(f > 0, f, 1.0), -1.0 / float(GN)), 0.0)
    return ff


@dataclass
class Problem:
    """Everything the Picard loop reads, in reference layouts and units (m, yr, Pa)."""
    name: str
    ewn: int
    nsn: int
    nz: int
    dx: float
    dy: float
    periodic: bool
    sigma: np.ndarray
    stagwbndsigma: np.ndarray
    thck: np.ndarray
    usrf: np.ndarray
    topg: np.ndarray
    thklim: float
    active_cell: np.ndarray
    active_vertex: np.ndarray
    xVertex: np.ndarray
    yVertex: np.ndarray
    stagusrf: np.ndarray
    stagthck: np.ndarray
    dusrf_dx: np.ndarray
    dusrf_dy: np.ndarray
    flwa: np.ndarray
    flwafact: np.ndarray
    beta_internal: np.ndarray
    umask_dirichlet: np.ndarray
    vmask_dirichlet: np.ndarray
    uvel: np.ndarray
    vvel: np.ndarray
    whichbabc: int
    ice_mask: np.ndarray = None
    floating_mask: np.ndarray = None
    ocean_mask: np.ndarray = None
    land_mask: np.ndarray = None
    whichefvs: int = HO_EFVS_NONLINEAR
    whichapprox: int = HO_APPROX_BP
    whichresid: int = HO_RESID_L2NORM
    whichsparse: int = HO_SPARSE_PCG_CHRONGEAR
    whichprecond: int = HO_PRECOND_SIA
    whichassemble_taud: int = 1
    whichassemble_beta: int = 1
    efvs_constant: float = 2336041.42829
    maxiter_nonlinear: int = 100
    extra: dict = field(default_factory=dict)

    @property
    def nx(self):
        return self.ewn + 2 * NHALO

    @property
    def ny(self):
        return self.nsn + 2 * NHALO

    @property
    def nnodes(self):
        return self.nz * (self.nx - 1) * (self.ny - 1)

    def owned_nodes(self):
        return self.ewn * self.nsn * self.nz


def _finish(name, ewn, nsn, nz, dx, dy, thck, usrf, topg, thklim, flwa_val, whichbabc,
            beta_vertex=None, periodic=True, flwa=None, f_ground=None, settled=False, **kw) -> Problem:
    nx, ny = ewn + 2 * NHALO, nsn + 2 * NHALO
    sigma = glide_sigma(nz)
    eus = 0.0
    ice, floating, ocean, land, active = get_masks(thck, topg, eus, thklim)
    ipl = ((active == 1) | (land == 1)).astype(np.int32)
    stagthck = stagger_margin1(thck, ipl)
    stagusrf = stagger_margin1(usrf, ipl)
    dsdx, dsdy = surface_gradient(usrf, thck, active, land, dx, dy, thklim, periodic=periodic)
    xV, yV, ac, av = vertex_geometry(nx, ny, dx, dy, active)
    if flwa is None:
        flwa = np.full((nz - 1, nx, ny), flwa_val, order="F")
    ff = flwafact_of(flwa, nz, nx, ny)
    if settled:                         # np.power, as settle() explains
        ff = np.asfortranarray(settle(ff))
    beta = fzeros((nx - 1, ny - 1))
    if whichbabc == HO_BABC_BETA_LARGE:
        beta[:] = 1.0e10
    elif whichbabc == HO_BABC_BETA_EXTERNAL and beta_vertex is not None:
        beta[:] = beta_vertex
    if f_ground is not None:            # calcbeta: beta * f_ground, floored at
        beta[:] = beta * f_ground       # beta_grounded_min = 1 Pa yr/m where f_ground > 0
        beta[:] = np.where((f_ground > 0) & (beta < 1.0), 1.0, beta)
    um = np.zeros((nz, nx - 1, ny - 1), dtype=np.int32, order="F")
    vm = np.zeros((nz, nx - 1, ny - 1), dtype=np.int32, order="F")
    if whichbabc == HO_BABC_NO_SLIP:
        um[nz - 1, :, :] = 1
        vm[nz - 1, :, :] = 1
    p = Problem(name=name, ewn=ewn, nsn=nsn, nz=nz, dx=dx, dy=dy, periodic=periodic,
                sigma=sigma, stagwbndsigma=stagwbnd(sigma), thck=thck, usrf=usrf, topg=topg,
                thklim=thklim, active_cell=ac, active_vertex=av, xVertex=xV, yVertex=yV,
                stagusrf=stagusrf, stagthck=stagthck, dusrf_dx=dsdx, dusrf_dy=dsdy,
                flwa=flwa, flwafact=ff, beta_internal=beta, umask_dirichlet=um,
                vmask_dirichlet=vm, uvel=fzeros((nz, nx - 1, ny - 1)),
                vvel=fzeros((nz, nx - 1, ny - 1)), whichbabc=whichbabc, ice_mask=ice,
                floating_mask=floating, ocean_mask=ocean, land_mask=land, **kw)
    if beta_vertex is not None:
        # calcbeta's beta_external (which_ho_babc 5): beta_internal before the f_ground weight
        p.extra["beta_external"] = np.asfortranarray(beta_vertex, dtype=np.float64)
    return p


def dome(ewn=31, nsn=31, nz=10, dx=2000.0, R=20000.0, H0=2000.0, A=1.0e-16, thklim=1.0,
         whichsparse=HO_SPARSE_PCG_STANDARD, whichprecond=HO_PRECOND_DIAG, **kw) -> Problem:
    """Parabolic dome, H = H0*sqrt(1-(r/R)^2) about the centre cell, flat bed, no-slip base."""
    ic, jc = (ewn + 1) // 2, (nsn + 1) // 2          # 1-based centre cell (16,16 for 31x31)
    ii = np.arange(1, ewn + 1)[:, None]
    jj = np.arange(1, nsn + 1)[None, :]
    r2 = ((ii - ic) * dx) ** 2 + ((jj - jc) * dx) ** 2
    q = 1.0 - r2 / R ** 2
    H = np.where(q > 0.0, H0 * np.sqrt(np.where(q > 0, q, 0.0)), 0.0)
    thck = pad_periodic(H)
    topg = fzeros(thck.shape)
    usrf = np.asfortranarray(topg + thck)
    return _finish("dome", ewn, nsn, nz, dx, dx, thck, usrf, topg, thklim, A, HO_BABC_NO_SLIP,
                   whichsparse=whichsparse, whichprecond=whichprecond, **kw)


def ismip_hom_a(ewn=160, nsn=160, nz=21, L=80000.0, A=1.0e-16, thklim=100.0,
                whichsparse=HO_SPARSE_PCG_CHRONGEAR, whichprecond=HO_PRECOND_SIA, **kw) -> Problem:
    """ISMIP-HOM experiment A: s = -x tan(0.5 deg), b = s - 1000 + 500 sin sin, frozen bed.

    Cell centres x = (i + 1/2) dx with i the 0-based global index (runISMIP_HOM.py:287-344).
    Halo cells are exact copies of their periodic images (parallel_halo), the surface
    shifted by the east-west offset.
    """
    dx = L / ewn
    x = (np.arange(ewn) + 0.5) * dx
    y = (np.arange(nsn) + 0.5) * dx
    om = 2.0 * math.pi / L
    alpha = 0.5 * math.pi / 180.0
    H = 1000.0 - 500.0 * np.sin(om * x)[:, None] * np.sin(om * y)[None, :]
    s = np.broadcast_to(-x[:, None] * math.tan(alpha), (ewn, nsn))
    thck = pad_periodic(H)
    # halo copies of the surface carry the east-west offset L tan(alpha)
    # (periodic_offset_ew, runISMIP_HOM.py:220-227), so it continues linearly
    usrf = pad_periodic(s)
    off = ewn * dx * math.tan(alpha)
    usrf[:NHALO, :] += off
    usrf[-NHALO:, :] -= off
    topg = np.asfortranarray(usrf - thck)
    return _finish("ismip_hom_a", ewn, nsn, nz, dx, dx, thck, usrf, topg, thklim, A,
                   HO_BABC_NO_SLIP, whichsparse=whichsparse, whichprecond=whichprecond, **kw)


def settle(a, decimals=3):
    """Rounds a generated field to a decimal grid (1 mm for 3 decimals).  numpy's exp, log
    and power take different SIMD paths on different CPUs and can differ in the last bit;
    rounding makes the synthetic inputs bit-identical from one host to the next, so a
    benchmark's iteration counts do not depend on which machine generated its problem."""
    return np.round(np.asarray(a, dtype=np.float64), decimals)


def gaussian_random_field(n1, n2, dx, corr, rng):
    """Periodic Gaussian random field with Gaussian covariance exp(-r^2/(2 corr^2)), unit variance."""
    k1 = np.fft.fftfreq(n1, d=dx) * 2 * math.pi
    k2 = np.fft.fftfreq(n2, d=dx) * 2 * math.pi
    kk = k1[:, None] ** 2 + k2[None, :] ** 2
    spec = np.exp(-0.25 * kk * corr ** 2)
    w = rng.standard_normal((n1, n2))
    f = np.real(np.fft.ifft2(np.fft.fft2(w) * np.sqrt(spec)))
    f -= f.mean()
    sd = f.std()
    return f / sd if sd > 0 else f


def grf_sheet(ewn=512, nsn=512, nz=11, dx=1000.0, seed=1, mean=1500.0, sd=400.0, corr=30000.0,
              beta_median=1.0e3, beta_sigma=1.0, A=1.0e-16, thklim=1.0,
              whichsparse=HO_SPARSE_PCG_CHRONGEAR, whichprecond=HO_PRECOND_SIA, **kw) -> Problem:
    """Ice sheet of GRF thickness on a flat bed, lognormal beta at vertices (babc 5).

    periodic=False gives the outflow form of the same sheet: the halo cells are zero, as
    parallel_halo leaves them on an outflow domain (parallel_slap.F90:1794-1860), so the ice
    ends at the domain edge, the owned vertices include the west and south boundary rows
    (staggered_ilo = staggered_lhalo, parallel_slap.F90:661-673) and the staggered halo
    is zeroed beyond the boundary (parallel_slap.F90:2856-2863).  The thickness is
    thinned over the last edge_cells cells toward the edge."""
    periodic = kw.pop("periodic", True)
    edge_cells = kw.pop("edge_cells", 5.0)
    rng = np.random.default_rng(seed)
    H = settle(mean + sd * gaussian_random_field(ewn, nsn, dx, corr, rng))
    if not periodic:
        # thinned toward the domain edge (to about 1 % in the boundary cells), so the ice
        # that ends there is a thin margin, not a cliff
        di = np.minimum(np.arange(ewn), np.arange(ewn)[::-1])[:, None]
        dj = np.minimum(np.arange(nsn), np.arange(nsn)[::-1])[None, :]
        H = settle(H * np.minimum(1.0, (np.minimum(di, dj) + 0.5) / edge_cells) ** 2)
    H = np.maximum(H, 10.0)
    thck = pad_periodic(H) if periodic else np.asfortranarray(np.pad(H, NHALO))
    topg = fzeros(thck.shape)
    usrf = np.asfortranarray(topg + thck)
    bo = settle(np.exp(np.log(beta_median) + beta_sigma * rng.standard_normal((ewn, nsn))))
    bv = np.zeros((ewn + 3, nsn + 3))
    bv[2:ewn + 2, 2:nsn + 2] = bo          # owned vertices 3..nx-2 (1-based)
    if not periodic:                       # and the west/south boundary rows 2 (1-based)
        bv[1, 1:nsn + 2] = bv[2, 1:nsn + 2]
        bv[1:ewn + 2, 1] = bv[1:ewn + 2, 2]
    bv = stag_halo(np.asfortranarray(bv), periodic)
    return _finish("grf_sheet" if periodic else "grf_sheet_outflow", ewn, nsn, nz, dx, dx, thck,
                   usrf, topg, thklim, A, HO_BABC_BETA_EXTERNAL, beta_vertex=bv,
                   whichsparse=whichsparse, whichprecond=whichprecond, settled=True,
                   periodic=periodic, **kw)


def slab(ewn=40, nsn=40, nz=11, dx=1000.0, H=1000.0, theta_deg=0.5, efvs=2336041.42829,
         **kw) -> Problem:
    """Inclined slab of uniform thickness, linear viscosity, frozen bed (tests/slab analogue)."""
    nx, ny = ewn + 2 * NHALO, nsn + 2 * NHALO
    ig = np.arange(nx) - NHALO
    x = (ig + 0.5) * dx
    th = math.radians(theta_deg)
    thck = np.full((nx, ny), H, order="F")
    usrf = np.asfortranarray(np.broadcast_to(-x[:, None] * math.tan(th), (nx, ny)).copy())
    topg = np.asfortranarray(usrf - thck)
    return _finish("slab", ewn, nsn, nz, dx, dx, thck, usrf, topg, 1.0, 1.0e-16,
                   HO_BABC_NO_SLIP, whichefvs=HO_EFVS_CONSTANT, efvs_constant=efvs, **kw)


def sliding_slab(ewn=24, nsn=20, nz=11, dx=1000.0, H=1000.0, theta_deg=0.5, beta=500.0,
                 A=1.0e-16, **kw) -> Problem:
    """Inclined slab of uniform thickness sliding over a uniform-beta bed (babc 5), periodic:
    every horizontal gradient of the solution vanishes, so the SSA solution is the basal
    balance u = rho_i g H tan(theta) / beta, v = 0."""
    nx, ny = ewn + 2 * NHALO, nsn + 2 * NHALO
    ig = np.arange(nx) - NHALO
    x = (ig + 0.5) * dx
    th = math.radians(theta_deg)
    thck = np.full((nx, ny), H, order="F")
    usrf = np.asfortranarray(np.broadcast_to(-x[:, None] * math.tan(th), (nx, ny)).copy())
    topg = np.asfortranarray(usrf - thck)
    bv = np.full((nx - 1, ny - 1), beta, order="F")
    return _finish("sliding_slab", ewn, nsn, nz, dx, dx, thck, usrf, topg, 1.0, A,
                   HO_BABC_BETA_EXTERNAL, beta_vertex=bv, **kw)


def grounded_fraction_vertex(ice, floating, land):
    """glissade_grounded_fraction with HO_GROUND_NO_GLP (glissade_grounding_line.F90:55-178):
    1 at a vertex next to land or grounded ice, else 0; the vertex mask tests
    land(i+1, j+1) twice and never land(i+1, j), as the reference's condition does."""
    ice, floating, land = (np.asarray(a) == 1 for a in (ice, floating, land))
    vm = (ice[:-1, 1:] | ice[1:, 1:] | ice[:-1, :-1] | ice[1:, :-1]
          | land[:-1, 1:] | land[1:, 1:] | land[:-1, :-1])
    cg = (ice & ~floating) | land
    gv = cg[:-1, 1:] | cg[1:, 1:] | cg[:-1, :-1] | cg[1:, :-1]
    return np.asfortranarray(np.where(vm & gv, 1.0, 0.0))


def ice_sheet(ewn=1200, nsn=1400, nz=11, dx=1000.0, seed=5, H_max=3500.0, H_edge=300.0, bed_sd=600.0,
              s_max=3300.0, shelf=0.0, beta_median=1.0e3, beta_sigma=1.0, A=1.0e-17,
              thklim=100.0, whichsparse=HO_SPARSE_PCG_CHRONGEAR, whichprecond=HO_PRECOND_SIA,
              **kw) -> Problem:
    """Large single-GPU ice sheet on an elliptical footprint (normalised radius r = 1 at
    0.42 of the domain).  Grounded part: a Vialov-type surface s = s_max (1 - r^(4/3))^(3/8)
    with a long-wavelength (150 km) random perturbation, over a bed that deepens toward
    the edge (400 - 1400 r^2 m) and carries random relief at 50 km and 15 km (sd bed_sd),
    so the thickness H = s - b is multiscale and random in [0, H_max] while the surface
    stays smooth (troughs deeper than the cap H_edge + (H_max - H_edge) (1 - r^(4/3))^(3/8)
    are filled, so the margin is at most H_edge = 300 m thick).  Flow factor A = 1e-17
    Pa^-3 yr^-1 (cold ice), which keeps a 3 km thick sheet's speeds glaciological.
    Floating fringe: ice is afloat wherever rho_i H < rho_o (-b) (thin margin ice over
    the deepening bed), with surface H (1 - rho_i / rho_o).  Optionally (shelf > 0) a
    shelf of width `shelf` (in r) beyond the grounded margin thins linearly from 500 m
    to 150 m where the sea is deep enough to float it.  It is off by default: floating
    ice has no basal drag, so its depth-uniform (plug) modes are held only by membrane
    stresses and the front corners by a single element; a 60 km shelf drives the
    Jacobi-scaled condition number of the Blatter-Pattyn system from about 1e4 to 1e8,
    and at 1 km the plain FP64 Picard loop (200 PCG iterations per solve) no longer
    settles.  Ice-free ocean beyond; lognormal beta on grounded vertices
    weighted by the grounded fraction; lateral ice-ocean pressure at the margin.  Floating
    ice not connected to grounded ice is removed (an unsupported floating patch has no
    unique velocity)."""
    from scipy import ndimage
    rng = np.random.default_rng(seed)
    Lx, Ly = ewn * dx, nsn * dx
    x = (np.arange(ewn) + 0.5) * dx - Lx / 2
    y = (np.arange(nsn) + 0.5) * dx - Ly / 2
    rr = np.sqrt((x[:, None] / (0.42 * Lx)) ** 2 + (y[None, :] / (0.42 * Ly)) ** 2)
    prof = np.clip(1.0 - np.clip(rr, 0.0, 1.0) ** (4.0 / 3.0), 0.0, None) ** (3.0 / 8.0)
    s = s_max * prof * (1.0 + 0.08 * gaussian_random_field(ewn, nsn, dx, 150e3, rng))
    bed = (400.0 - 1400.0 * rr ** 2
           + bed_sd * (0.8 * gaussian_random_field(ewn, nsn, dx, 50e3, rng)
                       + 0.6 * gaussian_random_field(ewn, nsn, dx, 15e3, rng)))
    # thickness where the smooth surface meets the rough bed; where it would exceed the
    # cap H_edge + (H_max - H_edge) * prof the bed is raised (troughs filled), so the
    # surface stays smooth and the margin is at most H_edge thick: no kilometre-high
    # cliffs at the ice edge, whose lateral ocean load on a single element per corner
    # drives the solve to 1e5 m/yr corner speeds
    bed = np.where(rr < 1.0, np.maximum(bed, s - (H_edge + (H_max - H_edge) * prof)), bed)
    s, bed = settle(s), settle(bed)
    H = np.where(rr < 1.0, np.clip(s - bed, 0.0, None), 0.0)
    H_shelf = np.where((rr >= 1.0) & (rr < 1.0 + shelf), 500.0 - 350.0 * (rr - 1.0) / max(shelf, 1e-12), 0.0)
    H = np.where((H <= thklim) & (RHOI * H_shelf < RHOO * (-bed)), H_shelf, H)
    fl = RHOI * H < RHOO * (-bed)
    # trim floating filaments: a floating cell is kept only inside a 2 x 2 block of ice.
    # A one-cell-wide floating strip, or a floating cell hanging on by a corner, bends or
    # turns with almost no resistance (no basal drag), which leaves near-zero eigenvalues
    # (Jacobi-scaled condition numbers of 1e7 and more) and Picard loops that do not settle
    for _ in range(64):
        ice = H > thklim
        blk = ice[:-1, :-1] & ice[1:, :-1] & ice[:-1, 1:] & ice[1:, 1:]
        sup = np.zeros(ice.shape, bool)
        sup[:-1, :-1] |= blk; sup[1:, :-1] |= blk; sup[:-1, 1:] |= blk; sup[1:, 1:] |= blk
        drop = ice & fl & ~sup
        if not drop.any():
            break
        H = np.where(drop, 0.0, H)
    grounded = (H > thklim) & ~fl
    lab, _ = ndimage.label(H > thklim)
    keep = np.unique(lab[grounded])
    keep = keep[keep > 0]
    H = np.where(np.isin(lab, keep), H, 0.0)
    s = np.where(RHOI * H < RHOO * (-bed), H * (1.0 - RHOI / RHOO), bed + H)
    thck = pad_periodic(H)
    topg = pad_periodic(bed)
    usrf = pad_periodic(s)
    ice, floating, _, land, _ = get_masks(thck, topg, 0.0, thklim)
    fg = grounded_fraction_vertex(ice, floating, land)
    bo = settle(np.exp(np.log(beta_median) + beta_sigma * rng.standard_normal((ewn, nsn))))
    bv = np.zeros((ewn + 3, nsn + 3))
    bv[2:ewn + 2, 2:nsn + 2] = bo
    bv = stag_halo(np.asfortranarray(bv), True)
    return _finish("ice_sheet", ewn, nsn, nz, dx, dx, thck, usrf, topg, thklim, A,
                   HO_BABC_BETA_EXTERNAL, beta_vertex=bv, f_ground=fg, settled=True,
                   whichsparse=whichsparse, whichprecond=whichprecond, **kw)


def build(name: str, **kw) -> Problem:
    if name in ("dome", "c1"):
        return dome(**kw)
    if name in ("ismip_hom_a", "c2"):
        return ismip_hom_a(**kw)
    if name in ("grf_sheet", "c3"):
        return grf_sheet(**kw)
    if name == "slab":
        return slab(**kw)
    if name in ("ice_sheet", "c5"):
        return ice_sheet(**kw)
    raise ValueError(f"unknown problem {name}")
