"""CPU oracle of the velocity solve's upstream inputs (numpy).

TEST INFRASTRUCTURE ONLY: the checker of the GPU kernels in
cism_amd/csrc/upstream.hip, used by tests/ alone.  Restates, with the
reference's loop ranges and operation order:

  get_masks            glissade_get_masks without a subgrid calving front
                       (libglissade/glissade_masks.F90:130-210)
  stagger_margin1      glissade_stagger, stagger_margin_in = 1
                       (libglissade/glissade_grid_operators.F90:118-146)
  surface_gradient     glissade_surface_elevation_gradient, HO_GRADIENT_CENTERED
                       with HO_GRADIENT_MARGIN_HYBRID and the max_slope limit
                       (glissade_grid_operators.F90:734-1168, branch 830-990)
  vertex_geometry      get_vertex_geometry (libglissade/glissade_velo_higher.F90:4163-4313)
  flwafact             0.5 A^(-1/n) on the cells the solve uses (velo_higher:1930-1939)
  stag_halo            staggered_parallel_halo (libglimmer/parallel_slap.F90:2807-2925)

Pins (tests/test_oracle_upstream.py): closed forms of each on fields whose
answer is known (linear surfaces, constant fields, hand-built masks), the
periodic index maps, and the invariant active vertex == corner of an active cell.
"""
from __future__ import annotations

import numpy as np

NHALO = 2
RHOI, RHOO = 910.0, 1028.0
GN = 3


def stag_halo(a, periodic=True):
    a = np.array(a, order="F", copy=True)
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
    """ice, floating, ocean, land, active (int32, cells)"""
    ice = (thck > thklim).astype(np.int32)
    ocean = ((topg < eus) & (ice == 0)).astype(np.int32)
    floating = ((topg - eus < (-RHOI / RHOO) * thck) & (ice == 1)).astype(np.int32)
    land = (topg >= eus).astype(np.int32)
    return ice, floating, ocean, land, ice.copy()


def stagger_margin1(var, mask):
    """vertex (i, j) = mean of var over the masked cells among (i..i+1, j..j+1), summed
    in the reference's order (i,j+1), (i+1,j+1), (i,j), (i+1,j)"""
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


def _edge_factor(thck_u, land_u, thklim):
    # (H - thklim)/0 with no ramp: 1 above thklim, 0 at or below it, on land edges only
    f = np.where(thck_u > thklim, 1.0, 0.0)
    return np.where(land_u == 1, f, 1.0)


def surface_gradient(usrf, thck, active, land, dx, dy, thklim, max_slope=1.0, periodic=True):
    """(ds/dx, ds/dy) at vertices: edge gradients where the upper cell is active and the
    lower one active or land (the upper cell's land edge factor), averaged to the vertex,
    limited to max_slope, halo by staggered_parallel_halo"""
    nx, ny = usrf.shape
    ex = np.zeros((nx, ny))
    up = usrf[:-1, :] > usrf[1:, :]
    s_u = np.where(up, usrf[:-1, :], usrf[1:, :])
    s_l = np.where(up, usrf[1:, :], usrf[:-1, :])
    sign = np.where(up, -1.0, 1.0)
    th_u = np.where(up, thck[:-1, :], thck[1:, :])
    la_u = np.where(up, land[:-1, :], land[1:, :])
    ac_u = np.where(up, active[:-1, :], active[1:, :])
    ac_l = np.where(up, active[1:, :], active[:-1, :])
    la_l = np.where(up, land[1:, :], land[:-1, :])
    ok = (ac_u == 1) & ((ac_l == 1) | (la_l == 1))
    ex[:-1, :] = np.where(ok, _edge_factor(th_u, la_u, thklim) * sign * (s_u - s_l) / dx, 0.0)
    ey = np.zeros((nx, ny))
    up = usrf[:, :-1] > usrf[:, 1:]
    s_u = np.where(up, usrf[:, :-1], usrf[:, 1:])
    s_l = np.where(up, usrf[:, 1:], usrf[:, :-1])
    sign = np.where(up, -1.0, 1.0)
    th_u = np.where(up, thck[:, :-1], thck[:, 1:])
    la_u = np.where(up, land[:, :-1], land[:, 1:])
    ac_u = np.where(up, active[:, :-1], active[:, 1:])
    ac_l = np.where(up, active[:, 1:], active[:, :-1])
    la_l = np.where(up, land[:, 1:], land[:, :-1])
    ok = (ac_u == 1) & ((ac_l == 1) | (la_l == 1))
    ey[:, :-1] = np.where(ok, _edge_factor(th_u, la_u, thklim) * sign * (s_u - s_l) / dy, 0.0)
    dsdx = 0.5 * (ex[:-1, :-1] + ex[:-1, 1:])
    dsdy = 0.5 * (ey[:-1, :-1] + ey[1:, :-1])
    dsdx = np.where(dsdx > 0.0, np.minimum(dsdx, max_slope), np.maximum(dsdx, -max_slope))
    dsdy = np.where(dsdy > 0.0, np.minimum(dsdy, max_slope), np.maximum(dsdy, -max_slope))
    return stag_halo(dsdx, periodic), stag_halo(dsdy, periodic)


def vertex_geometry(nx, ny, dx, dy, active_mask):
    """xVertex, yVertex, active_cell, active_vertex"""
    iv = np.arange(1, nx)[:, None]
    jv = np.arange(1, ny)[None, :]
    xV = np.asfortranarray(np.broadcast_to(dx * iv, (nx - 1, ny - 1)).astype(np.float64))
    yV = np.asfortranarray(np.broadcast_to(dy * jv, (nx - 1, ny - 1)).astype(np.float64))
    ac = np.zeros((nx, ny), dtype=np.int32, order="F")
    lo = NHALO
    ac[lo:nx - NHALO + 1, lo:ny - NHALO + 1] = active_mask[lo:nx - NHALO + 1, lo:ny - NHALO + 1] == 1
    av = np.zeros((nx - 1, ny - 1), dtype=np.int32, order="F")
    ci, cj = np.nonzero(ac)
    for di in (-1, 0):
        for dj in (-1, 0):
            av[ci + di, cj + dj] = 1
    return xV, yV, ac, av


def flwafact(flwa):
    """(nz-1, nx, ny): 0.5 A^(-1/3) on cells nhalo+1 .. n-nhalo+1 where A > 0, else 0"""
    nzm, nx, ny = flwa.shape
    ff = np.zeros((nzm, nx, ny), order="F")
    sl = (slice(None), slice(NHALO, nx - NHALO + 1), slice(NHALO, ny - NHALO + 1))
    f = flwa[sl]
    ff[sl] = np.where(f > 0.0, 0.5 * np.power(np.where(f > 0, f, 1.0), -1.0 / float(GN)), 0.0)
    return ff
