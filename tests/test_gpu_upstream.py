"""The solve's inputs computed on the GPU from the geometry (cism_glissade_velo_inputs)
against the CPU oracle (oracle/upstream_oracle.py)."""
import numpy as np
import pytest

import upstream_oracle as U
from cism_amd import configs, gpu as G

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["dome", "ismip", "grf", "sheet", "sheet_outflow"])
def test_inputs_match_oracle(name):
    p = {"dome": lambda: configs.dome(),
         "ismip": lambda: configs.ismip_hom_a(ewn=24, nsn=24, nz=11, L=12000.0),
         "grf": lambda: configs.grf_sheet(ewn=24, nsn=20, nz=7, seed=4),
         "sheet": lambda: configs.ice_sheet(ewn=36, nsn=42, nz=6, dx=33333.0, seed=7),
         "sheet_outflow": lambda: configs.ice_sheet(ewn=36, nsn=42, nz=6, dx=33333.0, seed=7)}[name]()
    periodic = name != "sheet_outflow"
    print(f"[upstream] {name} {p.ewn}x{p.nsn}x{p.nz} periodic={periodic}", flush=True)
    with G.Context(p.nx, p.ny, p.nz, periodic=periodic) as ctx:
        d = ctx.glissade_velo_inputs(p.dx, p.dy, p.thklim, 0.0, p.thck, p.usrf, p.topg, p.flwa)
    ice, fl, oc, la, act = U.get_masks(p.thck, p.topg, 0.0, p.thklim)
    for key, want in [("ice_mask", ice), ("floating_mask", fl), ("ocean_mask", oc), ("land_mask", la)]:
        assert np.array_equal(d[key], want), key
    ipl = ((act == 1) | (la == 1)).astype(np.int32)
    assert np.array_equal(d["stagthck"], U.stagger_margin1(p.thck, ipl))
    assert np.array_equal(d["stagusrf"], U.stagger_margin1(p.usrf, ipl))
    gx, gy = U.surface_gradient(p.usrf, p.thck, act, la, p.dx, p.dy, p.thklim, periodic=periodic)
    assert np.array_equal(d["dusrf_dx"], gx) and np.array_equal(d["dusrf_dy"], gy)
    xV, yV, ac, av = U.vertex_geometry(p.nx, p.ny, p.dx, p.dy, act)
    assert np.array_equal(d["xVertex"], xV) and np.array_equal(d["yVertex"], yV)
    assert np.array_equal(d["active_cell"], ac) and np.array_equal(d["active_vertex"], av)
    ff = U.flwafact(p.flwa)
    # pow on the device and in libm may differ in the last bit
    assert np.max(np.abs(d["flwafact"] - ff)) <= 2e-16 * np.max(np.abs(ff))
    assert np.array_equal(d["flwafact"] == 0, ff == 0)
