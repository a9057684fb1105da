"""The FP64 solver reading the FP64 matrix from its symmetric copy (cism_set_fp64_symmetric).

Each Picard iteration's symmetry pass (glissade_chk_symmetry, velo_higher:9111-9234)
leaves every coefficient pair of an owned active row holding one value unless it counts
an unrepaired pair; the pass then also writes the 14 slots of the symmetric storage, and
the FP64 PCG reads a coupling before the diagonal as the transpose of its neighbour's.
The products are the full storage's, term by term and in the same order, and the dot
partials sit in the same node-block order, so the whole solve must be the same bit for
bit: the check here is exact equality against the same solve on the full storage.
"""
import numpy as np
import pytest

import oracle as O
from cism_amd import configs

pytestmark = pytest.mark.gpu


def cases():
    sheet_std = configs.ice_sheet(ewn=36, nsn=42, nz=6, dx=33333.0, seed=7)
    sheet_std.whichassemble_beta = 0   # the standard basal assembly
    return {
        # periodic, no-slip bed: Dirichlet rows and column eliminations
        "ismip": configs.ismip_hom_a(ewn=24, nsn=24, nz=11, L=12000.0),
        "dome": configs.dome(),
        "sheet": configs.ice_sheet(ewn=36, nsn=42, nz=6, dx=33333.0, seed=7),
        "shelf": configs.ice_sheet(ewn=36, nsn=42, nz=6, dx=33333.0, seed=7, shelf=0.02),
        "sheet_standard_basal": sheet_std,
        "deep": configs.grf_sheet(ewn=16, nsn=14, nz=15, dx=2000.0, corr=8000.0, seed=5),
        # ice reaching the edge of an outflow domain
        "outflow": configs.grf_sheet(ewn=24, nsn=20, nz=7, dx=2000.0, corr=8000.0, seed=3, periodic=False),
    }


@pytest.fixture(scope="module")
def gpu():
    from cism_amd import gpu as G
    G.lib()
    return G


@pytest.mark.parametrize("name", list(cases()))
@pytest.mark.parametrize("chrongear", [1, 0])
def test_symmetric_copy_gives_the_same_solve(gpu, name, chrongear):
    p = cases()[name]
    print(f"\n[case] {name} ewn={p.ewn} nsn={p.nsn} nz={p.nz} chrongear={chrongear}", flush=True)
    lu, lv = O.loads(p)
    out = {}
    for on in (1, 0):
        with gpu.Context.for_problem(p) as ctx:
            ctx.set_fp64_symmetric(on)
            out[on] = ctx.velo_higher_solve_3d(p, lu, lv, whichsparse=3 if chrongear else 2)
    a, b = out[1], out[0]
    print(f"  picard {a['picard_iters']} / {b['picard_iters']}, linear {a['linear_iters']} / {b['linear_iters']}")
    assert a["picard_iters"] == b["picard_iters"] and a["linear_iters"] == b["linear_iters"]
    assert a["final_L2_norm"] == b["final_L2_norm"]
    for f in ("uvel", "vvel", "efvs"):
        assert np.array_equal(a[f], b[f]), f


def test_switch_rejects_other_values(gpu):
    p = configs.dome()
    with gpu.Context.for_problem(p) as ctx:
        with pytest.raises(RuntimeError):
            ctx.set_fp64_symmetric(2)


def test_fp64_matvec_timing_modes(gpu):
    """which 0 (the storage the FP64 solver uses), 5 (full) and 6 (symmetric, other launch
    order) all run on an assembled matrix"""
    p = configs.ice_sheet(ewn=36, nsn=42, nz=6, dx=33333.0, seed=7)
    lu, lv = O.loads(p)
    with gpu.Context.for_problem(p) as ctx:
        ctx.velo_higher_solve_3d(p, lu, lv, maxiter_nonlinear=1)
        for which in (0, 5, 6):
            assert ctx.time_kernel(which, 2) > 0.0
