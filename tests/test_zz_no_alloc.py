"""cism_compute_residual_vector_3d_noalloc (device arrays, no device allocation) against the
context entry point cism_compute_residual_vector_3d on the same systems: the residual
vectors bit for bit, the two norms to 1e-13 relative (the partial sums are added in
another fixed order).  Two calls in a row, the second on a smaller grid, so the second
reuses the scratch the first left behind.  The test's own device buffers come from the
HIP runtime the library links (libamdhip64.so.7), through ctypes."""
import ctypes as C

import numpy as np
import pytest

from cism_amd import configs, gpu as G
from helpers import assembled_system

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2   # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class Hip:
    """the few HIP runtime calls the test needs for its own buffers"""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipSetDevice.argtypes = [C.c_int]
        self.bufs = []
        assert self.L.hipSetDevice(0) == 0

    def put(self, a):
        a = np.ascontiguousarray(np.asarray(a).ravel(order="F"))
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), a.nbytes) == 0
        self.bufs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def get(self, p, n, dtype=np.float64):
        out = np.empty(n, dtype=dtype)
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _noalloc(hip, p, A, bu, bv, u, v):
    L = G.lib()
    f = L.cism_compute_residual_vector_3d_noalloc
    f.argtypes = [C.c_int] * 5 + [C.c_void_p] * 13
    ins = [hip.put(p.active_vertex.astype(np.int32))] + [hip.put(a) for a in A] + \
          [hip.put(a) for a in (bu, bv, u, v)]
    # outputs start as NaN: every entry must be written
    nan = np.full(p.nnodes, np.nan)
    ru, rv, norms = hip.put(nan), hip.put(nan), hip.put(np.full(2, np.nan))
    st = f(p.nx, p.ny, p.nz, int(p.periodic), 0, None, *ins, ru, rv, norms)
    assert st == 0, L.cism_last_error().decode()
    shape = (p.nz, p.nx - 1, p.ny - 1)
    out = (hip.get(ru, p.nnodes).reshape(shape, order="F"), hip.get(rv, p.nnodes).reshape(shape, order="F"),
           *hip.get(norms, 2))
    hip.free()
    return out


def test_noalloc_residual_matches_the_context_entry_point():
    big = configs.grf_sheet(ewn=24, nsn=20, nz=7, dx=2000.0, corr=8000.0, seed=3, periodic=False)
    small = configs.ismip_hom_a(ewn=9, nsn=9, nz=5, L=7200.0)
    assert small.nnodes < big.nnodes
    hip = Hip()
    for p in (big, small):
        print(f"[noalloc] {p.name} {p.ewn}x{p.nsn}x{p.nz}", flush=True)
        A, bu, bv, u, v, _ = assembled_system(p, seed=5)
        ru, rv = np.zeros(u.shape, order="F"), np.zeros(u.shape, order="F")
        with G.Context.for_problem(p) as ctx:
            L2, L2r = ctx.compute_residual_vector_3d(p.active_vertex, *A, bu, bv, u, v, ru, rv)
        nu, nv, nL2, nL2r = _noalloc(hip, p, A, bu, bv, u, v)
        assert np.array_equal(nu, ru) and np.array_equal(nv, rv)
        assert abs(nL2 - L2) <= 1e-13 * L2 and abs(nL2r - L2r) <= 1e-13 * L2r
