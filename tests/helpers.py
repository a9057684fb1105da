"""Shared helpers of the parity tests (oracle on the CPU, HIP library on the GPU)."""
import numpy as np

import oracle as O
from cism_amd import configs


def rel_max(a, b):
    """max |a-b| / max |b| (0 when both are zero)."""
    d = np.max(np.abs(a - b)) if a.size else 0.0
    s = np.max(np.abs(b)) if b.size else 0.0
    return 0.0 if d == 0.0 else d / (s if s > 0 else 1.0)


def rel_l2(a, b):
    n = np.linalg.norm(b.ravel())
    return np.linalg.norm((a - b).ravel()) / (n if n > 0 else 1.0)


def assembled_system(p, seed=None, velocity_scale=100.0):
    """Oracle A, b after assembly, basal BC, Dirichlet, halo and symmetry: the system the
    reference hands to its PCG (velo_higher:3011-3113)."""
    if seed is not None:
        rng = np.random.default_rng(seed)
        u = np.asfortranarray(velocity_scale * rng.standard_normal(p.uvel.shape))
        v = np.asfortranarray(velocity_scale * rng.standard_normal(p.vvel.shape))
        O.halo(p, u)
        O.halo(p, v)
    else:
        u, v = p.uvel.copy(order="F"), p.vvel.copy(order="F")
    st, A, efvs = O.assemble(p, u, v)
    assert st == 0
    if p.whichbabc != configs.HO_BABC_NO_SLIP:
        O.basal_local(p, A)
    lu, lv = O.load_gravity(p)
    bu, bv = lu.copy(order="F"), lv.copy(order="F")
    O.dirichlet(p, A, bu, bv, u, v)
    for a in A:
        O.halo(p, a)
    O.halo(p, bu)
    O.halo(p, bv)
    O.check_symmetry(p, A)
    return A, bu, bv, u, v, efvs


def node_index(p):
    """Dense numbering of the owned active nodes: returns (ids array (nz, nx-1, ny-1) with -1
    elsewhere, count)."""
    ids = -np.ones((p.nz, p.nx - 1, p.ny - 1), dtype=np.int64, order="F")
    lo = 2 if p.periodic else 1
    n = 0
    for j in range(lo, p.ny - 2):
        for i in range(lo, p.nx - 2):
            if p.active_vertex[i, j]:
                for k in range(p.nz):
                    ids[k, i, j] = n
                    n += 1
    return ids, n


def to_csr(p, A, active_vertex=None):
    """The 4 blocks as a (2N x 2N) scipy CSR over all nodes of the staggered grid (u block
    first), rows restricted to owned active vertices, as the matvec uses them."""
    import scipy.sparse as sp
    av = p.active_vertex if active_vertex is None else active_vertex
    nz, sx, sy = p.nz, p.nx - 1, p.ny - 1
    N = nz * sx * sy
    lo = 2 if p.periodic else 1
    rows, cols, vals = [], [], []
    for j in range(lo, p.ny - 2):
        for i in range(lo, p.nx - 2):
            if not av[i, j]:
                continue
            for k in range(nz):
                r = k + nz * (i + sx * j)
                for kA in (-1, 0, 1):
                    if not 0 <= k + kA < nz:
                        continue
                    for jA in (-1, 0, 1):
                        for iA in (-1, 0, 1):
                            m = (iA + 1) + 3 * (jA + 1) + 9 * (kA + 1)
                            cidx = (k + kA) + nz * ((i + iA) + sx * (j + jA))
                            for bi, (ro, co) in enumerate(((0, 0), (0, N), (N, 0), (N, N))):
                                v = A[bi][m, k, i, j]
                                if v != 0.0:
                                    rows.append(r + ro)
                                    cols.append(cidx + co)
                                    vals.append(v)
    return sp.csr_matrix((vals, (rows, cols)), shape=(2 * N, 2 * N))


def flat2(u, v):
    return np.concatenate([u.ravel(order="F"), v.ravel(order="F")])


def periodic_fold(p, M):
    """Folds the halo columns of M (rows: owned nodes) onto their periodic images, giving
    the square operator on owned nodes that the periodic solve represents.  Returns
    (Mf, owned_index) with owned_index the full-grid indices (u block then v)."""
    import scipy.sparse as sp
    nz, sx, sy = p.nz, p.nx - 1, p.ny - 1
    N = nz * sx * sy
    img = lambda i, s: i + s - 3 if i < 2 else (2 if i == s - 1 else i)   # noqa: E731
    owned = []
    for j in range(2, sy - 1):
        for i in range(2, sx - 1):
            for k in range(nz):
                owned.append(k + nz * (i + sx * j))
    owned = np.array(owned)
    pos = -np.ones(N, dtype=np.int64)
    pos[owned] = np.arange(len(owned))
    col_map = np.empty(N, dtype=np.int64)
    for j in range(sy):
        for i in range(sx):
            for k in range(nz):
                col_map[k + nz * (i + sx * j)] = pos[k + nz * (img(i, sx) + sx * img(j, sy))]
    no = len(owned)
    cm = np.concatenate([col_map, col_map + no])
    P = sp.csr_matrix((np.ones(2 * N), (np.arange(2 * N), cm)), shape=(2 * N, 2 * no))
    oi = np.concatenate([owned, owned + N])
    return (M[oi] @ P).tocsr(), oi
