"""Learned initial guess for the Blatter-Pattyn velocity solve (no reference counterpart).

A fully convolutional 4-level encoder-decoder (about 1.0M parameters) maps five
2D fields on the owned vertices of a periodic grid

    0  stagthck / 1000 m
    1  100 * d(usrf)/dx
    2  100 * d(usrf)/dy
    3  (ln(max(beta, 1e-3)) - 7) / 2        beta in Pa yr/m
    4  log10(column-mean A) + 16           A = flwa in Pa^-3 yr^-1

to u and v at each of nz <= 11 levels (channels 0..10 and 11..21, in units of
`uscale` m/yr) and two per-column log preconditioner scales (channels 22, 23).
The network runs in libcism_gpu.so (MFMA BF16 convolutions, FP32 accumulation,
hand-written forward/backward and Adam); this module holds the host side:
feature and target extraction, weight files and the prediction of a full
velocity field in the reference's (nz, nx-1, ny-1) layout.

Inactive vertices get zero features and zero predicted velocity; the solver's
own Dirichlet and halo rules apply on top of the guess.
"""
from __future__ import annotations

import ctypes
import json
import os

import numpy as np

from . import configs
from .gpu import lib, CismError

N_IN = 5
N_OUT = 24
NZ_MAX = 11
U_CH = 0
V_CH = 11
WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")

_c = ctypes
_P = _c.c_void_p
_fp = _c.POINTER(_c.c_float)
_dp = _c.POINTER(_c.c_double)


def _setup(L):
    if getattr(L, "_sg_ready", False):
        return L
    L.cism_surrogate_create.argtypes = [_c.c_int, _c.c_int, _c.c_int, _c.c_int, _c.POINTER(_P)]
    L.cism_surrogate_destroy.argtypes = [_P]
    L.cism_surrogate_num_params.argtypes = [_P]
    L.cism_surrogate_num_params.restype = _c.c_longlong
    L.cism_surrogate_get_params.argtypes = [_P, _fp]
    L.cism_surrogate_set_params.argtypes = [_P, _fp]
    L.cism_surrogate_forward.argtypes = [_P, _fp, _fp]
    L.cism_surrogate_loss_grad.argtypes = [_P, _fp, _fp, _fp, _c.c_int, _dp, _fp]
    L.cism_surrogate_train_step.argtypes = [_P, _fp, _fp, _fp, _c.c_int, _c.c_float, _dp]
    L.cism_surrogate_set_channel_weights.argtypes = [_P, _fp]
    L.cism_surrogate_last_forward_ms.argtypes = [_P]
    L.cism_surrogate_last_forward_ms.restype = _c.c_double
    L._sg_ready = True
    return L


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_fp)


def _check(L, st, what):
    if st != 0:
        raise CismError(st, f"{what}: {L.cism_last_error().decode()}")


class Surrogate:
    """The network on one GPU, for batches of `batch` grids of H x W owned vertices."""

    def __init__(self, H: int, W: int, batch: int = 1, device: int = 0):
        self.L = _setup(lib())
        self.H, self.W, self.B = H, W, batch
        h = _P()
        _check(self.L, self.L.cism_surrogate_create(H, W, batch, device, _c.byref(h)), "surrogate_create")
        self.h = h
        self.nparams = int(self.L.cism_surrogate_num_params(h))

    def close(self):
        if self.h:
            self.L.cism_surrogate_destroy(self.h)
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

    def _shape_in(self, x):
        x = np.asarray(x, dtype=np.float32)
        if x.shape != (self.B, self.H, self.W, N_IN):
            raise ValueError(f"input shape {x.shape} != {(self.B, self.H, self.W, N_IN)}")
        return x

    def get_params(self) -> np.ndarray:
        w = np.empty(self.nparams, dtype=np.float32)
        _check(self.L, self.L.cism_surrogate_get_params(self.h, w.ctypes.data_as(_fp)), "get_params")
        return w

    def set_params(self, w: np.ndarray):
        w, pw = _f32(w)
        if w.size != self.nparams:
            raise ValueError(f"{w.size} parameters given, network has {self.nparams}")
        _check(self.L, self.L.cism_surrogate_set_params(self.h, pw), "set_params")

    def set_channel_weights(self, w):
        w, pw = _f32(w)
        assert w.size == N_OUT
        _check(self.L, self.L.cism_surrogate_set_channel_weights(self.h, pw), "set_channel_weights")

    def forward(self, x) -> np.ndarray:
        x, px = _f32(self._shape_in(x))
        out = np.empty((self.B, self.H, self.W, N_OUT), dtype=np.float32)
        _check(self.L, self.L.cism_surrogate_forward(self.h, px, out.ctypes.data_as(_fp)), "forward")
        return out

    def last_forward_ms(self) -> float:
        """device time of the last forward pass (kernels only)"""
        return float(self.L.cism_surrogate_last_forward_ms(self.h))

    def loss_grad(self, x, tgt, mask, nch=N_OUT):
        x, px = _f32(self._shape_in(x))
        t, pt = _f32(tgt)
        m, pm = _f32(mask)
        assert t.shape == (self.B, self.H, self.W, N_OUT) and m.shape == (self.B, self.H, self.W)
        loss = _c.c_double()
        g = np.empty(self.nparams, dtype=np.float32)
        _check(self.L, self.L.cism_surrogate_loss_grad(self.h, px, pt, pm, nch, _c.byref(loss),
                                                       g.ctypes.data_as(_fp)), "loss_grad")
        return loss.value, g

    def train_step(self, x, tgt, mask, lr, nch=N_OUT) -> float:
        x, px = _f32(self._shape_in(x))
        t, pt = _f32(tgt)
        m, pm = _f32(mask)
        loss = _c.c_double()
        _check(self.L, self.L.cism_surrogate_train_step(self.h, px, pt, pm, nch, lr, _c.byref(loss)),
               "train_step")
        return loss.value


# ---------------------------------------------------------------------------
# fields <-> network tensors (owned vertices of a periodic grid, pixel (y, x) is
# vertex (i, j) = (x + 2, y + 2) in 0-based staggered indices)
# ---------------------------------------------------------------------------
def owned(a2d: np.ndarray, p: configs.Problem) -> np.ndarray:
    """(nx-1, ny-1) vertex field -> (nsn, ewn) image."""
    return np.asarray(a2d)[2:p.nx - 2, 2:p.ny - 2].T


def features(p: configs.Problem) -> np.ndarray:
    """(H, W, 5) float32 input of one problem."""
    av = owned(p.active_vertex, p) != 0
    f = np.zeros(av.shape + (N_IN,), dtype=np.float32)
    flwa = np.asarray(p.flwa)                                  # (nz-1, nx, ny) cells
    colA = flwa.mean(axis=0)
    # vertex (i, j) (0-based) is the NE corner of cell (i, j): mean of cells i..i+1, j..j+1
    vA = 0.25 * (colA[:-1, :-1] + colA[1:, :-1] + colA[:-1, 1:] + colA[1:, 1:])
    f[..., 0] = owned(p.stagthck, p) / 1000.0
    f[..., 1] = owned(p.dusrf_dx, p) * 100.0
    f[..., 2] = owned(p.dusrf_dy, p) * 100.0
    f[..., 3] = (np.log(np.maximum(owned(p.beta_internal, p), 1e-3)) - 7.0) / 2.0
    f[..., 4] = np.log10(np.maximum(owned(vA, p), 1e-30)) + 16.0
    f[~av] = 0.0
    return f


S_CH = 22          # log preconditioner scale, u and v components (channels 22, 23)


def encode(vel, uscale: float, target: str = "linear"):
    """velocity (m/yr) -> network units: vel / uscale ("linear") or
    sign(vel) log(1 + |vel| / uscale) ("log1p", which keeps the few fast margin
    vertices of an ice sheet from dominating the squared loss)"""
    if target == "log1p":
        return np.sign(vel) * np.log1p(np.abs(vel) / uscale)
    if target == "linear":
        return vel / uscale
    raise ValueError(f"unknown target encoding {target!r}")


# largest |output| decoded (network units): a diverging output is capped at about
# 2e6 m/yr (log1p, uscale 100) or 1e6 m/yr (linear, uscale 1000) rather than reaching inf
OUT_CLIP = {"log1p": 10.0, "linear": 1000.0}


def decode(out, uscale: float, target: str = "linear"):
    """network units -> velocity (m/yr), the inverse of encode (outputs clipped to
    +-OUT_CLIP[target] first)"""
    if target not in OUT_CLIP:
        raise ValueError(f"unknown target encoding {target!r}")
    out = np.clip(out, -OUT_CLIP[target], OUT_CLIP[target])
    if target == "log1p":
        return np.sign(out) * uscale * np.expm1(np.abs(out))
    return uscale * out


def targets(p: configs.Problem, uvel, vvel, uscale: float, su=None, sv=None, target: str = "linear"):
    """(H, W, 24) float32 target and (H, W) mask from a solved velocity field and,
    optionally, the per-column preconditioner scales of that solve (vertex fields)."""
    if p.nz > NZ_MAX:
        raise ValueError(f"nz = {p.nz} > {NZ_MAX}")
    av = owned(p.active_vertex, p) != 0
    t = np.zeros(av.shape + (N_OUT,), dtype=np.float32)
    u = np.asarray(uvel)[:, 2:p.nx - 2, 2:p.ny - 2].transpose(2, 1, 0)   # (H, W, nz)
    v = np.asarray(vvel)[:, 2:p.nx - 2, 2:p.ny - 2].transpose(2, 1, 0)
    t[..., U_CH:U_CH + p.nz] = encode(u, uscale, target)
    t[..., V_CH:V_CH + p.nz] = encode(v, uscale, target)
    if su is not None and sv is not None:
        t[..., S_CH] = np.log(owned(su, p))
        t[..., S_CH + 1] = np.log(owned(sv, p))
    t[~av] = 0.0
    return t, av.astype(np.float32)


def velocity_from_output(p: configs.Problem, out: np.ndarray, uscale: float, target: str = "linear"):
    """Network output (H, W, 24) -> uvel, vvel (nz, nx-1, ny-1), halos by periodic copy."""
    av = owned(p.active_vertex, p) != 0
    u = np.zeros((p.nz, p.nx - 1, p.ny - 1), order="F")
    v = np.zeros((p.nz, p.nx - 1, p.ny - 1), order="F")
    o = np.where(av[..., None], out, 0.0)
    u[:, 2:p.nx - 2, 2:p.ny - 2] = decode(o[..., U_CH:U_CH + p.nz], uscale, target).transpose(2, 1, 0)
    v[:, 2:p.nx - 2, 2:p.ny - 2] = decode(o[..., V_CH:V_CH + p.nz], uscale, target).transpose(2, 1, 0)
    for a in (u, v):
        for k in range(p.nz):
            a[k] = configs.stag_halo(np.asfortranarray(a[k]), p.periodic)
    return u, v


def scales_from_output(p: configs.Problem, out: np.ndarray):
    """Network output (H, W, 24) -> per-column preconditioner scales (vertex fields), 1 at
    inactive vertices, clamped to [1/16, 16]."""
    av = owned(p.active_vertex, p) != 0
    res = []
    for c in (S_CH, S_CH + 1):
        s = np.ones((p.nx - 1, p.ny - 1), order="F")
        s[2:p.nx - 2, 2:p.ny - 2] = np.where(av, np.exp(np.clip(out[..., c], -2.77, 2.77)), 1.0).T
        res.append(configs.stag_halo(s, p.periodic))
    return res


# ---------------------------------------------------------------------------
# weight files: float32 parameter vector + JSON metadata, in one .npz (no pickle)
# ---------------------------------------------------------------------------
def save_weights(path: str, params: np.ndarray, meta: dict):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez(path, params=np.asarray(params, dtype=np.float32), meta=np.array(json.dumps(meta)))


def load_weights(path: str):
    with np.load(path, allow_pickle=False) as z:
        return z["params"].astype(np.float32), json.loads(str(z["meta"]))


def export_weights_bin(npz_path: str, bin_path: str):
    """Writes a weight file for cism_learned_guess_load_file (e.g. for a Fortran host):
    b"CISMSG01", int32 nz, int32 encoding (0 linear, 1 log1p), float64 uscale, int64
    nparams, then the float32 parameters, little endian."""
    params, meta = load_weights(npz_path)
    enc = {"linear": 0, "log1p": 1}[meta.get("target", "linear")]
    with open(bin_path, "wb") as f:
        f.write(b"CISMSG01")
        f.write(np.array([int(meta.get("nz", NZ_MAX)), enc], dtype="<i4").tobytes())
        f.write(np.array([float(meta["uscale"])], dtype="<f8").tobytes())
        f.write(np.array([params.size], dtype="<i8").tobytes())
        f.write(np.asarray(params, dtype="<f4").tobytes())


def default_weights(family: str) -> str:
    return os.path.join(WEIGHTS_DIR, f"surrogate_{family}.npz")


class Predictor:
    """Holds a trained network for one grid shape and predicts initial guesses."""

    def __init__(self, p: configs.Problem, weights: str, device: int = 0):
        if not p.periodic:
            raise ValueError("the surrogate runs on periodic grids (circular padding)")
        params, self.meta = load_weights(weights)
        nz_w = int(self.meta.get("nz", NZ_MAX))
        if p.nz != nz_w:
            # the output channels hold nz_w levels of u, then nz_w of v: another nz would
            # read the wrong channels
            raise ValueError(f"the weights predict {nz_w} levels, the grid has nz = {p.nz}")
        self.uscale = float(self.meta["uscale"])
        self.target = self.meta.get("target", "linear")
        self.net = Surrogate(p.nsn, p.ewn, 1, device)
        self.net.set_params(params)

    def predict(self, p: configs.Problem, scales=False):
        """Initial guess (uvel, vvel); with scales=True also the per-column preconditioner
        scales (su, sv), when the weights were trained on them."""
        out = self.net.forward(features(p)[None])[0]
        u, v = velocity_from_output(p, out, self.uscale, self.target)
        if not scales:
            return u, v
        if not self.meta.get("scales", False):
            raise ValueError("these weights were not trained on preconditioner scales")
        su, sv = scales_from_output(p, out)
        return u, v, su, sv

    def close(self):
        self.net.close()
