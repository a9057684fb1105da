"""Trains the learned initial guess on velocity fields solved by the plain FP64 GPU path.

    python -m cism_amd.train_surrogate --family grf --n 256 --samples 160 --epochs 40 \
        --out cism_amd/weights/surrogate_grf.npz

Data: `samples` seeded problems of the chosen family (grf: Gaussian-random-field
thickness on a flat bed with lognormal beta, seeds 1000..; sheet: the ice-sheet
generator with a floating fringe), each solved on the GPU by the plain FP64
Picard + Chron-Gear path (the reference algorithm at its default tolerances).
Training: Adam on the masked MSE of u, v at every level (in units of uscale
m/yr), with the eight exact symmetries of the square grid as augmentation
(x-flip negates u and d(usrf)/dx, y-flip negates v and d(usrf)/dy, the
transpose swaps them).  Validation: relative L2 velocity error of the
prediction on held-out seeds.  Everything runs on the GPU through
libcism_gpu.so; the host only moves batches.
"""
from __future__ import annotations

import argparse
import math
import sys
import time

import numpy as np

from . import configs
from . import gpu as G
from . import surrogate as SG


def make_problem(family: str, ewn: int, nsn: int, seed: int):
    if family == "grf":
        return configs.grf_sheet(ewn=ewn, nsn=nsn, nz=11, dx=1000.0, seed=seed)
    if family == "sheet":
        return configs.ice_sheet(ewn=ewn, nsn=nsn, nz=11, dx=1000.0, seed=seed)
    raise SystemExit(f"unknown family {family}")


def solve_guessed(ctx, p, params, meta):
    """The fast path from the geometry with a learned initial guess (an earlier network):
    the same Picard fixed point under the reference's stopping rule, reached in fewer
    iterations than from zero."""
    if not getattr(ctx, "_guess_ready", False):
        ctx.learned_guess_load(params, int(meta.get("nz", 11)), 1 if meta.get("target") == "log1p" else 0,
                               float(meta["uscale"]))
        ctx._guess_ready = True
    bp = G.BasalPhysics(p.whichbabc)
    return ctx.glissade_velo_higher_solve(p, bp=bp, beta_external=p.extra["beta_external"], initial_guess=1,
                                          precision=1, adaptive_linear=1, inner_rtol=5e-2, anderson_m=3,
                                          efvs=False)


def solve_plain(ctx, p, log, chunk=10, precision=0):
    """The plain FP64 path at the reference's default tolerances (precision 0), or the
    mixed-precision path with inexact linear solves and Anderson acceleration to the same
    Picard stopping rule (precision 1), run in chunks of Picard iterations (a Picard
    iteration depends only on the current velocity, so a plain chunk restarts exactly) so
    that long solves report progress."""
    lu, lv = ctx.loads(p)
    u, v, done = p.uvel, p.vvel, 0
    fast = dict(adaptive_linear=1, anderson_m=4) if precision else {}
    while True:
        r = ctx.velo_higher_solve_3d(p, lu, lv, uvel=u, vvel=v, precision=precision,
                                     maxiter_nonlinear=chunk, **fast)
        u, v = r["uvel"], r["vvel"]
        done += r["picard_iters"]
        if r["picard_iters"] < chunk or done >= p.maxiter_nonlinear:
            r["picard_iters"] = done
            return r
        log(f"  {done} Picard, L2 {r['final_L2_norm']:.2e}")


def crop(x, t, m, c, margin, rng):
    """Random c x c window; the mask drops `margin` pixels at its edges, where the window's
    circular padding does not match the field outside it."""
    H, W = m.shape
    if not c or (c >= H and c >= W):
        return x, t, m
    y0, x0 = int(rng.integers(0, H - c + 1)), int(rng.integers(0, W - c + 1))
    sl = (slice(y0, y0 + c), slice(x0, x0 + c))
    mm = m[sl].copy()
    mm[:margin] = mm[-margin:] = 0
    mm[:, :margin] = mm[:, -margin:] = 0
    return x[sl], t[sl], mm


def augment(x, t, m, k):
    """Symmetry k in 0..7 of (H, W, C) inputs / targets and the (H, W) mask: bit 0 x-flip,
    bit 1 y-flip, bit 2 transpose."""
    x, t, m = x.copy(), t.copy(), m.copy()
    nz = SG.V_CH - SG.U_CH
    if k & 1:
        x, t, m = x[:, ::-1].copy(), t[:, ::-1].copy(), m[:, ::-1]
        x[..., 1] *= -1
        t[..., SG.U_CH:SG.U_CH + nz] *= -1
    if k & 2:
        x, t, m = x[::-1].copy(), t[::-1].copy(), m[::-1]
        x[..., 2] *= -1
        t[..., SG.V_CH:SG.V_CH + nz] *= -1
    if k & 4:
        x, t, m = x.transpose(1, 0, 2).copy(), t.transpose(1, 0, 2).copy(), m.T
        x[..., [1, 2]] = x[..., [2, 1]]
        u = t[..., SG.U_CH:SG.U_CH + nz].copy()
        t[..., SG.U_CH:SG.U_CH + nz] = t[..., SG.V_CH:SG.V_CH + nz]
        t[..., SG.V_CH:SG.V_CH + nz] = u
        t[..., [SG.S_CH, SG.S_CH + 1]] = t[..., [SG.S_CH + 1, SG.S_CH]]
    return np.ascontiguousarray(x), np.ascontiguousarray(t), np.ascontiguousarray(m)


def rel_err(pred, tgt, mask, nch):
    e = (pred[..., :nch] - tgt[..., :nch]) * mask[..., None]
    return float(np.linalg.norm(e) / max(np.linalg.norm(tgt[..., :nch] * mask[..., None]), 1e-300))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--family", default="grf")
    ap.add_argument("--n", type=int, default=256, help="square grid size")
    ap.add_argument("--ewn", type=int, default=0, help="grid size in x (overrides --n)")
    ap.add_argument("--nsn", type=int, default=0, help="grid size in y (overrides --n)")
    ap.add_argument("--crop", type=int, default=0, help="train on random crops of this size")
    ap.add_argument("--margin", type=int, default=16)
    ap.add_argument("--samples", type=int, default=160)
    ap.add_argument("--val", type=int, default=8)
    ap.add_argument("--seed0", type=int, default=1000)
    ap.add_argument("--val-seed0", type=int, default=3000)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--uscale", type=float, default=100.0)
    ap.add_argument("--out", required=True)
    ap.add_argument("--init", default=None, help="start from these weights")
    ap.add_argument("--candidates", type=int, default=3,
                    help="try up to this many seeds per wanted sample (skipping unconverged ones)")
    ap.add_argument("--target", default="linear", choices=["linear", "log1p"],
                    help="velocity encoding of the network output (log1p compresses fast margins)")
    ap.add_argument("--data-precision", type=int, default=0,
                    help="0: training solves on the plain FP64 path; 1: on the mixed-precision path")
    ap.add_argument("--scales", type=int, default=0,
                    help="also learn the per-column preconditioner scales (channels 22, 23)")
    ap.add_argument("--checkpoint", type=int, default=0, help="save the weights every this many epochs")
    ap.add_argument("--data-guess", default=None,
                    help="weights of an earlier network: the training solves start from its guess "
                         "(fast path from the geometry, the reference's stopping rule)")
    a = ap.parse_args(argv)
    t_start = time.perf_counter()
    nch = 24 if a.scales else 22
    ewn, nsn = a.ewn or a.n, a.nsn or a.n
    c = a.crop if a.crop else 0
    if c and (c > min(ewn, nsn) or c % 8):
        raise SystemExit("--crop must be a multiple of 8 no larger than the grid")
    th, tw = (c, c) if c else (nsn, ewn)

    guess_w = SG.load_weights(a.data_guess) if a.data_guess else None

    def gen(seed0, want, tag):
        """solves seeds seed0, seed0+1, ... and keeps the first `want` whose Picard loop
        converged (a solve that reaches the iteration cap is not a solution to learn);
        the samples go into arrays allocated once, at the first sample's shapes"""
        X = T = M = None
        used = []
        ctx = None
        s = seed0
        while len(used) < want and s < seed0 + want * a.candidates:
            p = make_problem(a.family, ewn, nsn, s)
            if ctx is None:
                ctx = G.Context.for_problem(p)
            t0 = time.perf_counter()
            if guess_w is not None:
                r = solve_guessed(ctx, p, *guess_w)
            else:
                r = solve_plain(ctx, p, lambda m: print(f"[train] {tag} seed {s}: {m}", flush=True),
                                precision=a.data_precision)
            umax = float(np.abs(r["uvel"]).max())
            ok = r["picard_iters"] < p.maxiter_nonlinear
            print(f"[train] {tag} {len(used) + ok}/{want} seed {s}: {r['picard_iters']} Picard, "
                  f"{time.perf_counter() - t0:.2f} s, max|u| {umax:.0f} m/yr"
                  + ("" if ok else " -- not converged, skipped"), flush=True)
            if ok:
                x = SG.features(p)
                su, sv = ctx.column_scale_target() if a.scales else (None, None)
                t, m = SG.targets(p, r["uvel"], r["vvel"], a.uscale, su, sv, target=a.target)
                if X is None:
                    X = np.empty((want,) + x.shape, x.dtype)
                    T = np.empty((want,) + t.shape, t.dtype)
                    M = np.empty((want,) + m.shape, m.dtype)
                n = len(used)
                X[n], T[n], M[n] = x, t, m
                used.append(s)
            s += 1
        ctx.close()
        if not used:
            raise SystemExit(f"no converged {tag} problem among the candidate seeds")
        n = len(used)
        return X[:n], T[:n], M[:n], used

    Xtr, Ttr, Mtr, seeds_tr = gen(a.seed0, a.samples, "data")
    Xva, Tva, Mva, seeds_va = gen(a.val_seed0, a.val, "val")
    print(f"[train] data ready in {time.perf_counter() - t_start:.0f} s", flush=True)

    net = SG.Surrogate(th, tw, a.batch)
    if a.init:
        w0, _ = SG.load_weights(a.init)
        net.set_params(w0)
    evalnet = SG.Surrogate(nsn, ewn, 1)

    def validate():
        """mean relative L2 error of the predicted velocities (decoded to m/yr) on the
        validation problems"""
        evalnet.set_params(net.get_params())
        dec = lambda t: SG.decode(t[..., :22], a.uscale, a.target)   # noqa: E731
        errs = [rel_err(dec(evalnet.forward(Xva[i:i + 1])[0]), dec(Tva[i]), Mva[i], 22)
                for i in range(len(Xva))]
        return float(np.mean(errs))

    def meta_base():
        return {"scales": bool(a.scales), "family": a.family, "ewn": ewn, "nsn": nsn, "crop": c, "nz": 11,
                "uscale": a.uscale, "samples": len(seeds_tr), "seeds": seeds_tr, "val_seeds": seeds_va,
                "target": a.target, "batch": a.batch, "lr": a.lr}

    rng = np.random.default_rng(0)
    n = len(Xtr)
    # one epoch sees each sample once (whole grids) or as many crops as would tile it
    per_sample = max(1, (nsn // th) * (ewn // tw))
    steps_per_epoch = max(1, n * per_sample // a.batch)
    total = a.epochs * steps_per_epoch
    step = 0
    hist = []
    e0 = validate()
    print(f"[train] initial validation rel L2 {e0:.3f}", flush=True)
    for ep in range(a.epochs):
        perm = rng.permutation(n)
        losses = []
        t0 = time.perf_counter()
        for b in range(steps_per_epoch):
            idx = [perm[(b * a.batch + q) % n] for q in range(a.batch)]
            xs, ts, ms = zip(*(augment(*crop(Xtr[i], Ttr[i], Mtr[i], c, a.margin, rng),
                                       int(rng.integers(0, 8))) for i in idx))
            lr = a.lr * (0.03 + 0.97 * 0.5 * (1 + math.cos(math.pi * step / total)))
            losses.append(net.train_step(np.stack(xs), np.stack(ts), np.stack(ms), lr, nch=nch))
            step += 1
        ve = validate()
        hist.append((ep, float(np.mean(losses)), ve))
        print(f"[train] epoch {ep + 1}/{a.epochs}: loss {np.mean(losses):.4f}, val rel L2 {ve:.4f}, "
              f"lr {lr:.2e}, {time.perf_counter() - t0:.1f} s", flush=True)
        if a.checkpoint and (ep + 1) % a.checkpoint == 0:
            SG.save_weights(a.out, net.get_params(), dict(meta_base(), epochs_done=ep + 1, history=hist))
    meta = {"scales": bool(a.scales), "family": a.family, "ewn": ewn, "nsn": nsn, "crop": c, "nz": 11, "uscale": a.uscale, "samples": len(seeds_tr),
            "seeds": seeds_tr, "val_seeds": seeds_va, "data_precision": a.data_precision,
            "data_guess": a.data_guess, "init": a.init,
            "target": a.target,
            "epochs": a.epochs, "batch": a.batch, "lr": a.lr, "val_rel_l2": hist[-1][2],
            "history": hist, "seconds": time.perf_counter() - t_start}
    SG.save_weights(a.out, net.get_params(), meta)
    print(f"[train] saved {a.out}: val rel L2 {hist[-1][2]:.4f} after {a.epochs} epochs, "
          f"{time.perf_counter() - t_start:.0f} s", flush=True)
    net.close()
    evalnet.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
