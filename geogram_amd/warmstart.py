"""Learned warm start of the semi-discrete optimal-transport solver (no
counterpart in geogram; it only changes the start of the damped Newton
iterations of ``OptimalTransportMap3d.optimize``).

A small MLP (6 -> 64 -> 64 -> 1, leaky ReLU, FP32; kernels in csrc/mlp.h)
maps a seed's features -- position in the box, log of its target mass over
the mean mass, log of the mean and max distance to its 16 nearest seeds over
the mean spacing h = (V/n)^(1/3) -- to its Laguerre weight w_i - mean(w)
(box units).  The weight field of a converged
transport plan is mostly a smooth potential that does not depend on n; the
part that follows mass fluctuations of single seeds scales with h^2 and is
left to Newton.

Training data are (features, converged weight) pairs that the solver itself
produces on Gaussian-mixture instances; the network is trained on the GPU by
``vgc_mlp_grad`` / ``vgc_mlp_adam``.  The solver keeps the plain zero start as
fallback: a predicted start that leaves a cell empty, or whose mass residual
is worse than the zero start's, is dropped (``vgc_ot_solve`` with
``VGC_OT_JUDGE_START``).

    python -m geogram_amd.warmstart generate --first 0 --count 100 --out part0.npz
    python -m geogram_amd.warmstart generate --first 100 --count 108 --out part1.npz
    python -m geogram_amd.warmstart train --data part0.npz part1.npz
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import _lib
from ._lib import VGC_MLP_FEATURES, VGC_MLP_PARAMS

_HERE = os.path.dirname(os.path.abspath(__file__))
PARAMS_FILE = os.path.join(_HERE, "warmstart_params.npy")
N_HID = 64
OFF_B1 = VGC_MLP_FEATURES * N_HID
OFF_W2 = OFF_B1 + N_HID
OFF_B2 = OFF_W2 + N_HID * N_HID
OFF_W3 = OFF_B2 + N_HID
OFF_B3 = OFF_W3 + N_HID


def init_params(seed: int = 0) -> np.ndarray:
    """He-initialised weights, zero biases."""
    rng = np.random.default_rng(seed)
    p = np.zeros(VGC_MLP_PARAMS, dtype=np.float32)
    p[:OFF_B1] = rng.normal(0.0, math.sqrt(2.0 / VGC_MLP_FEATURES), OFF_B1)
    p[OFF_W2:OFF_B2] = rng.normal(0.0, math.sqrt(2.0 / N_HID), N_HID * N_HID)
    p[OFF_W3:OFF_B3] = rng.normal(0.0, math.sqrt(1.0 / N_HID), N_HID)
    return p


def fold_normalisation(p, mu, sd, t_mu, t_sd) -> np.ndarray:
    """Parameters of a network trained on (x - mu) / sd -> (t - t_mu) / t_sd,
    rewritten to take raw features and return raw targets."""
    q = np.asarray(p, dtype=np.float64).copy()
    W1 = q[:OFF_B1].reshape(VGC_MLP_FEATURES, N_HID)
    mu = np.asarray(mu, dtype=np.float64)
    sd = np.asarray(sd, dtype=np.float64)
    q[OFF_B1:OFF_W2] -= (mu / sd) @ W1
    q[:OFF_B1] = (W1 / sd[:, None]).ravel()
    q[OFF_W3:OFF_B3] *= t_sd
    q[OFF_B3] = q[OFF_B3] * t_sd + t_mu
    return q.astype(np.float32)


def load_params(path: str | None = None) -> np.ndarray:
    p = np.load(path or PARAMS_FILE, allow_pickle=False)
    if p.dtype != np.float32 or p.shape != (VGC_MLP_PARAMS,):
        raise ValueError(f"{path or PARAMS_FILE}: expected {VGC_MLP_PARAMS} float32 values")
    return p


def predict_weights(ctx, points, nu, box=(0, 0, 0, 1, 1, 1), params=None):
    """Initial Laguerre weights (FP64, same kind of array as points)."""
    if params is None:
        params = load_params()
    feats = ctx.ot_features(points, nu, box)
    if _is_torch_cuda(points):
        import torch
        if isinstance(params, torch.Tensor):
            p = params.to(device=points.device, dtype=torch.float32).contiguous()
        else:
            p = torch.from_numpy(np.asarray(params, dtype=np.float32)).to(points.device)
        return ctx.mlp_forward(p, feats).to(torch.float64)
    if not isinstance(params, np.ndarray):
        params = params.cpu().numpy()
    return ctx.mlp_forward(np.asarray(params, dtype=np.float32), feats).astype(np.float64)


def _pieces(q, w_c, v_l, mg_l, v_v, mg_v, lo, hi, s_factor):
    """Convex pieces (m x 9: a, g, mu, b, R) and soft-max scale s of a coarse
    plan: seeds q, weights w_c, Laguerre and Voronoi volumes / first moments
    of the coarse cells in the box [lo, hi]."""
    if not (v_l > 0).all():
        raise RuntimeError("coarse_start: the coarse plan has an empty cell")
    c = mg_l / v_l[:, None]
    b = mg_v / v_v[:, None]
    mu = np.cbrt(v_l / v_v)
    r_vor = np.cbrt(3.0 * v_v / (4.0 * np.pi))
    r_lag = np.cbrt(3.0 * v_l / (4.0 * np.pi))
    s = s_factor * float(np.median(4.0 * r_vor * r_lag))
    dwall = np.minimum(c - lo, hi - c).min(axis=1)
    R = np.minimum(r_vor, 0.98 * dwall / mu)
    g = 2.0 * c
    r2 = ((q - b) ** 2).sum(axis=1)
    hq = np.where(r2 <= R * R, r2, R * (2.0 * np.sqrt(r2) - R))
    a = (q * q).sum(axis=1) - w_c - (g * q).sum(axis=1) - mu * hq
    return np.ascontiguousarray(np.column_stack([a, g, mu, b, R])), s, mu


def coarse_start(ctx, points, nu, box=(0, 0, 0, 1, 1, 1), ratio=256, s_factor=1.0,
                 coarse_tol=1e-3, min_coarse=64, seed=12345):
    """Newton start from a coarse transport plan, with no empty cell.

    A random subset of n // ratio seeds carries the masses of the fine seeds
    nearest to it and is solved to coarse_tol.  Coarse seed j gives one convex
    piece: the tangent of the coarse potential phi = |q|^2 - w at q_j, with
    slope 2 c_j (c_j the centroid of its Laguerre cell), plus a curvature
    mu_j = (laguerre volume / voronoi volume)^(1/3) about the centroid b_j of
    its Voronoi cell among the coarse seeds, so that the fine seeds around q_j
    spread over c_j's cell; the curvature is quadratic up to the Voronoi
    cell's radius and linear beyond.  The fine start is the soft maximum of
    all pieces (vgc_ot_convex_start), a convex function, so every fine cell
    is non-empty.  s: s_factor x the median over coarse seeds of
    4 x (Voronoi radius) x (Laguerre radius).  Returns (w, info).  Default
    ratio 256: measured on 1M 16-Gaussian seeds with log-normal masses, 10-12
    Newton steps and 2.3-2.6 s to 1e-4 in three trials (s_factor 1, 1.5, 2;
    two coarse subsets), against 10-13 steps and 2.2-4.4 s over three coarse
    subsets at ratio 128."""
    dev = _is_torch_cuda(points)
    n = int(points.shape[0])
    lo = np.asarray(box[:3], dtype=np.float64)
    hi = np.asarray(box[3:], dtype=np.float64)
    m = max(min_coarse, n // ratio)
    if m >= n:
        raise ValueError("coarse_start: too few seeds for a coarse level")
    sel = np.sort(np.random.default_rng(seed).choice(n, m, replace=False))
    if dev:
        import torch
        q_dev = points[torch.from_numpy(sel).to(points.device)].contiguous()
        nn, _ = ctx.knn_search(q_dev, points, 1)
        nn = nn[:, 0].cpu().numpy()
        nu_h = nu.cpu().numpy()
        q = q_dev.cpu().numpy()
    else:
        q = np.ascontiguousarray(np.asarray(points)[sel])
        nn, _ = ctx.knn_search(q, points, 1)
        nn = np.asarray(nn)[:, 0]
        nu_h = np.asarray(nu, dtype=np.float64)
    nu_c = np.bincount(nn.astype(np.int64), weights=nu_h, minlength=m)
    w_c, st = ctx.ot_solve(q, nu_c, box=box, tol=coarse_tol, inexact=True)
    r_l = ctx.cells(q, weights=w_c, box=box, neighbors=False)
    r_v = ctx.cells(q, box=box, neighbors=False)
    pieces, s, mu = _pieces(q, w_c, np.asarray(r_l.vol), np.asarray(r_l.mg), np.asarray(r_v.vol),
                            np.asarray(r_v.mg), lo, hi, s_factor)
    if dev:
        import torch
        pieces = torch.from_numpy(pieces).to(points.device)
    w = ctx.ot_convex_start(points, pieces, s)
    w = w - w.mean()
    info = {"coarse_seeds": m, "coarse_newton": int(st["newton_iters"]),
            "mu_fine": mu[nn.astype(np.int64)],
            "coarse_cell_passes": int(st["cell_passes"]), "coarse_converged": bool(st["converged"]),
            "s": s}
    return w, info


def coarse_start_batch(ctx, points, nu, inst_n, boxes, ratio=128, s_factor=1.0,
                       coarse_tol=1e-3, min_coarse=64, seed=12345):
    """coarse_start for independent instances (consecutive seed ranges, one
    box each, as vgc_ot_solve_batch takes them): the coarse plans are solved
    together by the batched solver, each instance's start is its own convex
    soft maximum.  Returns (w, info).  Default ratio 128: measured on 256
    instances of 2k to 200k seeds, 11.7 s to 1e-4 against 13.8 s at 256."""
    dev = _is_torch_cuda(points)
    inst_n = np.asarray(inst_n, dtype=np.int64)
    boxes = np.ascontiguousarray(boxes, dtype=np.float64).reshape(-1, 6)
    off = np.concatenate([[0], np.cumsum(inst_n)])
    nu_h = nu.cpu().numpy() if dev else np.asarray(nu, dtype=np.float64)
    qs, nus, sizes, nns = [], [], [], []
    for k, nk in enumerate(inst_n):
        m = max(min_coarse, int(nk) // ratio)
        if m >= nk:
            raise ValueError("coarse_start_batch: too few seeds for a coarse level")
        sel = np.sort(np.random.default_rng(seed + k).choice(int(nk), m, replace=False))
        xk = points[int(off[k]):int(off[k + 1])]
        if dev:
            import torch
            qk_dev = xk[torch.from_numpy(sel).to(points.device)].contiguous()
            nn, _ = ctx.knn_search(qk_dev, xk, 1)
            nn = nn[:, 0].cpu().numpy()
            qk = qk_dev.cpu().numpy()
        else:
            qk = np.ascontiguousarray(np.asarray(xk)[sel])
            nn = np.asarray(ctx.knn_search(qk, xk, 1)[0])[:, 0]
        qs.append(qk)
        nus.append(np.bincount(nn.astype(np.int64), weights=nu_h[int(off[k]):int(off[k + 1])],
                               minlength=m))
        sizes.append(m)
    q = np.ascontiguousarray(np.concatenate(qs))
    nu_c = np.concatenate(nus)
    w_c, sts = ctx.ot_solve_batch(q, nu_c, sizes, boxes, tol=coarse_tol, inexact=True)
    r_l = ctx.cells(q, weights=w_c, inst_n=sizes, boxes=boxes, neighbors=False)
    r_v = ctx.cells(q, inst_n=sizes, boxes=boxes, neighbors=False)
    v_l, mg_l = np.asarray(r_l.vol), np.asarray(r_l.mg)
    v_v, mg_v = np.asarray(r_v.vol), np.asarray(r_v.mg)
    coff = np.concatenate([[0], np.cumsum(sizes)])
    parts = []
    for k in range(len(inst_n)):
        a, b = int(coff[k]), int(coff[k + 1])
        pieces, s, _ = _pieces(q[a:b], w_c[a:b], v_l[a:b], mg_l[a:b], v_v[a:b], mg_v[a:b],
                               boxes[k, :3], boxes[k, 3:], s_factor)
        xk = points[int(off[k]):int(off[k + 1])]
        if dev:
            import torch
            pieces = torch.from_numpy(pieces).to(points.device)
        wk = ctx.ot_convex_start(xk, pieces, s)
        parts.append(wk - wk.mean())
    if dev:
        import torch
        w = torch.cat(parts)
    else:
        w = np.concatenate(parts)
    info = {"coarse_seeds": int(sum(sizes)),
            "coarse_newton_max": int(max(st["newton_iters"] for st in sts)),
            "coarse_converged": int(sum(bool(st["converged"]) for st in sts))}
    return w, info


def _is_torch_cuda(a):
    try:
        import torch
    except ImportError:
        return False
    return isinstance(a, torch.Tensor) and a.is_cuda


# ------------------------------------------------------------ instances
def make_instance(seed: int, n: int):
    """Seeds of a 16-component Gaussian mixture in the unit cube and
    log-normal (mu 0, sigma 0.5) target masses summing to 1."""
    from . import sample_points
    x = sample_points(2, seed, n)
    rng = np.random.default_rng(seed + 7919)
    nu = np.exp(rng.normal(0.0, 0.5, n))
    nu /= nu.sum()
    return x, nu


def generate(ctx, seeds, sizes, tol=1e-5, log=print):
    """Solve OT on each instance from the zero start; returns lists of
    features (n x 6 float32) and targets w - mean(w) (float32)."""
    feats, targets = [], []
    for s, n in zip(seeds, sizes):
        t0 = time.time()
        x, nu = make_instance(int(s), int(n))
        if _is_gpu_ctx(ctx):
            import torch
            xd = torch.from_numpy(x).cuda()
            nud = torch.from_numpy(nu).cuda()
            w, st = ctx.ot_solve(xd, nud, tol=tol, verbose=True)
            f = ctx.ot_features(xd, nud)
            t = (w - w.mean()).to(torch.float32)
        else:
            w, st = ctx.ot_solve(x, nu, tol=tol)
            f = ctx.ot_features(x, nu)
            t = (w - w.mean()).astype(np.float32)
        if st["converged"] != 1:
            log(f"instance seed={s} n={n}: not converged ({st['final_err']:.2e}), skipped")
            continue
        feats.append(f)
        targets.append(t)
        log(f"instance seed={s} n={n}: newton={st['newton_iters']} cg={st['cg_iters']} "
            f"{time.time() - t0:.2f}s")
    return feats, targets


def _is_gpu_ctx(ctx):
    return ctx.lib.prefix == "vgc_"


# -------------------------------------------------------------- training
def train(ctx, feats, targets, steps=4000, batch=32768, lr=2e-3, seed=0, log=print,
          val_feats=None, val_targets=None):
    """Adam on the mean squared error of normalised targets.  Arrays are torch
    CUDA tensors for a GPU context, numpy for the emulation.  Returns raw-unit
    parameters (normalisation folded in) and a history of losses."""
    gpu = _is_gpu_ctx(ctx)
    if gpu:
        import torch
        X = torch.cat(feats)
        T = torch.cat(targets)
        mu = X.double().mean(0).cpu().numpy()
        sd = X.double().std(0).cpu().numpy()
        t_mu, t_sd = 0.0, float(T.double().std())
        sd = np.where(sd > 0, sd, 1.0)
        Xn = ((X.double() - torch.from_numpy(mu).cuda()) / torch.from_numpy(sd).cuda()).float()
        Tn = (T.double() / t_sd).float()
        P = torch.from_numpy(init_params(seed)).cuda()
        M = torch.zeros_like(P)
        V = torch.zeros_like(P)
        G = torch.empty_like(P)
        gen = torch.Generator(device="cuda")
        gen.manual_seed(seed)
        N = X.shape[0]
    else:
        X = np.concatenate(feats).astype(np.float64)
        T = np.concatenate(targets).astype(np.float64)
        mu, sd = X.mean(0), X.std(0)
        sd = np.where(sd > 0, sd, 1.0)
        t_mu, t_sd = 0.0, float(T.std())
        Xn = ((X - mu) / sd).astype(np.float32)
        Tn = (T / t_sd).astype(np.float32)
        P = init_params(seed)
        M = np.zeros_like(P)
        V = np.zeros_like(P)
        rng = np.random.default_rng(seed)
        N = X.shape[0]
    hist = []
    for step in range(1, steps + 1):
        if gpu:
            idx = torch.randint(0, N, (min(batch, N),), device="cuda", generator=gen)
            xb = Xn.index_select(0, idx).contiguous()
            tb = Tn.index_select(0, idx).contiguous()
            G, loss = ctx.mlp_grad(P, xb, tb, G)
        else:
            idx = rng.integers(0, N, min(batch, N))
            G, loss = ctx.mlp_grad(P, Xn[idx], Tn[idx])
        # cosine decay to lr / 20
        lr_t = lr * (0.05 + 0.95 * 0.5 * (1.0 + math.cos(math.pi * (step - 1) / steps)))
        ctx.mlp_adam(P, G, M, V, step, lr_t)
        if step % max(1, steps // 20) == 0 or step == steps:
            hist.append((step, loss))
            log(f"step {step}: loss {loss:.4e} (normalised)")
    Pn = P.cpu().numpy() if gpu else P
    raw = fold_normalisation(Pn, mu, sd, t_mu, t_sd)
    out = {"loss": hist, "t_sd": t_sd}
    if val_feats is not None:
        out["val"] = [validate(ctx, raw, f, t) for f, t in zip(val_feats, val_targets)]
    return raw, out


def validate(ctx, params, f, t):
    """Relative L2 error of predicted weights w - mean(w) on one instance."""
    if _is_gpu_ctx(ctx):
        import torch
        y = ctx.mlp_forward(torch.from_numpy(params).cuda(), f).double()
        y = y - y.mean()
        t = t.double()
        return float(torch.linalg.norm(y - t) / torch.linalg.norm(t))
    y = ctx.mlp_forward(params, f).astype(np.float64)
    y -= y.mean()
    t = np.asarray(t, dtype=np.float64)
    return float(np.linalg.norm(y - t) / np.linalg.norm(t))


def instance_list(count, seed=1000, nmin=50_000, nmax=200_000):
    """The fixed list of training instances: seeds seed + i, sizes log-uniform
    in [nmin, nmax] drawn from rng(seed)."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(math.log(nmin), math.log(nmax), count)).astype(np.int64)
    return seed + np.arange(count), sizes


def generate_rows(ctx, seeds, sizes, rows, tol=1e-4, log=print):
    """Solve each instance, keep `rows` random (features, target) rows of it
    (all rows if rows <= 0).  Returns numpy arrays and per-instance stats."""
    F, T, I, stats = [], [], [], []
    for k, (s, n) in enumerate(zip(seeds, sizes)):
        f, t = generate(ctx, [s], [n], tol, log)
        if not f:
            continue
        if _is_gpu_ctx(ctx):
            f, t = f[0].cpu().numpy(), t[0].cpu().numpy()
        else:
            f, t = f[0], t[0]
        if 0 < rows < len(t):
            pick = np.sort(np.random.default_rng(int(s)).choice(len(t), rows, replace=False))
            f, t = f[pick], t[pick]
        F.append(f)
        T.append(t)
        I.append(np.full(len(t), int(s), dtype=np.int64))
        stats.append((int(s), int(n)))
    return (np.concatenate(F).astype(np.float32), np.concatenate(T).astype(np.float32),
            np.concatenate(I), np.array(stats, dtype=np.int64).reshape(-1, 2))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("mode", choices=["generate", "train", "all"])
    ap.add_argument("--instances", type=int, default=200, help="training instances")
    ap.add_argument("--val", type=int, default=8, help="held-out instances after them")
    ap.add_argument("--first", type=int, default=0, help="generate: first instance index")
    ap.add_argument("--count", type=int, default=None, help="generate: number of instances")
    ap.add_argument("--rows", type=int, default=8000, help="rows kept per instance (0: all)")
    ap.add_argument("--nmin", type=int, default=50_000)
    ap.add_argument("--nmax", type=int, default=200_000)
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--data", nargs="*", default=[], help="train: .npz files from generate")
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=32768)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--emulated", action="store_true", help="CPU emulation (small runs only)")
    ap.add_argument("--out", default=PARAMS_FILE, help="generate: .npz; train/all: params .npy")
    ap.add_argument("--report", default=None, help="JSON training report")
    a = ap.parse_args(argv)
    from . import Context

    def log(msg):
        print(msg, flush=True)

    ctx = Context.emulated() if a.emulated else Context(0)
    seeds, sizes = instance_list(a.instances + a.val, a.seed, a.nmin, a.nmax)
    if a.mode in ("generate", "all"):
        lo = a.first if a.mode == "generate" else 0
        hi = len(seeds) if a.count is None or a.mode == "all" else min(len(seeds), lo + a.count)
        t0 = time.time()
        parts = []
        for b in range(lo, hi, 10):  # saved every 10 instances
            e = min(hi, b + 10)
            parts.append(generate_rows(ctx, seeds[b:e], sizes[b:e], a.rows, a.tol, log))
            F, T, I, st = (np.concatenate([p[k] for p in parts]) for k in range(4))
            if a.mode == "generate":
                np.savez(a.out, feats=F, targets=T, inst=I, stats=st)
            log(f"generated {len(st)} instances, {len(T)} rows in {time.time() - t0:.1f}s")
        if a.mode == "generate":
            return 0
        datasets = [(F, T, I)]
    else:
        datasets = []
        for path in a.data:
            z = np.load(path, allow_pickle=False)
            datasets.append((z["feats"], z["targets"], z["inst"]))
    F = np.concatenate([d[0] for d in datasets])
    T = np.concatenate([d[1] for d in datasets])
    I = np.concatenate([d[2] for d in datasets])
    val_seeds = set(int(s) for s in seeds[a.instances:])
    is_val = np.isin(I, list(val_seeds))
    tr_f, tr_t = F[~is_val], T[~is_val]
    va = [(F[I == s], T[I == s]) for s in sorted(val_seeds) if np.any(I == s)]
    if not a.emulated:
        import torch
        tr_f, tr_t = [torch.from_numpy(tr_f).cuda()], [torch.from_numpy(tr_t).cuda()]
        va = [(torch.from_numpy(f).cuda(), torch.from_numpy(t).cuda()) for f, t in va]
    else:
        tr_f, tr_t = [tr_f], [tr_t]
    t1 = time.time()
    params, rep = train(ctx, tr_f, tr_t, a.steps, a.batch, a.lr, a.seed, log,
                        [f for f, _ in va], [t for _, t in va])
    rep.update({"train_instances": int(len(set(I[~is_val].tolist()))),
                "train_rows": int((~is_val).sum()), "val_instances": len(va),
                "t_train_s": time.time() - t1, "steps": a.steps, "batch": a.batch,
                "lr": a.lr, "tol": a.tol, "seed": a.seed, "rows_per_instance": a.rows})
    np.save(a.out, params, allow_pickle=False)
    log(json.dumps(rep))
    if a.report:
        with open(a.report, "w") as fh:
            json.dump(rep, fh, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
