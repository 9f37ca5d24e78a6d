"""Generate training tuples with the FMM itself and train the proposer on the GPU.

  python -m fmm2d_amd.train_proposer --problems 200 --seed 21 --out tuples_a.npz
  python -m fmm2d_amd.train_proposer --data a_tuples.json,b_tuples.json --out proposer.npz

For each random problem (Laplace lfmm2d or Helmholtz hfmm2d; points uniform in
a square, 4-64 Gaussian clusters, on the curve r = 1 + 0.3 cos 5t, or half on
that curve and half in 16 clusters on the unit circle; N log-uniform in
[1e4, nmax] (1e7 by default); zk log-uniform in [1, ZKMAX]; eps in {1e-4, 1e-6, 1e-9}) the FMM
runs on a grid of parameters -- order reduction r_f at the levels whose
default order is below 40, r_c at the levels at or above it (the
high-frequency levels of Helmholtz problems; Laplace orders never reach 40),
and leaf capacity 2^e times the default; mixed precision for eps >= 1e-6,
FP64 below, as the proposals are used -- and each run's GPU time (the
driver's own event timing, timeinfo[11], the minimum of REPEATS calls after a
warm-up) and relative L2 error of pot+grad against the FP64 direct sum at 256
sampled sources are recorded.  The target of a problem is the fastest setting
whose error is at most eps/2: per level, r_c where the default order is at
least 40 and r_f elsewhere, and e.  The network is then fitted by the
library's device kernels (Adam on the squared error).  Every tuple (features,
parameters, error, time) is saved next to the weights.
"""
import argparse
import json
import time

import numpy as np

from . import _native as nat
from . import fmm2d as F
from . import proposer as PR

R_GRID = [0.0, 0.1, 0.2, 0.3, 0.4, 0.5]
RC_GRID = [0.0, 0.2, 0.4]  # reduction at levels whose default order is >= P_HF
P_HF = 40
E_GRID = [-1, 0, 1, 2, 3]
NSAMPLE = 256
REPEATS = 2  # timed calls per setting (after one warm-up call); the minimum is kept
ZKMAX = 1000.0
DISTS = ["uniform", "cluster", "curve", "curvecluster"]


def make_points(rng, n, dist):
    if dist == "uniform":
        return rng.random((2, n))
    if dist == "cluster":
        k = int(rng.integers(4, 65))
        cen = rng.uniform(0.1, 0.9, (2, k))
        sig = np.exp(rng.uniform(np.log(1e-3), np.log(3e-2), k))
        lab = rng.integers(0, k, n)
        p = cen[:, lab] + sig[lab] * rng.standard_normal((2, n))
        m = n // 10
        p[:, :m] = rng.random((2, m))
        return p
    if dist == "curvecluster":
        m = n // 2
        t = rng.uniform(0, 2 * np.pi, m)
        r = 1 + 0.3 * np.cos(5 * t)
        curve = np.stack([r * np.cos(t), r * np.sin(t)]) + 1e-4 * rng.standard_normal((2, m))
        a = rng.uniform(0, 2 * np.pi, 16)
        lab = rng.integers(0, 16, n - m)
        sig = np.exp(rng.uniform(np.log(1e-3), np.log(1e-2), 16))
        cl = np.stack([np.cos(a), np.sin(a)])[:, lab] + sig[lab] * rng.standard_normal((2, n - m))
        return np.concatenate([curve, cl], axis=1)
    t = rng.uniform(0, 2 * np.pi, n)
    r = 1 + 0.3 * np.cos(5 * t)
    return np.stack([r * np.cos(t), r * np.sin(t)]) + 1e-4 * rng.standard_normal((2, n))


def problem_inputs(rng, dists=("uniform", "cluster", "curve"), zkmax=300.0, nmax=1e7):
    """One random problem of the training distribution: (kind, dist, n, eps, zk,
    sources, charges, sampled indices).  Every random draw of a problem is
    made here, in this order, so a seed replays the same problems (the
    defaults are those of the first shipped tuple set)."""
    kind = "h" if rng.random() < 0.5 else "l"
    dist = list(dists)[int(rng.integers(0, len(dists)))]
    n = int(np.exp(rng.uniform(np.log(1e4), np.log(nmax))))
    eps = float([1e-4, 1e-6, 1e-9][int(rng.integers(0, 3))])
    zk = complex(np.exp(rng.uniform(0, np.log(zkmax)))) if kind == "h" else None
    src = make_points(rng, n, dist)
    ch = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    idx = np.sort(rng.choice(n, NSAMPLE, replace=False))
    return kind, dist, n, eps, zk, src, ch, idx


def precision_for(eps):
    """The arithmetic a proposal is used with: mixed precision (FP32 M2L and near
    field) down to eps = 1e-6, FP64 below."""
    return "mixed" if eps >= 1e-6 else "fp64"


def run_problem(rng, dists=DISTS, zkmax=ZKMAX, nmax=1e7):
    kind, dist, n, eps, zk, src, ch, idx = problem_inputs(rng, dists, zkmax, nmax)
    (p, g, _), _ = F._direct(kind, zk, src, src[:, idx], ch, None, None, 2, 1, 1e-16)
    ex = np.concatenate([p[0], g[0].ravel()])
    nd0 = PR.default_ndiv(eps, n, 0)
    feats, bs0 = PR.tree_features(src, None, eps, zk, kind, nd0)
    p0 = PR.default_nterms(kind, eps, [bs0 / 2.0 ** l for l in range(52)], zk)
    hf = [l for l in range(PR.NLEV) if p0[l] >= P_HF]
    prec = precision_for(eps)
    # the inputs stay on the device for every run of the grid; only the
    # sampled outputs come back
    import torch
    dev = F._DevInputs(src, n, 1, ch, None, None, None, True)
    sel = torch.from_numpy(idx).to(dev.src.device)
    bufs = None
    rows = []
    for e in E_GRID:
        for rc in (RC_GRID if hf else [None]):
            for r in R_GRID:
                y = np.zeros(PR.NOUT)
                y[:PR.NLEV] = r
                if rc is not None:
                    y[hf] = rc
                y[PR.NLEV] = e
                nterms, ndiv = PR.decode(y, p0, nd0)
                opts = {"nterms": nterms, "ndiv": ndiv, "precision": prec,
                        "timeinfo": np.zeros(16)}
                bufs, _ = F._fmm_device(kind, eps, zk, dev, 2, 0, 1, dict(opts), bufs)
                best = None
                for _ in range(REPEATS):
                    bufs, ier = F._fmm_device(kind, eps, zk, dev, 2, 0, 1, opts, bufs)
                    t = float(opts["timeinfo"][11])
                    best = t if best is None else min(best, t)
                # device layouts: pot (ns, 1), grad (ns, 2, 1)
                pot = bufs["pot"][sel, 0].cpu().numpy()
                grad = bufs["grad"][sel, :, 0].cpu().numpy().T
                got = np.concatenate([pot, grad.ravel()])
                err = float(np.linalg.norm(got - ex) / np.linalg.norm(ex))
                rows.append({"r": r, "rc": rc, "e": e, "time": best, "err": err,
                             "ier": int(ier)})
    ok = [w for w in rows if w["err"] <= 0.5 * eps and w["ier"] == 0]
    tgt = min(ok, key=lambda w: w["time"]) if ok else {"r": 0.0, "rc": None, "e": 0}
    default = [w for w in rows if w["r"] == 0.0 and w["e"] == 0 and not w["rc"]][0]
    target = [0.0] * PR.NOUT
    for l in range(PR.NLEV):
        target[l] = tgt["rc"] if (tgt["rc"] is not None and l in hf) else tgt["r"]
    target[PR.NLEV] = tgt["e"]
    return {"kind": kind, "dist": dist, "n": n, "eps": eps, "gen": [list(dists), zkmax, nmax],
            "precision": prec,
            "zk": None if zk is None else [zk.real, zk.imag], "features": feats.tolist(),
            "grid": rows, "target": [tgt["r"], tgt["e"]], "target_vec": target,
            "default": default, "chosen": tgt}


def fit(feats, targets, steps=3000, lr=3e-3, seed=0):
    """Fit the network on the device (Glorot-uniform start, Adam on the
    squared error); returns (params, [(step, loss)])."""
    import torch
    L = nat.lib()
    dev = nat.torch_device()
    n = L.fmm2d_mlp_nparams(PR.NIN, PR.NOUT)
    rng = np.random.default_rng(seed)
    parts = []
    for fi, fo in ((PR.NIN, 64), (64, 64), (64, PR.NOUT)):
        a = np.sqrt(6.0 / (fi + fo))
        parts.append(rng.uniform(-a, a, fi * fo))
        parts.append(np.zeros(fo))
    p = torch.from_numpy(np.concatenate(parts).astype(np.float32)).to(dev)
    m = torch.zeros(n, dtype=torch.float32, device=dev)
    v = torch.zeros_like(m)
    g = torch.zeros_like(m)
    loss = torch.zeros(1, dtype=torch.float32, device=dev)
    x = torch.from_numpy(np.asarray(feats, dtype=np.float32)).to(dev)
    t = torch.from_numpy(np.asarray(targets, dtype=np.float32)).to(dev)
    b = x.shape[0]
    hist = []
    for step in range(1, steps + 1):
        nat.check(L.fmm2d_mlp_loss_grad(PR.NIN, PR.NOUT, nat.ptr(p), b, nat.ptr(x), nat.ptr(t),
                                        None, nat.ptr(g), nat.ptr(loss), nat.stream_ptr()),
                  "mlp_loss_grad")
        nat.check(L.fmm2d_mlp_adam(n, nat.ptr(p), nat.ptr(g), nat.ptr(m), nat.ptr(v), lr, 0.9,
                                   0.999, 1e-8, step, nat.stream_ptr()), "mlp_adam")
        if step % 100 == 0 or step == 1:
            hist.append((step, float(loss.cpu())))
    return p.cpu().numpy(), hist


def refeature(problems, seed):
    """Recompute the features of stored problems by replaying their inputs
    from the generator seed (checks that each replayed problem is the stored
    one); the measured tuples and targets do not depend on the features."""
    rng = np.random.default_rng(seed)
    for i, pb in enumerate(problems):
        gen = pb.get("gen", [("uniform", "cluster", "curve"), 300.0])
        nmax = gen[2] if len(gen) > 2 else 1e7
        kind, dist, n, eps, zk, src, ch, idx = problem_inputs(rng, gen[0], gen[1], nmax)
        if (kind, dist, n, eps) != (pb["kind"], pb["dist"], pb["n"], pb["eps"]):
            raise ValueError(f"problem {i}: replay gives {(kind, dist, n, eps)}, stored "
                             f"{(pb['kind'], pb['dist'], pb['n'], pb['eps'])}")
        nd0 = PR.default_ndiv(eps, n, 0)
        feats, _ = PR.tree_features(src, None, eps, zk, kind, nd0)
        pb["features"] = feats.tolist()
        print(f"refeature {i}: {kind} {dist} n={n}", flush=True)


def targets_of(problems):
    y = np.zeros((len(problems), PR.NOUT), dtype=np.float32)
    for i, pb in enumerate(problems):
        if "target_vec" in pb:
            y[i] = pb["target_vec"]
        else:  # one reduction at every level
            y[i, :PR.NLEV] = pb["target"][0]
            y[i, PR.NLEV] = pb["target"][1]
    return y


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--problems", type=int, default=160)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--out", default=PR.WEIGHTS)
    ap.add_argument("--data", default=None, help="reuse tuples from this JSON file")
    ap.add_argument("--nmax", type=float, default=1e7,
                    help="largest problem size (log-uniform from 1e4)")
    ap.add_argument("--tuples-only", action="store_true",
                    help="generate and save the tuples without fitting")
    ap.add_argument("--refeature", action="store_true",
                    help="with --data: replay the stored problems from --seed (the seed they "
                         "were generated with), recompute their features with the current "
                         "feature function and refit; the targets are kept")
    args = ap.parse_args()
    if args.data:
        problems = []
        for path in args.data.split(","):
            with open(path) as f:
                problems += json.load(f)
        if args.refeature:
            refeature(problems, args.seed)
    else:
        rng = np.random.default_rng(args.seed)
        problems = []
        t0 = time.time()
        for i in range(args.problems):
            problems.append(run_problem(rng, nmax=args.nmax))
            pb = problems[-1]
            print(f"problem {i}: {pb['kind']} {pb['dist']} n={pb['n']} eps={pb['eps']:.0e} "
                  f"target r={pb['target'][0]} e={pb['target'][1]} "
                  f"t={pb['chosen'].get('time', 0):.4f}s default={pb['default']['time']:.4f}s "
                  f"({time.time() - t0:.0f}s)", flush=True)
            if args.tuples_only and (i + 1) % 25 == 0:  # keep what a cut-off run has done
                with open(args.out.replace(".npz", "_tuples.json"), "w") as f:
                    json.dump(problems, f)
    if args.tuples_only:
        with open(args.out.replace(".npz", "_tuples.json"), "w") as f:
            json.dump(problems, f)
        print("saved tuples", flush=True)
        return
    feats = np.array([pb["features"] for pb in problems], dtype=np.float32)
    params, hist = fit(feats, targets_of(problems), steps=args.steps)
    for s, l in hist[-3:]:
        print(f"step {s}: loss {l:.5f}", flush=True)
    np.savez(args.out, params=params, nin=PR.NIN, nout=PR.NOUT)
    with open(args.out.replace(".npz", "_tuples.json"), "w") as f:
        json.dump(problems, f)
    with open(args.out.replace(".npz", "_loss.json"), "w") as f:
        json.dump(hist, f)
    print("saved", args.out, flush=True)


if __name__ == "__main__":
    main()
