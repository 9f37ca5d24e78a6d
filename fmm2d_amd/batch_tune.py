"""Leaf-capacity autotuner for batches of independent FMMs.

The reference takes the leaf capacity from lndiv2d / hndiv2d or from the
caller (laplace/lfmm2d_ndiv.f:29-32, helmholtz/hfmm2d_ndiv.f:19-22, the
`ndiv` of fmm2d_opts here); it has no batched driver and no tuner.  For a
batch, problems are grouped by kind, eps, arithmetic, size (a factor of 2 in
N) and, for Helmholtz, the number of wavelengths across the domain (a factor
of 2); one representative per group (median N) is timed with the default
leaf capacity times each of FACTORS, and the fastest factor is applied to the
whole group.  The leaf capacity does not enter the error bound (the orders
come from the reference's rules for the tree's box sizes), so the tuned calls
keep the reference's accuracy; the batch test checks every instance against
the FP64 direct sum.

Arithmetic: "auto" runs eps >= 1e-6 in mixed precision (FP32 M2L on MFMA,
FP32 near field summed in FP64) and smaller eps in FP64, as the single-problem
pipelines do.
"""
import time

import numpy as np

from . import _native as nat
from . import fmm2d as F

FACTORS = (1.0, 2.0, 4.0, 8.0, 16.0)


def precision_for(eps):
    return "mixed" if eps >= 1e-6 else "fp64"


def _size(pb):
    src = np.asarray(pb["sources"])
    ns = src.shape[1]
    nt = 0 if pb.get("targets") is None or not pb.get("pgt") else np.asarray(pb["targets"]).shape[1]
    return src, ns, nt


def group_key(pb, precision="auto"):
    """(kind, eps, arithmetic, log2 N, log2 wavelengths across) of a problem."""
    src, ns, nt = _size(pb)
    prec = precision_for(pb["eps"]) if precision == "auto" else precision
    waves = 0
    if pb["kind"] == "h":
        diam = float(max(np.ptp(src[0]), np.ptp(src[1])))
        waves = int(np.floor(np.log2(1.0 + abs(complex(pb["zk"])) * diam / (2 * np.pi))))
    return (pb["kind"], float(pb["eps"]), prec, int(np.floor(np.log2(ns + nt))), waves)


def _time_call(pb, opts, repeats):
    """Best wall time of `repeats` calls (after one warm-up) with inputs on the device."""
    import torch
    src, ns, nt = _size(pb)
    nd = int(pb.get("nd", 1))
    pg, pgt = int(pb.get("pg", 0)), int(pb.get("pgt", 0))
    dev = F._DevInputs(src, ns, nd, pb.get("charges"), pb.get("dipstr"), pb.get("dipvec"),
                       pb.get("targets") if pgt > 0 else None, pb["kind"] != "c",
                       real=(pb["kind"] == "r"))
    zk = pb.get("zk")
    bufs, ier = F._fmm_device(pb["kind"], pb["eps"], zk, dev, pg, pgt, nd, dict(opts))
    nat.check(ier, "tune")
    best = None
    for _ in range(repeats):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        F._fmm_device(pb["kind"], pb["eps"], zk, dev, pg, pgt, nd, dict(opts), bufs=bufs)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    return best


def tune(problems, factors=FACTORS, repeats=2, precision="auto"):
    """{group key: {"factor": f, "times": {factor: seconds}, "n": group size}}."""
    from .proposer import default_ndiv
    groups = {}
    for k, pb in enumerate(problems):
        groups.setdefault(group_key(pb, precision), []).append(k)
    table = {}
    for key, members in groups.items():
        sizes = [_size(problems[k])[1] + _size(problems[k])[2] for k in members]
        rep = problems[members[int(np.argsort(sizes)[len(sizes) // 2])]]
        _, ns, nt = _size(rep)
        nd0 = default_ndiv(rep["eps"], ns, nt)
        times = {}
        for f in factors:
            times[f] = _time_call(rep, {"precision": key[2],
                                        "ndiv": max(4, int(round(nd0 * f)))}, repeats)
        best = min(times, key=times.get)
        table[key] = {"factor": best, "times": times, "n": len(members)}
    return table


def apply(problems, table, precision="auto"):
    """Copies of the problems with "opts" (leaf capacity and arithmetic) set
    from a table of tune(); problems of groups the table lacks keep the
    reference's leaf capacity."""
    from .proposer import default_ndiv
    out = []
    for pb in problems:
        key = group_key(pb, precision)
        _, ns, nt = _size(pb)
        opts = {"precision": key[2]}
        if key in table:
            opts["ndiv"] = max(4, int(round(default_ndiv(pb["eps"], ns, nt) * table[key]["factor"])))
        out.append(dict(pb, opts=opts))
    return out
