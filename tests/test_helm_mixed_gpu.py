"""Mixed-precision hfmm2d (fmm2d_opts.precision = 1): FP32 M2L on MFMA and, for
real wavenumbers, an FP32 near field (FP32 Hankel pair, csrc/hankel32.h) with
FP64 sums.  At eps = 1e-6 the relative L2 error of pot and grad against the FP64
direct sum at sampled points stays below eps, and the outputs stay within eps
of the FP64 FMM.  Cases: the Helmholtz workload's shape (uniform sources and
separate targets, zk = 100, charges and dipoles), the proposer's evaluation
shape (curve plus Gaussian clusters, zk = 500, potential and gradient at the
sources), a complex wavenumber (FP32 M2L, FP64 near field) and duplicate /
1e-11-close sources (the out-of-line FP64 pairs).  The direct sums come from
the FP64 direct evaluator, pinned to the reference's h2d_direct kernels."""
import numpy as np
import pytest

from fmm2d_amd import fmm2d as F

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


def _curve_cluster(rng, n):
    m = n // 2
    t = rng.uniform(0, 2 * np.pi, m)
    r = 1 + 0.3 * np.cos(5 * t)
    curve = np.stack([r * np.cos(t), r * np.sin(t)]) + 1e-4 * rng.standard_normal((2, m))
    cen = np.stack([np.cos(rng.uniform(0, 2 * np.pi, 16)), np.sin(rng.uniform(0, 2 * np.pi, 16))])
    lab = rng.integers(0, 16, n - m)
    sig = np.exp(rng.uniform(np.log(1e-3), np.log(1e-2), 16))
    cl = cen[:, lab] + sig[lab] * rng.standard_normal((2, n - m))
    return np.concatenate([curve, cl], axis=1)


@pytest.mark.parametrize("case", ["c2", "c5", "complex", "duplicates"])
def test_helmholtz_mixed_vs_direct(case):
    rng = np.random.default_rng(31)
    eps = 1e-6
    if case == "c2":
        n, zk = 40000, 100.0
        src = rng.random((2, n))
        trg = rng.random((2, 3000))
        ch = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2)
        ds = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2)
        th = rng.uniform(0, 2 * np.pi, n)
        dv = np.stack([np.cos(th), np.sin(th)])
        pg, pgt = 0, 2
    else:
        n = 40000
        src = _curve_cluster(rng, n) if case != "complex" else rng.random((2, n))
        zk = {"c5": 500.0, "complex": 60.0 + 3.0j, "duplicates": 80.0}[case]
        if case == "duplicates":
            src[:, 100:200] = src[:, :100]
            src[:, 200:300] = src[:, :100] + 1e-11
        trg = None
        ch = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2)
        ds = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2)
        dv = rng.standard_normal((2, n))
        pg, pgt = 2, 0
    print(f"hfmm2d mixed case {case}: n={n} zk={zk}", flush=True)
    args = ("h", eps, zk, src, ch, ds, dv, trg, pg, pgt, 1)
    mixed, ier = F._fmm(*args, opts={"precision": "mixed"})
    assert ier == 0
    fp64, ier = F._fmm(*args)
    assert ier == 0
    if pgt:
        (p, g, _), _ = F._direct("h", zk, src, trg, ch, ds, dv, 2, 1, 1e-16)
        assert _rel(mixed["pottarg"], p[0]) < eps
        assert _rel(mixed["gradtarg"], g[0]) < eps
    else:
        idx = np.sort(rng.choice(n, 600, replace=False))
        if case == "duplicates":
            idx = np.unique(np.concatenate([idx, np.arange(300)]))
        (p, g, _), _ = F._direct("h", zk, src, src[:, idx], ch, ds, dv, 2, 1, 1e-16)
        # the direct sum includes a target's own source pair only through the
        # threshold rule; the FMM drops it (and its duplicates) by the same rule
        assert _rel(mixed["pot"][idx], p[0]) < eps, _rel(mixed["pot"][idx], p[0])
        assert _rel(mixed["grad"][:, idx], g[0]) < eps
    for k in mixed:
        assert _rel(mixed[k], fp64[k]) < eps, (k, _rel(mixed[k], fp64[k]))
