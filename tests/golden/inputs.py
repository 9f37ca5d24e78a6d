"""Seeded input generators shared by the fixture writer (make_golden.py) and
the GPU tests that compare against those fixtures.

The fixtures store the reference's outputs and a SHA-256 of the generated
inputs; a test regenerates the inputs from the same seed and first checks
that hash, so a generator that changed is reported as such and not as a
parity failure.
"""
import hashlib

import numpy as np


def clustered(n, seed):
    """Clustered sources: 90% from 64 isotropic Gaussians (centres U[0.05,0.95]^2,
    sigma log-uniform in [1e-3, 1e-2]), 10% uniform in [0,1]^2; real charges N(0,1).
    Returns sources (2,n) and charges (n,) float64."""
    rng = np.random.default_rng(seed)
    nb = n // 10
    nc = n - nb
    cen = rng.uniform(0.05, 0.95, (64, 2))
    sig = np.exp(rng.uniform(np.log(1e-3), np.log(1e-2), 64))
    lab = rng.integers(0, 64, nc)
    pts = cen[lab] + sig[lab, None] * rng.standard_normal((nc, 2))
    pts = np.concatenate([pts, rng.random((nb, 2))])
    q = rng.standard_normal(n)
    return np.ascontiguousarray(pts.T), q


def helmholtz_st(n, seed):
    """n sources and n targets uniform in [0,1]^2, charges and dipole strengths
    CN(0,1), dipole directions at uniform random angles.
    Returns sources (2,n), targets (2,n), charges (n,), dipstr (n,), dipvec (2,n)."""
    rng = np.random.default_rng(seed)
    src = rng.random((2, n))
    trg = rng.random((2, n))
    ch = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2)
    ds = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2)
    th = rng.uniform(0, 2 * np.pi, n)
    dv = np.stack([np.cos(th), np.sin(th)])
    return src, trg, ch, ds, dv


def digest(*arrays):
    """SHA-256 (hex) of the arrays' bytes in C order."""
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()
