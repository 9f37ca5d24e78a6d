"""Learned parameter proposer for the FMM drivers.

A perceptron 20 -> 64 -> 64 -> 17 (tanh, tanh, linear; FP32) run and trained
by the library's own device kernels (fmm2d_mlp_forward / _loss_grad / _adam,
csrc/mlp.hip).  It maps a problem to expansion orders per level and a leaf
capacity; the drivers take them through fmm2d_opts.nterms / ndiv.  The
reference's rules (laplace/l2dterms.f:38, helmholtz/h2dterms.f:35,
laplace/lndiv2d.f) stay the fallback: a proposal whose error against the
FP64 direct sum on sampled points exceeds eps is rejected.

Inputs (20): log10 N / 7, log10(1 + |zk| boxsize0) / 3, -log10 eps / 15, a
Helmholtz flag, and the fraction of points held by leaves at levels 0..15 of
the tree refined with the default leaf capacity (no level restriction;
computed on the device from the sorted path keys).
Outputs (17): r_l, l = 0..15, the fractional reduction of level l's default
order (nterms_l = max(2, round(p_l (1 - r_l))), r_l clipped to [0, 0.7]), and
log2 of the leaf capacity relative to the default one (clipped to [-1.5, 3]).
"""
import os

import numpy as np

from . import _native as nat

NIN, NOUT, NLEV = 20, 17, 16
MARGIN = 0.05  # taken off the predicted reductions (the training grid step is 0.1)
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "proposer.npz")


def default_ndiv(eps, ns, nt):
    """lndiv2d / hndiv2d (laplace/lndiv2d.f)."""
    import ctypes as C
    L = nat.lib()
    nd, idf = C.c_int(0), C.c_int(0)
    nat.check(L.fmm2d_ndiv(float(eps), int(ns), int(nt), 0, C.byref(nd), C.byref(idf)), "ndiv")
    return nd.value


def default_nterms(kind, eps, boxsize, zk=None):
    """Per-level orders of the reference's rules for the given box sizes."""
    import ctypes as C
    L = nat.lib()
    out = []
    for bs in boxsize:
        v = C.c_int(0)
        if kind == "h":
            z = complex(zk)
            nat.check(L.fmm2d_h2dterms(float(bs), z.real, z.imag, float(eps), C.byref(v)),
                      "h2dterms")
        else:
            nat.check(L.fmm2d_l2dterms(float(eps), C.byref(v)), "l2dterms")
        out.append(v.value)
    return out


def tree_features(sources, targets, eps, zk, kind, ndiv):
    """Feature vector (NIN floats) and the root box size for host points
    (copied to the device; see device_features)."""
    import torch
    dev = nat.torch_device()
    src = torch.from_numpy(np.ascontiguousarray(np.asarray(sources, dtype=np.float64).T)).to(dev)
    ns = src.shape[0]
    trg, nt = None, 0
    if targets is not None:
        trg = torch.from_numpy(np.ascontiguousarray(np.asarray(targets, dtype=np.float64).T)).to(
            dev)
        nt = trg.shape[0]
    return device_features(nat.ptr(src), ns, nat.ptr(trg), nt, eps, zk, kind, ndiv)


class Prepared:
    """An fmm2d_prepare handle: the bounding square and sorted path keys of
    device-resident points, shared by the proposer's features
    (fmm2d_prepared_leaf_histogram) and the FMM's tree build (opts
    "prepared").  The point arrays must stay unchanged while it is open, and
    it is used on the stream it was made on."""

    def __init__(self, src_ptr, ns, targ_ptr=None, nt=0, stream=None):
        import ctypes as C
        self.L = nat.lib()
        h = C.c_void_p()
        self.stream = nat.stream_ptr() if stream is None else stream
        nat.check(self.L.fmm2d_prepare(src_ptr, int(ns), targ_ptr, int(nt), C.byref(h),
                                       self.stream), "prepare")
        self.handle = h.value

    def close(self):
        if self.handle:
            self.L.fmm2d_prepare_free(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()


def device_features(src_ptr, ns, targ_ptr, nt, eps, zk, kind, ndiv, stream=None,
                    prepared=None):
    """Feature vector (NIN floats) and the root box size for device-resident
    points (sources(2,ns), targets(2,nt) as device pointers): log10 N / 7,
    log10(1 + |zk| boxsize0) / 3, -log10 eps / 15, a Helmholtz flag, and the
    fraction of points held by the leaves of levels 0..15 of the tree refined
    while a box holds more than ndiv sources (fmm2d_key_leaf_histogram: from
    the sorted path keys, no tree built and nothing copied to the host but the
    16 counts).  prepared: a Prepared of the same points, whose keys are used
    instead of sorting them here (the same counts)."""
    import ctypes as C
    L = nat.lib()
    hist = np.zeros(NLEV, dtype=np.float64)
    bs0 = C.c_double(0.0)
    stream = nat.stream_ptr() if stream is None else stream
    if prepared is not None:
        nat.check(L.fmm2d_prepared_leaf_histogram(prepared.handle, int(ndiv), NLEV,
                                                  hist.ctypes.data, C.byref(bs0), stream),
                  "prepared_leaf_histogram")
    else:
        nat.check(L.fmm2d_key_leaf_histogram(src_ptr, int(ns), targ_ptr, int(nt), int(ndiv),
                                             NLEV, hist.ctypes.data, C.byref(bs0), stream),
                  "key_leaf_histogram")
    hist = hist / max(1.0, hist.sum())
    kz = 0.0 if kind != "h" else abs(complex(zk)) * bs0.value
    f = np.concatenate([[np.log10(ns + nt) / 7.0, np.log10(1.0 + kz) / 3.0,
                         -np.log10(eps) / 15.0, 1.0 if kind == "h" else 0.0], hist])
    return f.astype(np.float32), bs0.value


def decode(y, p_default, ndiv_default, margin=0.0):
    """Network output -> (nterms per level, ndiv); margin is taken off the
    predicted order reductions (a proposal's safety allowance)."""
    r = np.clip(np.asarray(y[:NLEV], dtype=np.float64) - margin, 0.0, 0.7)
    nl = len(p_default)
    rr = np.array([r[min(l, NLEV - 1)] for l in range(nl)])
    nterms = [int(max(2, round(p * (1.0 - x)))) for p, x in zip(p_default, rr)]
    e = float(np.clip(y[NLEV], -1.5, 3.0))
    ndiv = int(np.clip(round(ndiv_default * 2.0 ** e), 4, 400))
    return nterms, ndiv


class Proposer:
    """Holds the network parameters on the device."""

    def __init__(self, params=None, path=WEIGHTS):
        import torch
        if params is None:
            with np.load(path) as z:
                params = z["params"]
        self.L = nat.lib()
        assert len(params) == self.L.fmm2d_mlp_nparams(NIN, NOUT)
        self.params = torch.from_numpy(np.asarray(params, dtype=np.float32)).to(
            nat.torch_device())

    def predict(self, feats):
        """Network outputs for a batch of feature rows (numpy [batch][NIN])."""
        import torch
        x = torch.from_numpy(np.ascontiguousarray(feats, dtype=np.float32)).to(
            self.params.device)
        b = x.shape[0]
        y = torch.zeros((b, NOUT), dtype=torch.float32, device=x.device)
        nat.check(self.L.fmm2d_mlp_forward(NIN, NOUT, nat.ptr(self.params), b, nat.ptr(x),
                                           nat.ptr(y), nat.stream_ptr()), "mlp_forward")
        return y.cpu().numpy()

    def propose(self, kind, sources, eps, targets=None, zk=None):
        """dict(nterms=[...per level...], ndiv=k) for host points."""
        import torch
        dev = nat.torch_device()
        src = torch.from_numpy(np.ascontiguousarray(np.asarray(sources, dtype=np.float64).T)).to(
            dev)
        trg, nt = None, 0
        if targets is not None:
            trg = torch.from_numpy(np.ascontiguousarray(
                np.asarray(targets, dtype=np.float64).T)).to(dev)
            nt = trg.shape[0]
        return self.propose_device(kind, nat.ptr(src), src.shape[0], eps,
                                   targ_ptr=nat.ptr(trg), nt=nt, zk=zk)

    def propose_device(self, kind, src_ptr, ns, eps, targ_ptr=None, nt=0, zk=None,
                       nlevels_max=52, prepared=None):
        """propose() for device-resident points: features from the device leaf
        histogram, the network on the device, orders for levels
        0..nlevels_max-1 (the drivers read the entries of their tree's
        levels).  prepared: a Prepared of the same points (the FMM call that
        follows takes it too, in its opts)."""
        nd0 = default_ndiv(eps, ns, nt)
        f, bs0 = device_features(src_ptr, ns, targ_ptr, nt, eps, zk, kind, nd0,
                                 prepared=prepared)
        y = self.predict(f[None, :])[0]
        bsx = [bs0 / 2.0 ** l for l in range(nlevels_max)]
        p0 = default_nterms(kind, eps, bsx, zk)
        nterms, ndiv = decode(y, p0, nd0, margin=MARGIN)
        return {"nterms": nterms, "ndiv": ndiv}


def fmm_validated(kind, eps, sources, charges=None, dipstr=None, dipvec=None, zk=None,
                  pg=2, proposer=None, nsample=256, seed=0, opts=None):
    """Run the FMM at sources with the network's proposal, check its relative
    L2 error (pot and grad) against the FP64 direct sum at nsample sampled
    sources, and fall back to the reference rules if it exceeds eps.

    Returns (outputs dict, info) with info = {"used": "proposal" | "default",
    "err": error of the returned outputs, "proposal": {...}, "err_proposal"}.
    """
    from . import fmm2d as F
    src = np.asarray(sources, dtype=np.float64)
    n = src.shape[1]
    prop = (proposer or Proposer()).propose(kind, src, eps, zk=zk)
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(n, min(nsample, n), replace=False))
    (p, g, _), _ = F._direct(kind, zk, src, src[:, idx], charges, dipstr, dipvec, 2, 1, 1e-16)

    def err_of(res):
        gp, gg = res["pot"][idx], res["grad"][..., idx].ravel()
        if kind == "c":  # Re(pot) only, as comperr's Cauchy mode
            gp, ep = gp.real, p[0].real
        else:
            ep = p[0]
        got, want = np.concatenate([gp, gg]), np.concatenate([ep, g[0].ravel()])
        return float(np.linalg.norm(got - want) / np.linalg.norm(want))

    o = dict(opts or {})
    o.update(nterms=prop["nterms"], ndiv=prop["ndiv"])
    res, _ = F._fmm(kind, eps, zk, src, charges, dipstr, dipvec, None, pg, 0, 1, opts=o)
    e = err_of(res)
    info = {"used": "proposal", "err": e, "err_proposal": e, "proposal": prop}
    if e > eps:
        res, _ = F._fmm(kind, eps, zk, src, charges, dipstr, dipvec, None, pg, 0, 1,
                        opts=dict(opts or {}))
        info.update(used="default", err=err_of(res))
    return res, info
