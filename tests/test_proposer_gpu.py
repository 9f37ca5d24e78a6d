"""The trained proposer (fmm2d_amd/data/proposer.npz) on problems it was not
trained on: its proposals meet eps against the FP64 direct sum and, summed
over the problems, run faster than the reference's nterms / ndiv rules (both in
the arithmetic a proposal is used with: mixed precision down to eps = 1e-6)."""
import numpy as np
import pytest

from fmm2d_amd import fmm2d as F
from fmm2d_amd import proposer as PR
from fmm2d_amd.train_proposer import make_points, precision_for

pytestmark = pytest.mark.gpu

# (kind, dist, n, eps, zk): held out (seed differs from the training run's)
HELD_OUT = [("l", "cluster", 120000, 1e-6, None), ("h", "uniform", 60000, 1e-6, 40.0),
            ("l", "curve", 80000, 1e-9, None), ("h", "cluster", 50000, 1e-4, 120.0)]


def _time(kind, eps, zk, src, ch, opts):
    best = None
    F._fmm(kind, eps, zk, src, ch, None, None, None, 2, 0, 1, opts=dict(opts))
    for _ in range(3):
        o = dict(opts)
        o["timeinfo"] = np.zeros(16)
        F._fmm(kind, eps, zk, src, ch, None, None, None, 2, 0, 1, opts=o)
        t = float(o["timeinfo"][11])
        best = t if best is None else min(best, t)
    return best


def test_proposals_meet_eps_and_beat_default_rules():
    rng = np.random.default_rng(1234)
    prop = PR.Proposer()
    t_prop = t_def = 0.0
    for kind, dist, n, eps, zk in HELD_OUT:
        src = make_points(rng, n, dist)
        ch = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        zkc = None if zk is None else complex(zk)
        # the arithmetic the proposals are trained for and used with
        prec = {"precision": precision_for(eps)}
        res, info = PR.fmm_validated(kind, eps, src, ch, zk=zkc, proposer=prop, opts=prec)
        assert info["used"] == "proposal", info
        assert info["err"] <= eps, info
        p = info["proposal"]
        tp = _time(kind, eps, zkc, src, ch, dict(prec, nterms=p["nterms"], ndiv=p["ndiv"]))
        td = _time(kind, eps, zkc, src, ch, prec)
        print(f"{kind} {dist} n={n} eps={eps:.0e}: proposal {tp*1e3:.2f} ms "
              f"(err {info['err']:.2e}), default {td*1e3:.2f} ms")
        t_prop += tp
        t_def += td
    assert t_prop < t_def, (t_prop, t_def)


@pytest.mark.parametrize("kind", ["l", "h"])
def test_prepared_handle_leaves_fmm_unchanged(kind):
    """An fmm2d_prepare handle (the sorted keys the proposer's features use)
    handed to the FMM in opts gives the same tree and results as the FMM's own
    build; a handle for other arrays is ignored."""
    from fmm2d_amd import _native as nat
    rng = np.random.default_rng(3)
    n, nt = 150000, 4000
    src = make_points(rng, n, "cluster")
    trg = rng.uniform(src.min(), src.max(), (2, nt))
    ch = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    zk = 30.0 if kind == "h" else None
    eps = 1e-6
    print(f"prepared handle: kind {kind} n {n} nt {nt}", flush=True)
    dev = F._DevInputs(src, n, 1, ch, None, None, trg, True)
    other = F._DevInputs(src[:, : n // 2], n // 2, 1, ch[: n // 2], None, None, None, True)
    runs = {}
    for tag in ("own", "prepared", "other"):
        prep = None
        if tag == "prepared":
            prep = PR.Prepared(nat.ptr(dev.src), n, nat.ptr(dev.targ), nt)
        elif tag == "other":
            prep = PR.Prepared(nat.ptr(other.src), n // 2)
        opts = {"precision": "mixed", "prepared": prep}
        bufs, ier = F._fmm_device(kind, eps, zk, dev, 2, 2, 1, opts)
        assert ier == 0
        if prep is not None:
            prep.close()
        runs[tag] = ({k: v.cpu().numpy() for k, v in bufs.items() if v is not None},
                     (opts["nlevels"], opts["nboxes"]))
    ref, shape = runs["own"]
    for tag in ("prepared", "other"):
        got, sh = runs[tag]
        assert sh == shape, tag
        for k, v in ref.items():
            err = np.linalg.norm(got[k] - v) / np.linalg.norm(v)
            assert err < 1e-13, (tag, k, err)
