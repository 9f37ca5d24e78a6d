"""fmm2d_batch (many independent FMMs per call, no counterpart in the
reference): every problem of a mixed batch -- all four kinds, several sizes,
flags and nd -- gives the same bits as the same problem run alone through the
single-problem API (the FMM's sums run in a fixed order).  A batch tuned by
batch_tune (leaf capacity per group, mixed precision at eps >= 1e-6) is
checked instance by instance against the FP64 direct sum."""
import numpy as np
import pytest

from fmm2d_amd import fmm2d as F

pytestmark = pytest.mark.gpu


def _problems():
    rng = np.random.default_rng(21)
    out = []
    specs = [("l", 3000, 1, 2, 0), ("h", 2000, 1, 2, 2), ("r", 5000, 2, 1, 1),
             ("c", 4000, 1, 3, 0), ("l", 800, 2, 0, 3), ("h", 6000, 1, 0, 1),
             ("c", 1500, 1, 1, 1), ("l", 20000, 1, 2, 0)]
    for k, (kind, n, nd, pg, pgt) in enumerate(specs):
        src = rng.random((2, n))
        real = kind == "r"
        ch = rng.standard_normal((nd, n)) + (0 if real else 1j * rng.standard_normal((nd, n)))
        ds = rng.standard_normal((nd, n)) + (0 if real else 1j * rng.standard_normal((nd, n)))
        pb = dict(kind=kind, eps=1e-9 if k % 2 else 1e-6, sources=src, pg=pg, pgt=pgt, nd=nd,
                  charges=ch[0] if nd == 1 else ch, dipstr=ds[0] if nd == 1 else ds)
        if kind != "c":
            dv = rng.standard_normal((nd, 2, n))
            pb["dipvec"] = dv[0] if nd == 1 else dv
        if pgt:
            pb["targets"] = rng.random((2, 700))
        if kind == "h":
            pb["zk"] = 30.0 if k == 1 else 5.0 + 0.5j
        out.append(pb)
    return out


def _single(pb):
    args = {k: v for k, v in pb.items() if k != "kind"}
    return {"l": F.lfmm2d, "h": F.hfmm2d, "r": F.rfmm2d, "c": F.cfmm2d}[pb["kind"]](**args)


def _rel(a, b):
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


@pytest.mark.parametrize("nstreams", [1, 4])
def test_batch_matches_single_calls(nstreams):
    probs = _problems()
    outs = F.fmm2d_batch(probs, nstreams=nstreams)
    assert len(outs) == len(probs)
    for pb, out in zip(probs, outs):
        ref = _single(pb)
        assert out.ier == 0
        for name in ("pot", "grad", "hess", "pottarg", "gradtarg", "hesstarg"):
            a, b = getattr(out, name), getattr(ref, name)
            assert (a is None) == (b is None), (pb["kind"], name)
            if a is not None:
                assert a.shape == b.shape and a.dtype == b.dtype, (pb["kind"], name)
                assert np.array_equal(a, b), (pb["kind"], name, _rel(a, b))


def _c4_like(count, seed=4):
    """The batched configuration's mix at smaller N: disk or perturbed curve,
    lfmm2d / hfmm2d alternating, zk log-uniform in [1, 300], eps 1e-6 / 1e-9."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(np.exp(rng.uniform(np.log(1e3), np.log(3e4))))
        t = rng.uniform(0, 2 * np.pi, n)
        if rng.random() < 0.5:
            r = np.sqrt(rng.random(n))
            src = np.stack([r * np.cos(t), r * np.sin(t)])
        else:
            r = 1 + 0.3 * np.cos(5 * t)
            src = np.stack([r * np.cos(t), r * np.sin(t)]) + 1e-4 * rng.standard_normal((2, n))
        ch = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2)
        pb = dict(kind="h" if k % 2 else "l", eps=float(rng.choice([1e-6, 1e-9])), sources=src,
                  charges=ch, pg=2)
        if k % 2:
            pb["zk"] = float(np.exp(rng.uniform(0, np.log(300.0))))
        out.append(pb)
    return out


def test_tuned_batch_within_eps_of_direct_sums():
    from fmm2d_amd import batch_tune as BT
    probs = _c4_like(24)
    print(f"tuned batch: {len(probs)} instances", flush=True)
    table = BT.tune(probs, repeats=1)
    assert all(v["factor"] in BT.FACTORS for v in table.values())
    tuned = BT.apply(probs, table)
    assert all(p["opts"]["precision"] == ("mixed" if p["eps"] >= 1e-6 else "fp64") for p in tuned)
    outs = F.fmm2d_batch(tuned, nstreams=4)
    rng = np.random.default_rng(2)
    for pb, out in zip(tuned, outs):
        assert out.ier == 0
        ns = pb["sources"].shape[1]
        idx = np.sort(rng.choice(ns, 40, replace=False))
        zk = pb.get("zk", 0.0)
        (p, g, _), _ = F._direct(pb["kind"], zk, pb["sources"], pb["sources"][:, idx],
                                 pb["charges"], None, None, 2, 1, 1e-14)
        e_p = _rel(out.pot[idx], p[0])
        e_g = _rel(out.grad[:, idx], g[0])
        assert e_p < pb["eps"] and e_g < pb["eps"], (pb["kind"], pb["eps"], ns, e_p, e_g)
