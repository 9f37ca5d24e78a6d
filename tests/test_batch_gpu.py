"""fmm2d_batch (many independent FMMs per call, no counterpart in the
reference): every problem of a mixed batch -- all four kinds, several sizes,
flags and nd -- matches the same problem run alone through the single-problem
API.  The list-3 sums take FP64 atomic adds in LDS, whose order can vary
between runs, so the bar is 1e-12 relative rather than bit equality."""
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
                assert _rel(a, b) < 1e-12, (pb["kind"], name, _rel(a, b))
