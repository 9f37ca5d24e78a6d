"""GPU FMM parity: against the reference's own FMM outputs and against the
direct sum.

tests/golden/fmm_ref.npz holds cfmm2d / lfmm2d / hfmm2d outputs of the
reference library on 1200 sources (300 of them in a tight cluster) and 700
targets, for several (charge, dipole, pg, pgt) cases, nd = 1 and 2, at two
precisions.  Our FMM uses the same tree, lists and expansion orders, so it
must agree with the reference far below eps: the bar is max(1e-3*eps, 1e-13) relative
L2 (FP64).  Against the direct sum the bar is the reference's own test
bar: relative L2 error < eps (test/laplace/test_cfmm2d.f:56-205).
"""
import numpy as np
import pytest

from fmm2d_amd import fmm2d as F

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


CASES1 = [(1, 0, 1, 0), (1, 0, 2, 2), (0, 1, 0, 2), (1, 1, 3, 3)]


@pytest.mark.parametrize("kind", ["c", "l"])
@pytest.mark.parametrize("eps", [1e-6, 1e-12])
@pytest.mark.parametrize("nd,case", [(1, c) for c in CASES1] + [(2, (1, 1, 2, 2))])
def test_laplace_matches_reference(golden_dir, kind, eps, nd, case):
    g = np.load(f"{golden_dir}/fmm_ref.npz")
    ifc, ifd, pg, pgt = case
    src, trg = g["src"], g["trg"]
    ch, ds, dv = g[f"nd{nd}_charge"], g[f"nd{nd}_dipstr"], g[f"nd{nd}_dipvec"]
    res, ier = F._fmm(kind, eps, None, src, ch if ifc else None, ds if ifd else None,
                      dv if ifd else None, trg if pgt else None, pg, pgt, nd)
    assert ier == 0
    tag = f"nd{nd}_{kind}_{eps:.0e}_{ifc}{ifd}{pg}{pgt}_"
    for k, v in res.items():
        ref = g[tag + k]
        if nd == 1:
            ref = ref[0]
        e = _rel(v, ref)
        assert e < max(1e-3 * eps, 1e-13), (k, e)


def test_lfmm2d_public_api_vs_direct():
    rng = np.random.default_rng(3)
    n = 20000
    src = rng.random((2, n))
    ch = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    out = F.lfmm2d(eps=1e-6, sources=src, charges=ch, pg=2)
    assert out.pot.shape == (n,) and out.grad.shape == (2, n)
    m = 400
    ex = F.l2ddir(sources=src, targets=src[:, :m], charges=ch, pgt=2)
    err = F.comperr(ntest=m, out=out, outex=_as_src(ex), pg=2)
    assert err < 1e-6, err


def test_cfmm2d_targets_vs_direct():
    rng = np.random.default_rng(4)
    n, nt = 30000, 5000
    src = rng.random((2, n))
    src[:, : n // 2] = 0.5 + 0.01 * rng.standard_normal((2, n // 2))
    trg = rng.random((2, nt))
    ch = rng.uniform(-1, 1, n) + 0j
    ds = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    out = F.cfmm2d(eps=1e-9, sources=src, charges=ch, dipstr=ds, targets=trg, pgt=3)
    ex = F.c2ddir(sources=src, targets=trg, charges=ch, dipstr=ds, pgt=3)
    assert _rel(out.pottarg.real, ex.pottarg.real) < 1e-9
    assert _rel(out.gradtarg, ex.gradtarg) < 1e-9
    assert _rel(out.hesstarg, ex.hesstarg) < 1e-9


def _as_src(ex):
    o = F.Output()
    o.pot = ex.pottarg
    o.grad = ex.gradtarg
    o.hess = ex.hesstarg
    return o
