"""The reference's case matrix (docs/fortran-c.rst:30-66; test/laplace/test_lfmm2d.f):
charges / dipoles / both x pot / +grad / +hess x at sources / at targets / at both,
for lfmm2d, cfmm2d and rfmm2d (27 cases each) and hfmm2d (18: pot and grad), every
output against the direct evaluators (themselves equal to the reference's direct
kernels to 1e-12).  FMM at eps = 1e-9; bar 1e-8 relative L2 (1e-7 for Hessians)."""
import numpy as np
import pytest

from fmm2d_amd import fmm2d as F

pytestmark = pytest.mark.gpu

KINDS = {"l": 3, "c": 3, "r": 3, "h": 2}
EPS = 1e-9


def _rel(a, b):
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


@pytest.mark.parametrize("kind", sorted(KINDS))
@pytest.mark.parametrize("ifc,ifd", [(1, 0), (0, 1), (1, 1)])
def test_case_matrix_vs_direct(kind, ifc, ifd):
    rng = np.random.default_rng(5 + 7 * ifc + 3 * ifd)
    ns, nt = 1500, 700
    src = rng.random((2, ns))
    src[:, :400] = 0.3 + 0.01 * rng.standard_normal((2, 400))
    trg = rng.random((2, nt))
    real = kind in ("r", "c")  # Cauchy charges real: see the cfmm2d Re(pot) finding
    ch = (rng.standard_normal(ns) + (0 if real else 1j * rng.standard_normal(ns))) if ifc \
        else None
    ds = (rng.standard_normal(ns) + (0 if kind == "r" else 1j * rng.standard_normal(ns))) \
        if ifd else None
    dv = rng.standard_normal((2, ns)) if (ifd and kind != "c") else None
    zk = 20.0 + 0.0j if kind == "h" else None
    names = {"pot": 0, "grad": 1, "hess": 2}
    checked = 0
    for lvl in range(1, KINDS[kind] + 1):
        for pg, pgt in ((lvl, 0), (0, lvl), (lvl, lvl)):
            res, ier = F._fmm(kind, EPS, zk, src, ch, ds, dv, trg if pgt else None, pg, pgt, 1)
            assert ier == 0, (kind, pg, pgt)
            for where, flag, pts in (("", pg, src), ("targ", pgt, trg)):
                if not flag:
                    assert all(k + where not in res for k in names), (kind, pg, pgt)
                    continue
                ex, _ = F._direct(kind, zk, src, pts, ch, ds, dv, flag, 1, 1e-14)
                for k, j in names.items():
                    key = k + where
                    if j >= flag:
                        assert key not in res, (kind, pg, pgt, key)
                        continue
                    got, ref = res[key], ex[j][0]
                    if kind == "c" and k == "pot":
                        got, ref = got.real, ref.real
                    tol = 1e-7 if k == "hess" else 1e-8
                    assert _rel(got, ref) < tol, (kind, pg, pgt, key, _rel(got, ref))
                    checked += 1
    assert checked == 4 * sum(range(1, KINDS[kind] + 1))
