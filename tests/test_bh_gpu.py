"""Biharmonic and Stokes sums on the GPU (fmm2d_bh2d_direct, fmm2d_st2d_direct,
fmm2d_bhfmm2d, fmm2d_stfmm2d through the Python mirror) against the
reference's outputs in tests/golden/bh_ref.npz (the reference Fortran built
from source by oracle/ref.mk) and against the oracle's direct sums."""
import numpy as np
import pytest

from fmm2d_amd import fmm2d as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a).ravel(), np.asarray(b).ravel()
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("nd", [1, 2])
def test_bh_st_direct_match_reference(golden_dir, nd):
    g = np.load(f"{golden_dir}/bh_ref.npz")
    pre = f"d{nd}_"
    src, trg = g[pre + "src"], g[pre + "trg"]
    ch, dp = g[pre + "charge"], g[pre + "dip"]
    nt = trg.shape[1]
    for ifc, ifd in ((1, 0), (0, 1), (1, 1)):
        for pg in (1, 2):
            c = (ch[0] if nd == 1 else ch) if ifc else None
            d = (dp[0] if nd == 1 else dp) if ifd else None
            out = F.bh2ddir(sources=src, targets=trg, charges=c, dipoles=d, pgt=pg, nd=nd,
                            thresh=1e-14)
            tag = f"{pre}bh_{'c' if ifc else ''}{'d' if ifd else ''}_{pg}_"
            assert out.pottarg.shape == ((nt,) if nd == 1 else (nd, nt))
            assert _rel(out.pottarg, g[tag + "pottarg"]) < 1e-12, tag
            if pg == 2:
                assert out.gradtarg.shape == ((3, nt) if nd == 1 else (nd, 3, nt))
                assert _rel(out.gradtarg, g[tag + "gradtarg"]) < 1e-12, tag
    sk, sl, sv = g[pre + "stoklet"], g[pre + "strslet"], g[pre + "strsvec"]
    for tag, (a, b, c) in (("s", (sk, None, None)), ("st", (sk, sl, sv)), ("t", (None, sl, sv))):
        sq = (lambda x: None if x is None else (x[0] if nd == 1 else x))
        out = F.st2ddir(eps=1e-6, sources=src, stoklet=sq(a), strslet=sq(b), strsvec=sq(c),
                        targets=trg, ifppreg=3, ifppregtarg=3, nd=nd, thresh=1e-14)
        for k in ("pottarg", "pretarg", "gradtarg"):
            assert _rel(getattr(out, k), g[f"{pre}st_{tag}_{k}"]) < 1e-12, (tag, k)
        # at the sources: the self term excluded by thresh
        ex = O.st_direct(nd, src, src, a, b, c, thresh=1e-14)
        for k, kk in (("pot", "pottarg"), ("pre", "pretarg"), ("grad", "gradtarg")):
            assert _rel(getattr(out, k), ex[kk]) < 1e-12, (tag, k)


def _case(g, prefix):
    return {k[len(prefix):]: g[k] for k in g.files if k.startswith(prefix)}


@pytest.mark.parametrize("tag", ["c", "d", "cd"])
@pytest.mark.parametrize("name", ["u6", "c10", "u9nd2"])
def test_bhfmm2d_matches_reference(golden_dir, name, tag):
    """pot / grad at the sources and targets within the requested eps of the
    direct sum, and no further from it than a small multiple of the reference
    FMM's own error."""
    g = np.load(f"{golden_dir}/bh_ref.npz")
    c = _case(g, f"fb_{name}_")
    nd, eps, src, trg = int(c["nd"]), float(c["eps"]), c["src"], c["trg"]
    ch = c["charge"] if "c" in tag else None
    dp = c["dip"] if "d" in tag else None
    sq = (lambda x: None if x is None else (x[0] if nd == 1 else x))
    out = F.bhfmm2d(eps=eps, sources=src, charges=sq(ch), dipoles=sq(dp), targets=trg, pg=2,
                    pgt=2, nd=nd)
    assert out.ier == 0
    exs = O.bh_direct(nd, src, src, ch, dp, 2, thresh=1e-14)
    ext = O.bh_direct(nd, src, trg, ch, dp, 2, thresh=1e-14)
    for mine, ref, ex in ((out.pot, c[f"{tag}_pot"], exs["pottarg"]),
                          (out.grad, c[f"{tag}_grad"], exs["gradtarg"]),
                          (out.pottarg, c[f"{tag}_pottarg"], ext["pottarg"]),
                          (out.gradtarg, c[f"{tag}_gradtarg"], ext["gradtarg"])):
        e_mine, e_ref = _rel(mine, ex), _rel(ref, ex)
        assert e_mine < max(eps, 10 * e_ref), (e_mine, e_ref)


@pytest.mark.parametrize("tag", ["s", "st", "t"])
@pytest.mark.parametrize("name", ["u6", "c10", "u9nd2"])
def test_stfmm2d_matches_reference(golden_dir, name, tag):
    """Velocity, pressure and velocity gradient at the sources and targets within
    eps of the direct sum (and near the reference FMM's own error); the
    reference's constant -sum(charge1) and self term included."""
    g = np.load(f"{golden_dir}/bh_ref.npz")
    c = _case(g, f"fs_{name}_")
    nd, eps, src, trg = int(c["nd"]), float(c["eps"]), c["src"], c["trg"]
    sk = c["stoklet"] if "s" in tag else None
    sl = c["strslet"] if "t" in tag else None
    sv = c["strsvec"] if "t" in tag else None
    sq = (lambda x: None if x is None else (x[0] if nd == 1 else x))
    out = F.stfmm2d(eps=eps, sources=src, stoklet=sq(sk), strslet=sq(sl), strsvec=sq(sv),
                    targets=trg, ifppreg=3, ifppregtarg=3, nd=nd)
    assert out.ier == 0
    exs = O.st_direct(nd, src, src, sk, sl, sv, thresh=1e-14)
    ext = O.st_direct(nd, src, trg, sk, sl, sv, thresh=1e-14)
    for k, ex in (("pot", exs["pottarg"]), ("pre", exs["pretarg"]), ("grad", exs["gradtarg"]),
                  ("pottarg", ext["pottarg"]), ("pretarg", ext["pretarg"]),
                  ("gradtarg", ext["gradtarg"])):
        mine, ref = getattr(out, k), c[f"{tag}_{k}"]
        assert mine.shape == ref.shape, k
        e_mine, e_ref = _rel(mine, ex), _rel(ref, ex)
        assert e_mine < max(eps, 10 * e_ref), (k, e_mine, e_ref)


@pytest.mark.parametrize("eps", [1e-6, 1e-9])
def test_bhfmm2d_large_sampled(eps):
    """N = 2e5 clustered points: the FMM against the GPU direct sum (pinned to the
    reference above) at 200 sampled sources and 200 targets."""
    rng = np.random.default_rng(7)
    n = 200_000
    cen = rng.random((2, 16))
    src = cen[:, rng.integers(0, 16, n)] + 0.03 * rng.standard_normal((2, n))
    trg = rng.uniform(src.min(), src.max(), (2, 200))
    ch = rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))
    dp = rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))
    out = F.bhfmm2d(eps=eps, sources=src, charges=ch, dipoles=dp, targets=trg, pg=2, pgt=2)
    idx = np.sort(rng.choice(n, 200, replace=False))
    exs = F.bh2ddir(sources=src, targets=src[:, idx], charges=ch, dipoles=dp, pgt=2,
                    thresh=1e-14)
    ext = F.bh2ddir(sources=src, targets=trg, charges=ch, dipoles=dp, pgt=2, thresh=1e-14)
    assert _rel(out.pot[idx], exs.pottarg) < eps
    assert _rel(out.grad[:, idx], exs.gradtarg) < eps
    assert _rel(out.pottarg, ext.pottarg) < eps
    assert _rel(out.gradtarg, ext.gradtarg) < eps
