"""Write the golden fixtures under tests/golden/ from the reference library.

Runs in the build container only: it needs oracle/_ref/libfmm2d_ref.so,
compiled from the reference Fortran sources by oracle/ref.mk, and mpmath.
Each fixture holds inputs and the reference's outputs; nothing of the
reference's source is stored.

  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))
from oracle import ref_binding as R  # noqa: E402
sys.path.insert(0, HERE)
import inputs as I  # noqa: E402


def direct_fixtures():
    """Direct kernels c2d/l2d/h2d for all 27 variants each, nd in {1, 3}."""
    rng = np.random.default_rng(11)
    out = {}
    ns, nt = 41, 29
    for nd in (1, 3):
        src = rng.random((2, ns))
        trg = rng.random((2, nt))
        trg[:, 0] = src[:, 3]          # coincident pair: excluded by thresh
        trg[:, 1] = src[:, 5] + 1e-17  # within thresh of a source
        ch = rng.standard_normal((nd, ns)) + 1j * rng.standard_normal((nd, ns))
        ds = rng.standard_normal((nd, ns)) + 1j * rng.standard_normal((nd, ns))
        dv = rng.standard_normal((nd, 2, ns))
        pre = f"nd{nd}_"
        out[pre + "src"] = src
        out[pre + "trg"] = trg
        out[pre + "charge"] = ch
        out[pre + "dipstr"] = ds
        out[pre + "dipvec"] = dv
        for kind in "clh":
            zk = 7.5 + 0.25j if kind == "h" else None
            out[pre + "zk"] = np.array(7.5 + 0.25j)
            for ifc, ifd in ((1, 0), (0, 1), (1, 1)):
                for pgt in (1, 2, 3):
                    r = R.direct(kind, nd, src, trg, charges=ch if ifc else None,
                                 dipstr=ds if ifd else None, dipvec=dv if ifd else None,
                                 pgt=pgt, thresh=1e-14, zk=zk)
                    tag = f"{pre}{kind}_{'c' if ifc else ''}{'d' if ifd else ''}_{pgt}_"
                    for k, v in r.items():
                        out[tag + k] = v
    np.savez_compressed(os.path.join(HERE, "direct_ref.npz"), **out)


def hankel_fixture():
    """hank103 and mpmath values of H0, H1 on a complex grid."""
    import mpmath as mp
    mp.mp.dps = 40
    zs = list(np.logspace(-8, 4, 97))
    for r in (0.3, 1.0, 2.49, 2.51, 3.7, 8.0, 16.9, 17.1, 20.0, 55.0):
        for th in np.linspace(0.0, np.pi, 9):
            zs.append(r * np.exp(1j * th))
    for r in (0.5, 2.0, 5.0, 30.0):
        for th in (-0.05, -0.2):
            zs.append(r * np.exp(1j * th))
    # deeper in the lower half-plane (Im zk < 0): every regime of the evaluator
    for r in (0.4, 1.5, 2.2, 4.0, 6.5, 9.0, 13.0, 16.5, 18.0, 25.0):
        for th in (-0.5, -1.0, -np.pi / 2, -2.0, -2.6, -3.1):
            zs.append(r * np.exp(1j * th))
    z = np.array(zs, dtype=np.complex128)
    h0r, h1r = R.hank103(z)

    def hank(nu, v):
        # mpmath forms H = J + iY, which cancels by ~exp(2 Im z): raise the
        # working precision accordingly
        with mp.workdps(40 + int(abs(v.imag) / 1.1)):
            return complex(mp.hankel1(nu, mp.mpc(v)))

    h0m = np.array([hank(0, v) for v in z])
    h1m = np.array([hank(1, v) for v in z])
    np.savez_compressed(os.path.join(HERE, "hankel.npz"), z=z, h0_ref=h0r, h1_ref=h1r,
                        h0_mp=h0m, h1_mp=h1m)


def point_sets():
    """Named point sets (sources, targets) for the tree and FMM fixtures."""
    rng = np.random.default_rng(5)
    sets = {}
    sets["uniform"] = (rng.random((2, 3000)), rng.random((2, 1500)))
    # clusters deep enough to exercise the level-restriction passes
    c = []
    for cx, cy, s, n in ((0.3, 0.3, 2e-3, 1500), (0.71, 0.64, 1e-2, 1200), (0.5, 0.9, 3e-4, 600)):
        c.append(np.array([[cx], [cy]]) + s * rng.standard_normal((2, n)))
    c.append(rng.random((2, 300)))
    sets["cluster"] = (np.concatenate(c, axis=1), rng.random((2, 700)))
    t = rng.uniform(0, 2 * np.pi, 2500)
    r = 1 + 0.3 * np.cos(5 * t)
    curve = np.stack([r * np.cos(t), r * np.sin(t)]) + 1e-4 * rng.standard_normal((2, 2500))
    sets["curve"] = (curve, None)
    # exact duplicates (more than ndiv of them): refinement runs to nlmax = 51
    d = [np.full((2, 12), 0.3141592653589793), 0.77 + 1e-11 * rng.standard_normal((2, 30)),
         rng.random((2, 300))]
    sets["deep"] = (np.concatenate(d, axis=1), None)
    return sets


def tree_fixtures():
    out = {}
    for name, (src, trg) in point_sets().items():
        for ndiv in (5, 15):
            tr = R.tree(src, trg, ndiv)
            li = R.lists(tr)
            pre = f"{name}_{ndiv}_"
            for k in ("nlevels", "nboxes", "laddr", "ilevel", "iparent", "nchild", "ichild",
                      "ncoll", "coll", "centers", "boxsize", "isrc", "isrcse", "itarg",
                      "itargse"):
                out[pre + k] = np.asarray(tr[k])
            for k in range(1, 5):
                n = li[f"nlist{k}"]
                L = li[f"list{k}"].copy()
                for b in range(tr["nboxes"]):
                    row = np.sort(L[: n[b], b])
                    L[:, b] = 0
                    L[: n[b], b] = row
                out[pre + f"nlist{k}"] = n
                out[pre + f"list{k}"] = L
        out[name + "_src"] = src
        if trg is not None:
            out[name + "_trg"] = trg
    np.savez_compressed(os.path.join(HERE, "tree_ref.npz"), **out)


def fmm_fixtures():
    """cfmm2d / lfmm2d / hfmm2d outputs of the reference on small problems."""
    rng = np.random.default_rng(7)
    out = {}
    ns, nt = 1200, 700
    src = rng.random((2, ns))
    src[:, :300] = 0.2 + 3e-3 * rng.standard_normal((2, 300))
    trg = rng.random((2, nt))
    out["src"], out["trg"] = src, trg
    for nd in (1, 2):
        ch = rng.standard_normal((nd, ns)) + 1j * rng.standard_normal((nd, ns))
        ds = rng.standard_normal((nd, ns)) + 1j * rng.standard_normal((nd, ns))
        dv = rng.standard_normal((nd, 2, ns))
        out[f"nd{nd}_charge"], out[f"nd{nd}_dipstr"], out[f"nd{nd}_dipvec"] = ch, ds, dv
        cases = [(1, 0, 1, 0), (1, 0, 2, 2), (0, 1, 0, 2), (1, 1, 3, 3)] if nd == 1 else \
                [(1, 1, 2, 2)]
        for kind in ("c", "l", "h"):
            for eps in ((1e-6, 1e-12) if kind != "h" else (1e-6, 1e-10)):
                for ifc, ifd, pg, pgt in cases:
                    if kind == "h" and (pg > 2 or pgt > 2):
                        continue
                    kw = dict(nd=nd, eps=eps, sources=src, charges=ch if ifc else None,
                              dipstr=ds if ifd else None, dipvec=dv if ifd else None,
                              targets=trg if pgt else None, ifpgh=pg, ifpghtarg=pgt)
                    if kind == "c":
                        kw.pop("dipvec")
                        r = R.cfmm2d(**kw)
                    elif kind == "l":
                        r = R.lfmm2d(**kw)
                    else:
                        r = R.hfmm2d(zk=12.0 + 0.0j, **kw)
                    tag = f"nd{nd}_{kind}_{eps:.0e}_{ifc}{ifd}{pg}{pgt}_"
                    for k, v in r.items():
                        if k in ("pot", "grad", "hess", "pottarg", "gradtarg", "hesstarg"):
                            out[tag + k] = v
                    out[tag + "ier"] = np.array(r["ier"])
    np.savez_compressed(os.path.join(HERE, "fmm_ref.npz"), **out)


# (zk, eps, nd, ifc, ifd, pg, pgt): hessians; complex zk; Im(zk) large enough
# that the coarse levels drop out (hfmm2d.f:711); zk = 600, where levels 2 and
# 3 are in the reference's high-frequency (FFT diagonal form) regime.
HFMM_CASES = [
    (12.0 + 0.0j, 1e-6, 1, 1, 1, 3, 3),
    (12.0 + 0.0j, 1e-10, 1, 1, 1, 3, 3),
    (12.0 + 3.0j, 1e-9, 1, 1, 0, 2, 2),
    (12.0 + 3.0j, 1e-9, 1, 0, 1, 3, 3),
    (5.0 + 100.0j, 1e-9, 1, 1, 1, 2, 2),
    (600.0 + 0.0j, 1e-6, 1, 1, 0, 2, 2),
    (600.0 + 0.0j, 1e-9, 2, 1, 1, 3, 3),
]


def hfmm_fixtures():
    """hfmm2d outputs of the reference beyond fmm_ref.npz (same points)."""
    base = np.load(os.path.join(HERE, "fmm_ref.npz"))
    src, trg = base["src"], base["trg"]
    out = {"src": src, "trg": trg}
    for i, (zk, eps, nd, ifc, ifd, pg, pgt) in enumerate(HFMM_CASES):
        ch, ds, dv = base[f"nd{nd}_charge"], base[f"nd{nd}_dipstr"], base[f"nd{nd}_dipvec"]
        r = R.hfmm2d(zk=zk, nd=nd, eps=eps, sources=src, charges=ch if ifc else None,
                     dipstr=ds if ifd else None, dipvec=dv if ifd else None,
                     targets=trg if pgt else None, ifpgh=pg, ifpghtarg=pgt)
        for k, v in r.items():
            if k in ("pot", "grad", "hess", "pottarg", "gradtarg", "hesstarg"):
                out[f"c{i}_{k}"] = v
        out[f"c{i}_ier"] = np.array(r["ier"])
    np.savez_compressed(os.path.join(HERE, "hfmm_ref.npz"), **out)


# (eps, nd, ifc, ifd, pg, pgt) for lfmm2d with real charges and dipole
# strengths (zero imaginary parts)
LFMM_REAL_CASES = [
    (1e-6, 1, 1, 0, 2, 2),
    (1e-12, 1, 1, 1, 3, 3),
    (1e-9, 1, 0, 1, 1, 2),
    (1e-9, 2, 1, 1, 2, 0),
]


def lfmm_real_fixtures():
    """lfmm2d outputs of the reference for real densities (same points)."""
    base = np.load(os.path.join(HERE, "fmm_ref.npz"))
    src, trg = base["src"], base["trg"]
    out = {"src": src, "trg": trg}
    for i, (eps, nd, ifc, ifd, pg, pgt) in enumerate(LFMM_REAL_CASES):
        ch = base[f"nd{nd}_charge"].real + 0j
        ds = base[f"nd{nd}_dipstr"].real + 0j
        dv = base[f"nd{nd}_dipvec"]
        r = R.lfmm2d(nd=nd, eps=eps, sources=src, charges=ch if ifc else None,
                     dipstr=ds if ifd else None, dipvec=dv if ifd else None,
                     targets=trg if pgt else None, ifpgh=pg, ifpghtarg=pgt)
        for k, v in r.items():
            if k in ("pot", "grad", "hess", "pottarg", "gradtarg", "hesstarg"):
                out[f"c{i}_{k}"] = v
        out[f"c{i}_ier"] = np.array(r["ier"])
    np.savez_compressed(os.path.join(HERE, "lfmm_real_ref.npz"), **out)


# (eps, nd, ifc, ifd, pg, pgt) for rfmm2d (real densities and outputs)
RFMM_CASES = [
    (1e-6, 1, 1, 0, 2, 2),
    (1e-12, 1, 1, 1, 3, 3),
    (1e-9, 1, 0, 1, 1, 2),
    (1e-9, 2, 1, 1, 3, 0),
]


def rfmm_fixtures():
    """rfmm2d and r2d_direct outputs of the reference (points of fmm_ref.npz)."""
    base = np.load(os.path.join(HERE, "fmm_ref.npz"))
    src, trg = base["src"], base["trg"]
    out = {"src": src, "trg": trg}
    for i, (eps, nd, ifc, ifd, pg, pgt) in enumerate(RFMM_CASES):
        ch = base[f"nd{nd}_charge"].real.copy()
        ds = base[f"nd{nd}_dipstr"].real.copy()
        dv = base[f"nd{nd}_dipvec"]
        r = R.rfmm2d(nd=nd, eps=eps, sources=src, charges=ch if ifc else None,
                     dipstr=ds if ifd else None, dipvec=dv if ifd else None,
                     targets=trg if pgt else None, ifpgh=pg, ifpghtarg=pgt)
        for k, v in r.items():
            if k in ("pot", "grad", "hess", "pottarg", "gradtarg", "hesstarg"):
                out[f"c{i}_{k}"] = v
        out[f"c{i}_ier"] = np.array(r["ier"])
    # r2d_direct: all 9 variants, nd = 1 and 3, on small point sets
    rng = np.random.default_rng(17)
    for nd in (1, 3):
        s_ = rng.random((2, 37))
        t_ = rng.random((2, 23))
        t_[:, 0] = s_[:, 2]
        ch = rng.standard_normal((nd, 37))
        ds = rng.standard_normal((nd, 37))
        dv = rng.standard_normal((nd, 2, 37))
        out[f"dir{nd}_src"], out[f"dir{nd}_trg"] = s_, t_
        out[f"dir{nd}_charge"], out[f"dir{nd}_dipstr"], out[f"dir{nd}_dipvec"] = ch, ds, dv
        for ifc, ifd in ((1, 0), (0, 1), (1, 1)):
            for pgt in (1, 2, 3):
                r = R.direct("r", nd, s_, t_, charges=ch if ifc else None,
                             dipstr=ds if ifd else None, dipvec=dv if ifd else None, pgt=pgt,
                             thresh=1e-14)
                for k, v in r.items():
                    out[f"dir{nd}_{ifc}{ifd}{pgt}_{k}"] = v
    np.savez_compressed(os.path.join(HERE, "rfmm_ref.npz"), **out)


# The bench workload's point distribution at a size the reference runs in
# seconds: lfmm2d with real charges at eps 1e-6, default leaf capacity and
# leaf capacity 37 (lfmm2d_ndiv, laplace/lfmm2d_ndiv.f:29), and rfmm2d.
HEADLINE = dict(n=120000, seed=31, nsample=3000, sample_seed=32, eps=1e-6, ndiv=37)


def headline_fixtures():
    """Reference lfmm2d / lfmm2d_ndiv / rfmm2d / rfmm2d_ndiv outputs (pot, grad at
    the sources, pg = 2) and the reference direct sum, at sampled sources."""
    h = HEADLINE
    src, q = I.clustered(h["n"], h["seed"])
    idx = np.sort(np.random.default_rng(h["sample_seed"]).choice(h["n"], h["nsample"],
                                                                 replace=False))
    out = {"input_sha256": np.array(I.digest(src, q)), "idx": idx}
    eps, nd = h["eps"], h["ndiv"]
    runs = {
        "l_default": R.lfmm2d(nd=1, eps=eps, sources=src, charges=q + 0j, ifpgh=2),
        "l_ndiv": R.lfmm2d(nd=1, eps=eps, sources=src, charges=q + 0j, ifpgh=2, ndiv=nd),
        "r_default": R.rfmm2d(nd=1, eps=eps, sources=src, charges=q, ifpgh=2),
        "r_ndiv": R.rfmm2d(nd=1, eps=eps, sources=src, charges=q, ifpgh=2, ndiv=nd),
    }
    for tag, r in runs.items():
        assert r["ier"] == 0, (tag, r["ier"])
        out[tag + "_pot"] = r["pot"][0, idx]
        out[tag + "_grad"] = r["grad"][0][:, idx]
    d = R.direct("l", 1, src, src[:, idx], charges=q + 0j, pgt=2, thresh=1e-16)
    out["direct_pot"] = d["pottarg"][0]
    out["direct_grad"] = d["gradtarg"][0]
    np.savez_compressed(os.path.join(HERE, "headline_ref.npz"), **out)


# The Helmholtz workload (hfmm2d, zk = 100, eps 1e-9, charges and dipoles,
# pot+grad at separate targets) at 1e4 sources + 1e4 targets.
HELM_C2 = dict(n=10000, seed=2, zk=100.0, eps=1e-9, nsample=2000, sample_seed=3)


def helm_c2_fixture():
    """Reference hfmm2d outputs (pottarg, gradtarg at sampled targets)."""
    h = HELM_C2
    src, trg, ch, ds, dv = I.helmholtz_st(h["n"], h["seed"])
    idx = np.sort(np.random.default_rng(h["sample_seed"]).choice(h["n"], h["nsample"],
                                                                 replace=False))
    r = R.hfmm2d(zk=h["zk"] + 0j, nd=1, eps=h["eps"], sources=src, charges=ch[None],
                 dipstr=ds[None], dipvec=dv[None], targets=trg, ifpgh=0, ifpghtarg=2)
    assert r["ier"] == 0
    out = {"input_sha256": np.array(I.digest(src, trg, ch, ds, dv)), "idx": idx,
           "pottarg": r["pottarg"][0, idx], "gradtarg": r["gradtarg"][0][:, idx]}
    np.savez_compressed(os.path.join(HERE, "helm_c2_ref.npz"), **out)


def mps_cases():
    """Inputs of the hfmm2d_mps fixtures: per case (name, nd, eps, zk, centers
    (2,n), rscales, nterms, mpole, impole (0-based), sources (2,n) and charges
    (nd,n) of the point charges the expansions were formed from).

    grid: the reference's own test (test/helmholtz/test_hfmm2d_mps.f90): a 10x10
    grid of sources h = 1/11 apart, centres shifted by h/100 in x, zk = 2 pi,
    order 30, rscale = |zk| h/100, eps 0.5e-9 -- here with random complex charges
    at every source.  cluster: 300 centres (two Gaussian clusters and a uniform
    background, centres at least 0.01 apart), nd = 2, zk = 40 + 0.5i, orders
    6..20, one source per centre at a random offset below 1e-3, eps 1e-8,
    expansions stored in shuffled order with gaps between them."""
    rng = np.random.default_rng(11)
    cases = []
    n1 = 10
    h = 1.0 / (n1 + 1)
    src = np.array([[h * i, h * j] for i in range(1, n1 + 1) for j in range(1, n1 + 1)]).T
    cen = src.copy()
    cen[0] += h / 100
    zk = 2 * np.pi + 0j
    n = src.shape[1]
    rs = np.full(n, min(abs(zk) * h / 100, 1.0))
    nt = np.full(n, 30, dtype=np.int32)
    ch = (rng.standard_normal((1, n)) + 1j * rng.standard_normal((1, n)))
    cases.append(("grid", 1, 0.5e-9, zk, cen, rs, nt, ch, src))
    n = 300
    a = 0.3 + 0.05 * rng.standard_normal((2, 4000))
    b = np.array([[0.7], [0.6]]) + 0.08 * rng.standard_normal((2, 4000))
    c = rng.random((2, 2000))
    pool = np.concatenate([a, b, c], axis=1)[:, rng.permutation(10000)]
    keep = []
    for k in range(pool.shape[1]):  # centres at least 0.01 apart
        if all(np.hypot(*(pool[:, k] - pool[:, j])) >= 0.01 for j in keep):
            keep.append(k)
        if len(keep) == n:
            break
    cen = pool[:, keep]
    off = rng.standard_normal((2, n))
    off *= (1e-3 * rng.random(n)) / np.linalg.norm(off, axis=0)
    src = cen + off
    zk = 40.0 + 0.5j
    rs = np.minimum(abs(zk) * np.linalg.norm(off, axis=0) * 1.5, 1.0)
    nt = rng.integers(6, 21, n).astype(np.int32)
    ch = (rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n)))
    cases.append(("cluster", 2, 1e-8, zk, cen, rs, nt, ch, src))
    out = []
    for name, nd, eps, zk, cen, rs, nt, ch, src in cases:
        n = cen.shape[1]
        blocks = []
        for i in range(n):
            per = [R.h2dformmpc(zk, rs[i], src[:, i:i + 1], ch[d, i:i + 1], cen[:, i], int(nt[i]))
                   for d in range(nd)]
            blocks.append(np.stack(per, axis=1).ravel())  # (2m+1, nd) -> density fastest
        if name == "cluster":
            order = rng.permutation(n)
            gaps = rng.integers(0, 3, n)
        else:
            order = np.arange(n)
            gaps = np.zeros(n, dtype=int)
        imp = np.zeros(n, dtype=np.int64)
        pos = 0
        for k in order:
            pos += gaps[k]
            imp[k] = pos
            pos += blocks[k].size
        mp = np.zeros(pos + 2, np.complex128)
        for k in range(n):
            mp[imp[k]:imp[k] + blocks[k].size] = blocks[k]
        out.append((name, nd, eps, zk, cen, rs, nt, mp, imp, src, ch))
    return out


def mps_fixtures():
    """Reference hfmm2d_mps outputs (FMM, and hfmm2d_mps_ndiv with ndiv = n:
    every pair directly) for mps_cases(), with the inputs (the expansions come
    from the reference's h2dformmpc, so the test reads them from here)."""
    out = {}
    for name, nd, eps, zk, cen, rs, nt, mp, imp, src, ch in mps_cases():
        n = cen.shape[1]
        loc, ier = R.hfmm2d_mps(nd, eps, zk, cen, rs, nt, mp, imp + 1)
        assert ier == 0
        locd, ier = R.hfmm2d_mps(nd, eps, zk, cen, rs, nt, mp, imp + 1, ndiv=n)
        assert ier == 0
        out[name + "_input_sha256"] = np.array(I.digest(cen, rs, nt, mp, imp))
        for key, val in dict(nd=nd, eps=eps, zk=zk, centers=cen, rscales=rs, nterms=nt,
                             mpole=mp, impole=imp, sources=src, charges=ch).items():
            out[f"{name}_{key}"] = np.asarray(val)
        out[name + "_local"] = loc
        out[name + "_local_direct"] = locd
    np.savez_compressed(os.path.join(HERE, "mps_ref.npz"), **out)


def bh_cases():
    """(name, nd, eps, sources, targets, charges, dipoles) for the biharmonic FMM;
    (name, nd, eps, sources, targets, stoklet, strslet, strsvec) for Stokes."""
    rng = np.random.default_rng(23)
    bh, st = [], []
    for name, nd, eps, n, dist in (("u6", 1, 1e-6, 1200, "u"), ("c10", 1, 1e-10, 1200, "c"),
                                   ("u9nd2", 2, 1e-9, 800, "u")):
        if dist == "u":
            src = rng.random((2, n))
        else:  # 8 Gaussian clusters plus 10% uniform, offset from the origin
            cen = rng.uniform(2.0, 3.0, (2, 8))
            src = cen[:, rng.integers(0, 8, n)] + 0.02 * rng.standard_normal((2, n))
            src[:, : n // 10] = rng.uniform(2.0, 3.0, (2, n // 10))
        trg = rng.uniform(src.min(), src.max(), (2, 400))
        ch = rng.standard_normal((nd, 2, n)) + 1j * rng.standard_normal((nd, 2, n))
        dp = rng.standard_normal((nd, 3, n)) + 1j * rng.standard_normal((nd, 3, n))
        bh.append((name, nd, eps, src, trg, ch, dp))
        sk = rng.standard_normal((nd, 2, n))
        sl = rng.standard_normal((nd, 2, n))
        sv = rng.standard_normal((nd, 2, n))
        st.append((name, nd, eps, src, trg, sk, sl, sv))
    return bh, st


def bh_fixtures():
    """Reference biharmonic and Stokes outputs: the six bh2d_direct variants and
    the three st2ddirect variants (nd 1 and 2; a coincident target), and bhfmm2d /
    stfmm2d on bh_cases() (charges, dipoles or both; Stokeslets, stresslets or
    both), with the inputs."""
    out = {}
    rng = np.random.default_rng(29)
    ns, nt = 53, 37
    for nd in (1, 2):
        src = rng.random((2, ns))
        trg = rng.random((2, nt))
        trg[:, 0] = src[:, 4]  # coincident: excluded by thresh
        ch = rng.standard_normal((nd, 2, ns)) + 1j * rng.standard_normal((nd, 2, ns))
        dp = rng.standard_normal((nd, 3, ns)) + 1j * rng.standard_normal((nd, 3, ns))
        sk, sl, sv = (rng.standard_normal((nd, 2, ns)) for _ in range(3))
        pre = f"d{nd}_"
        for key, val in dict(src=src, trg=trg, charge=ch, dip=dp, stoklet=sk, strslet=sl,
                             strsvec=sv).items():
            out[pre + key] = val
        for ifc, ifd in ((1, 0), (0, 1), (1, 1)):
            for pg in (1, 2):
                r = R.bh_direct(nd, src, trg, ch if ifc else None, dp if ifd else None, pg,
                                thresh=1e-14)
                tag = f"{pre}bh_{'c' if ifc else ''}{'d' if ifd else ''}_{pg}_"
                for k, v in r.items():
                    out[tag + k] = v
        for tag, args in (("s", (sk, None, None)), ("st", (sk, sl, sv)), ("t", (None, sl, sv))):
            r = R.st_direct(nd, src, trg, *args, thresh=1e-14)
            for k, v in r.items():
                out[f"{pre}st_{tag}_{k}"] = v
    bh, st = bh_cases()
    for name, nd, eps, src, trg, ch, dp in bh:
        out[f"fb_{name}_input_sha256"] = np.array(I.digest(src, trg, ch, dp))
        for key, val in dict(nd=nd, eps=eps, src=src, trg=trg, charge=ch, dip=dp).items():
            out[f"fb_{name}_{key}"] = np.asarray(val)
        for tag, c, d in (("c", ch, None), ("d", None, dp), ("cd", ch, dp)):
            r = R.bhfmm2d(nd, eps, src, c, d, trg, 2, 2)
            assert r["ier"] == 0
            for k in ("pot", "grad", "pottarg", "gradtarg"):
                out[f"fb_{name}_{tag}_{k}"] = r[k]
    for name, nd, eps, src, trg, sk, sl, sv in st:
        out[f"fs_{name}_input_sha256"] = np.array(I.digest(src, trg, sk, sl, sv))
        for key, val in dict(nd=nd, eps=eps, src=src, trg=trg, stoklet=sk, strslet=sl,
                             strsvec=sv).items():
            out[f"fs_{name}_{key}"] = np.asarray(val)
        for tag, args in (("s", (sk, None, None)), ("st", (sk, sl, sv)), ("t", (None, sl, sv))):
            r = R.stfmm2d(nd, eps, src, *args, trg, 3, 3)
            assert r["ier"] == 0
            for k in ("pot", "pre", "grad", "pottarg", "pretarg", "gradtarg"):
                out[f"fs_{name}_{tag}_{k}"] = r[k]
    np.savez_compressed(os.path.join(HERE, "bh_ref.npz"), **out)


if __name__ == "__main__":
    which = sys.argv[1:] or ["direct", "hankel", "tree", "fmm", "hfmm", "lreal", "rfmm",
                             "headline", "helmc2", "mps", "bh"]
    if "direct" in which:
        direct_fixtures()
    if "hankel" in which:
        hankel_fixture()
    if "tree" in which:
        tree_fixtures()
    if "fmm" in which:
        fmm_fixtures()
    if "hfmm" in which:
        hfmm_fixtures()
    if "lreal" in which:
        lfmm_real_fixtures()
    if "rfmm" in which:
        rfmm_fixtures()
    if "headline" in which:
        headline_fixtures()
    if "helmc2" in which:
        helm_c2_fixture()
    if "mps" in which:
        mps_fixtures()
    if "bh" in which:
        bh_fixtures()
    for f in sorted(os.listdir(HERE)):
        print(f, os.path.getsize(os.path.join(HERE, f)))
