"""CPU tests: the C oracle pinned against the reference's own outputs and
against closed forms, so that it can serve as the GPU library's checker.

Pins (see oracle/oracle.h): the reference fixtures in tests/golden (generated
by tests/golden/make_golden.py from the reference Fortran built by
oracle/ref.mk), mpmath Hankel values, single-source closed forms, the
Laplace / Helmholtz equations by finite differences and the Wronskian.
"""
import numpy as np
import pytest

from oracle import oracle as O


def _rel(a, b):
    return np.linalg.norm((np.asarray(a) - np.asarray(b)).ravel()) / np.linalg.norm(np.asarray(b).ravel())


@pytest.mark.parametrize("nd", [1, 3])
@pytest.mark.parametrize("kind", ["c", "l", "h"])
def test_oracle_direct_matches_reference(golden_dir, kind, nd):
    g = np.load(f"{golden_dir}/direct_ref.npz")
    pre = f"nd{nd}_"
    src, trg = g[pre + "src"], g[pre + "trg"]
    ch, ds, dv = g[pre + "charge"], g[pre + "dipstr"], g[pre + "dipvec"]
    zk = complex(g[pre + "zk"])
    for ifc, ifd in ((1, 0), (0, 1), (1, 1)):
        for pgt in (1, 2, 3):
            tag = f"{pre}{kind}_{'c' if ifc else ''}{'d' if ifd else ''}_{pgt}_"
            r = O.direct(kind, nd, src, trg, ch if ifc else None, ds if ifd else None,
                         dv if ifd else None, pgt, 1e-14, zk)
            for k, v in r.items():
                assert _rel(v, g[tag + k]) < 1e-13, (tag, k)


@pytest.mark.parametrize("ndiv", [5, 15])
@pytest.mark.parametrize("name", ["uniform", "cluster", "curve", "deep"])
def test_oracle_tree_bit_exact(golden_dir, name, ndiv):
    g = np.load(f"{golden_dir}/tree_ref.npz")
    src = g[name + "_src"]
    trg = g[name + "_trg"] if name + "_trg" in g.files else None
    t = O.tree(src, trg, ndiv)
    pre = f"{name}_{ndiv}_"
    assert t["nlevels"] == int(g[pre + "nlevels"]) and t["nboxes"] == int(g[pre + "nboxes"])
    keys = ["laddr", "ilevel", "iparent", "nchild", "ichild", "ncoll", "coll", "centers",
            "boxsize", "isrc", "isrcse"] + (["itarg", "itargse"] if trg is not None else [])
    for k in keys:
        assert np.array_equal(t[k], g[pre + k]), k
    for k in range(1, 5):
        L, n = t[f"list{k}"].copy(), t[f"nlist{k}"]
        for b in range(t["nboxes"]):
            L[: n[b], b] = np.sort(L[: n[b], b])
            L[n[b]:, b] = 0
        assert np.array_equal(n, g[pre + f"nlist{k}"]) and np.array_equal(L, g[pre + f"list{k}"])


def test_oracle_tree_invariants():
    """Properties the reference's tree must have, independent of fixtures."""
    rng = np.random.default_rng(9)
    src = np.concatenate([0.4 + 1e-3 * rng.standard_normal((2, 800)), rng.random((2, 400))], axis=1)
    t = O.tree(src, None, 10)
    nb, isrc, se = t["nboxes"], t["isrc"], t["isrcse"]
    assert sorted(isrc.tolist()) == list(range(1, src.shape[1] + 1))  # a permutation
    leaves = np.where(t["nchild"] == 0)[0]
    cnt = se[1, leaves] - se[0, leaves] + 1
    assert cnt.sum() == src.shape[1]
    assert np.all((cnt <= 10) | (t["ilevel"][leaves] == t["nlevels"]))
    # points lie inside their leaf and leaf points keep input order (stable sort)
    bs = t["boxsize"]
    for b in leaves:
        idx = isrc[se[0, b] - 1: se[1, b]] - 1
        assert np.all(np.diff(idx) > 0)
        c, h = t["centers"][:, b], bs[t["ilevel"][b]] / 2
        assert np.all(np.abs(src[:, idx] - c[:, None]) <= h * (1 + 1e-12))
    # level restriction: adjacent leaves differ by at most one level (list 1)
    for b in leaves:
        for j in t["list1"][: t["nlist1"][b], b] - 1:
            assert abs(int(t["ilevel"][j]) - int(t["ilevel"][b])) <= 1


def test_oracle_hankel_vs_mpmath(golden_dir):
    g = np.load(f"{golden_dir}/hankel.npz")
    z = g["z"]
    dom = (z.imag >= 0) & (z.imag <= 3.0)
    h0, h1 = O.hank01(z[dom])
    assert _rel(h0, g["h0_mp"][dom]) < 1e-13
    e0 = np.abs(h0 - g["h0_mp"][dom]) / np.abs(g["h0_mp"][dom])
    e1 = np.abs(h1 - g["h1_mp"][dom]) / np.abs(g["h1_mp"][dom])
    assert e0.max() < 5e-13 and e1.max() < 5e-13, (e0.max(), e1.max())


def test_oracle_hankel_wronskian():
    x = np.linspace(0.05, 60.0, 400)
    h0, h1 = O.hank01(x)
    # J1 Y0 - J0 Y1 = 2 / (pi x)
    w = h1.real * h0.imag - h0.real * h1.imag
    assert np.max(np.abs(w * np.pi * x / 2 - 1)) < 1e-12


def test_oracle_direct_closed_forms():
    s = np.array([[0.3], [0.2]])
    t = np.array([[1.1], [-0.4]])
    c = np.array([0.7 - 0.2j])
    z = complex(t[0, 0] - s[0, 0], t[1, 0] - s[1, 0])
    r = O.direct("l", 1, s, t, charges=c, pgt=3)
    assert abs(r["pottarg"][0, 0] - c[0] * np.log(abs(z))) < 1e-15
    g = r["gradtarg"][0, :, 0]
    assert np.allclose(g, c[0] * np.array([z.real, z.imag]) / abs(z) ** 2, rtol=1e-14)
    rc = O.direct("c", 1, s, t, charges=c, dipstr=c, pgt=3)
    assert abs(rc["gradtarg"][0, 0] - c[0] / z + c[0] / z ** 2) < 1e-14
    assert abs(rc["hesstarg"][0, 0] - (-c[0] / z ** 2 + 2 * c[0] / z ** 3)) < 1e-14


def test_oracle_pde_by_finite_differences():
    """Laplace potentials are harmonic; Helmholtz ones satisfy lap u + k^2 u = 0."""
    rng = np.random.default_rng(2)
    s = rng.random((2, 5))
    ch = rng.standard_normal(5) + 1j * rng.standard_normal(5)
    ds = rng.standard_normal(5) + 1j * rng.standard_normal(5)
    dv = rng.standard_normal((2, 5))
    x0, h = np.array([2.0, 1.7]), 1e-3
    pts = np.array([x0, x0 + [h, 0], x0 - [h, 0], x0 + [0, h], x0 - [0, h]]).T
    for kind, zk in (("l", 0.0), ("h", 3.0)):
        u = O.direct(kind, 1, s, pts, charges=ch, dipstr=ds, dipvec=dv, pgt=3, zk=zk)
        p = u["pottarg"][0]
        lap = (p[1] + p[2] + p[3] + p[4] - 4 * p[0]) / h ** 2
        assert abs(lap + zk ** 2 * p[0]) < 1e-5 * max(1.0, abs(p[0])), kind
        # hessian trace agrees with the same Laplacian
        hs = u["hesstarg"][0, :, 0]
        assert abs(hs[0] + hs[2] + zk ** 2 * p[0]) < 1e-10 * max(1.0, abs(p[0]))


def test_oracle_rules_match_reference():
    ref = pytest.importorskip("oracle.ref_binding")
    if not ref.available():
        pytest.skip("reference library not built here")
    for eps in (0.6, 1e-2, 1e-3, 1e-6, 5e-7, 1e-9, 1e-12, 1e-15, 1e-17):
        assert O.l2dterms(eps) == ref.l2dterms(eps)[0]
        assert O.lndiv2d(eps, 100, 50) == ref.ndiv(eps, 100, 50)[0]


def test_oracle_detects_planted_mistakes(golden_dir):
    """The fixture comparisons are sharp enough to catch a plausible slip: a
    sign error in a dipole term or a transposed coordinate changes the result
    by far more than the tolerance used above."""
    g = np.load(f"{golden_dir}/direct_ref.npz")
    src, trg = g["nd1_src"], g["nd1_trg"]
    ds, dv = g["nd1_dipstr"], g["nd1_dipvec"]
    ref = g["nd1_l_d_2_gradtarg"]
    good = O.direct("l", 1, src, trg, None, ds, dv, 2, 1e-14)["gradtarg"]
    assert _rel(good, ref) < 1e-13
    flipped = O.direct("l", 1, src, trg, None, -ds, dv, 2, 1e-14)["gradtarg"]
    swapped = O.direct("l", 1, src[::-1], trg[::-1], None, ds, dv, 2, 1e-14)["gradtarg"]
    assert _rel(flipped, ref) > 1.0 and _rel(swapped, ref) > 1e-2


# ---------------------------------------------------------------- proposer network
def test_mlp_oracle_gradient_by_finite_differences():
    """oracle/mlp_ref.py: the analytic gradient of the weighted squared-error
    loss equals central differences of the loss (float64)."""
    from oracle import mlp_ref as M
    nin, nout, batch = 5, 3, 7
    rng = np.random.default_rng(1)
    p = M.init(nin, nout, seed=2) + 0.05 * rng.standard_normal(M.layout(nin, nout)[1])
    x = rng.standard_normal((batch, nin))
    t = rng.standard_normal((batch, nout))
    w = rng.uniform(0.5, 2.0, (batch, nout))
    loss, g = M.loss_grad(p, x, t, w, nin, nout)
    idx = rng.choice(len(p), 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (M.loss_grad(p + e, x, t, w, nin, nout)[0] -
              M.loss_grad(p - e, x, t, w, nin, nout)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_mlp_oracle_closed_forms():
    """Zero hidden weights: the output is b3 + W3 tanh(b2) independent of x;
    one Adam step from zero moments moves each parameter by lr sign(g)."""
    from oracle import mlp_ref as M
    nin, nout = 4, 2
    p = np.zeros(M.layout(nin, nout)[1])
    W1, b1, W2, b2, W3, b3 = M.unpack(p, nin, nout)
    b2[:] = 0.3
    W3[:] = 0.5
    b3[:] = [1.0, -1.0]
    y, _ = M.forward(p, np.random.default_rng(0).standard_normal((3, nin)), nin, nout)
    expect = np.array([1.0, -1.0]) + 0.5 * 64 * np.tanh(0.3)
    assert np.allclose(y, expect[None, :], rtol=1e-14)
    g = np.array([2.0, -3.0, 0.5])
    p2, m, v = M.adam(np.zeros(3), g, np.zeros(3), np.zeros(3), 0.1, 0.9, 0.999, 0.0, 1)
    assert np.allclose(p2, -0.1 * np.sign(g), rtol=1e-12)


def test_leaf_histogram_restatements_agree():
    """The key-window restatement of the proposer's leaf histogram (the
    device's algorithm) equals box-by-box subdivision, with and without
    targets, on uniform, clustered and curve sets and leaf capacities 1..40;
    a planted off-by-one window fails."""
    from oracle import leaf_hist_ref as H
    rng = np.random.default_rng(0)
    for trial in range(18):
        ns = int(rng.integers(1, 2500))
        nt = int(rng.integers(0, 400)) if trial % 2 else 0
        if trial % 3 == 0:
            src = rng.random((2, ns))
        elif trial % 3 == 1:
            src = 0.5 + 0.01 * rng.standard_normal((2, ns))
        else:
            t = rng.uniform(0, 2 * np.pi, ns)
            src = np.stack([np.cos(t), np.sin(t)]) + 1e-4 * rng.standard_normal((2, ns))
        trg = rng.random((2, nt)) if nt else None
        ndiv = int(rng.integers(1, 41))
        assert np.array_equal(H.windows(src, trg, ndiv), H.brute(src, trg, ndiv)), trial
    # a window one source short (ndiv - 1) refines boxes of exactly ndiv sources
    src = rng.random((2, 3000))
    assert not np.array_equal(H.windows(src, None, 14), H.brute(src, None, 15))


# ---------------------------------------------------------------- biharmonic / Stokes
@pytest.mark.parametrize("nd", [1, 2])
def test_oracle_bh_st_direct_match_reference(golden_dir, nd):
    """orc_bh_direct / orc_st_direct against the reference's bh2d_direct* (all six
    variants) and st2ddirect* (Stokeslet, both, stresslet), with a coincident
    target excluded by thresh."""
    g = np.load(f"{golden_dir}/bh_ref.npz")
    pre = f"d{nd}_"
    src, trg = g[pre + "src"], g[pre + "trg"]
    ch, dp = g[pre + "charge"], g[pre + "dip"]
    for ifc, ifd in ((1, 0), (0, 1), (1, 1)):
        for pg in (1, 2):
            r = O.bh_direct(nd, src, trg, ch if ifc else None, dp if ifd else None, pg, 1e-14)
            tag = f"{pre}bh_{'c' if ifc else ''}{'d' if ifd else ''}_{pg}_"
            for k, v in r.items():
                assert _rel(v, g[tag + k]) < 1e-13, (tag, k)
    sk, sl, sv = g[pre + "stoklet"], g[pre + "strslet"], g[pre + "strsvec"]
    for tag, args in (("s", (sk, None, None)), ("st", (sk, sl, sv)), ("t", (None, sl, sv))):
        r = O.st_direct(nd, src, trg, *args, thresh=1e-14)
        for k, v in r.items():
            assert _rel(v, g[f"{pre}st_{tag}_{k}"]) < 1e-13, (tag, k)


def test_oracle_bh_derivatives_by_finite_differences():
    """grad1 + grad2 = d vel/dz and grad3 = d vel/dzbar (central differences);
    with c2 = d2 = d3 = 0 the velocity is analytic and grad2 = 0."""
    rng = np.random.default_rng(3)
    src = rng.random((2, 6)) * 0.5
    x0 = np.array([[0.9], [0.8]])
    ch = rng.standard_normal((1, 2, 6)) + 1j * rng.standard_normal((1, 2, 6))
    dp = rng.standard_normal((1, 3, 6)) + 1j * rng.standard_normal((1, 3, 6))
    h = 1e-5

    def vel(p, c, d):
        return O.bh_direct(1, src, p, c, d, 1)["pottarg"][0, 0]
    for c, d in ((ch, dp), (ch * [[1], [0]], dp * [[1], [0], [0]])):
        g = O.bh_direct(1, src, x0, c, d, 2)["gradtarg"][0, :, 0]
        dx = (vel(x0 + [[h], [0]], c, d) - vel(x0 - [[h], [0]], c, d)) / (2 * h)
        dy = (vel(x0 + [[0], [h]], c, d) - vel(x0 - [[0], [h]], c, d)) / (2 * h)
        assert abs((dx - 1j * dy) / 2 - (g[0] + g[1])) < 1e-7 * abs(g).max()
        assert abs((dx + 1j * dy) / 2 - g[2]) < 1e-7 * abs(g).max()
    assert abs(g[1]) == 0.0


@pytest.mark.parametrize("which", ["stokeslet", "stresslet"])
def test_oracle_stokes_equations(which):
    """The velocity gradient equals central differences of the velocity, the
    velocity is divergence-free, and lap u = grad p (Stokes equations) to the
    O(H^2) accuracy of the five-point stencil."""
    rng = np.random.default_rng(4)
    src = rng.random((2, 5)) * 0.5
    x0 = np.array([[0.9], [0.8]])
    sk, sl, sv = (rng.standard_normal((1, 2, 5)) for _ in range(3))
    args = (sk, None, None) if which == "stokeslet" else (None, sl, sv)

    def at(a, b, step):
        o = O.st_direct(1, src, x0 + [[a * step], [b * step]], *args)
        return o["pottarg"][0, :, 0], o["pretarg"][0, 0], o["gradtarg"][0, :, :, 0]
    u0, p0, g0 = at(0, 0, 1.0)
    h = 1e-5
    fd = np.array([(at(1, 0, h)[0] - at(-1, 0, h)[0]) / (2 * h),
                   (at(0, 1, h)[0] - at(0, -1, h)[0]) / (2 * h)])  # fd[a, b] = d_a u_b
    assert np.abs(fd - g0).max() < 1e-7 * np.abs(g0).max()
    assert abs(g0[0, 0] + g0[1, 1]) < 1e-12 * np.abs(g0).max()
    H = 2e-3
    lap = (at(1, 0, H)[0] + at(-1, 0, H)[0] + at(0, 1, H)[0] + at(0, -1, H)[0] - 4 * u0) / H ** 2
    gp = np.array([(at(1, 0, H)[1] - at(-1, 0, H)[1]) / (2 * H),
                   (at(0, 1, H)[1] - at(0, -1, H)[1]) / (2 * H)])
    assert np.abs(lap - gp).max() < 1e-4 * np.abs(gp).max()
