"""GPU parity of the benchmark pipelines' own kernel instantiations.

The benchmark runs lfmm2d on clustered points with REAL charges at nd = 1 and
eps = 1e-6 (only the real Cauchy half is computed: the near-field kernel with
one real density, FP32 in mixed precision), with the reference's leaf capacity
and orders and with the learned ones (leaf capacity 37, orders 14: the M2L
then takes the LDS-staged FP32 translate path of orders <= 15).  These tests
run those instantiations on the same point distribution at 1.2e5 points:

  * FP64 against the reference's own FMM (lfmm2d and lfmm2d_ndiv with ndiv 37,
    laplace/lfmm2d_ndiv.f:29; rfmm2d and rfmm2d_ndiv): same tree, lists and
    orders, so the bar is max(1e-3 eps, 1e-13) relative L2 of pot+grad;
  * mixed precision, default and learned parameters, FP64 with the learned
    parameters: relative L2 error of pot+grad against the reference's FP64
    direct sum below eps (the reference's own test bar,
    test/laplace/test_lfmm2d.f), and against the reference FMM below eps.

tests/golden/headline_ref.npz (tests/golden/make_golden.py:headline_fixtures)
holds the reference's outputs at 3000 sampled sources and a hash of the
inputs, which golden/inputs.py regenerates from the seed.
"""
import numpy as np
import pytest

from golden import inputs as I
from fmm2d_amd import fmm2d as F

pytestmark = pytest.mark.gpu

# make_golden.py:HEADLINE
N, SEED, EPS, NDIV = 120000, 31, 1e-6, 37
LEARNED = {"nterms": [14] * 30, "ndiv": NDIV}


@pytest.fixture(scope="module")
def case(golden_dir):
    g = np.load(f"{golden_dir}/headline_ref.npz")
    src, q = I.clustered(N, SEED)
    assert I.digest(src, q) == str(g["input_sha256"]), "input generator changed"
    return g, src, q


def _rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def _pg(res, idx, real=False):
    p, g = res["pot"][idx], res["grad"][:, idx]
    return np.concatenate([p.ravel(), g.ravel()])


def _ref(g, tag):
    return np.concatenate([g[tag + "_pot"].ravel(), g[tag + "_grad"].ravel()])


def _direct(g, real=False):
    d = np.concatenate([g["direct_pot"].ravel(), g["direct_grad"].ravel()])
    return d.real if real else d


@pytest.mark.parametrize("opts,tag", [(None, "l_default"), ({"ndiv": NDIV}, "l_ndiv")])
def test_lfmm2d_real_fp64_matches_reference(case, opts, tag):
    g, src, q = case
    res, ier = F._fmm("l", EPS, None, src, q + 0j, None, None, None, 2, 0, 1, opts=opts)
    assert ier == 0
    got = _pg(res, g["idx"])
    assert _rel(got, _ref(g, tag)) < max(1e-3 * EPS, 1e-13)
    # real charges: the imaginary parts are exactly those of the reference (zero)
    assert np.array_equal(got.imag == 0, _ref(g, tag).imag == 0)


@pytest.mark.parametrize("opts,tag", [(None, "r_default"), ({"ndiv": NDIV}, "r_ndiv")])
def test_rfmm2d_fp64_matches_reference(case, opts, tag):
    g, src, q = case
    res, ier = F._fmm("r", EPS, None, src, q, None, None, None, 2, 0, 1, opts=opts)
    assert ier == 0
    assert _rel(_pg(res, g["idx"]), _ref(g, tag)) < max(1e-3 * EPS, 1e-13)


PIPELINES = {
    "mixed": {"precision": "mixed"},
    "mixed+learned": dict(LEARNED, precision="mixed"),
    "fp64+learned": dict(LEARNED, precision="fp64"),
}


@pytest.mark.parametrize("name", list(PIPELINES))
def test_lfmm2d_pipelines_vs_direct(case, name):
    """The benchmark's pipelines on real charges: error against the reference's
    FP64 direct sum and against the reference FMM within eps."""
    g, src, q = case
    res, ier = F._fmm("l", EPS, None, src, q + 0j, None, None, None, 2, 0, 1,
                      opts=dict(PIPELINES[name]))
    assert ier == 0
    got = _pg(res, g["idx"])
    e_dir = _rel(got, _direct(g))
    e_ref = _rel(got, _ref(g, "l_ndiv" if "learned" in name else "l_default"))
    assert e_dir < EPS, (name, e_dir)
    assert e_ref < EPS, (name, e_ref)
    assert np.all(np.isfinite(got))


@pytest.mark.parametrize("name", ["mixed", "mixed+learned"])
def test_rfmm2d_mixed_vs_direct(case, name):
    g, src, q = case
    res, ier = F._fmm("r", EPS, None, src, q, None, None, None, 2, 0, 1,
                      opts=dict(PIPELINES[name]))
    assert ier == 0
    got = _pg(res, g["idx"])
    assert _rel(got, _direct(g, real=True)) < EPS
    assert _rel(got, _ref(g, "r_ndiv" if "learned" in name else "r_default")) < EPS


def test_learned_orders_reach_the_lds_translate_path(case):
    """Orders 14 give 15 coefficients per expansion: the FP32 M2L then stages
    source rows in LDS (the path of orders <= 15); the default orders (18) do
    not.  Both must give the same result within eps, and the work counters must
    report the orders that were asked for."""
    g, src, q = case
    o = dict(PIPELINES["mixed+learned"], count_work=1)
    res, ier = F._fmm("l", EPS, None, src, q + 0j, None, None, None, 2, 0, 1, opts=o)
    assert ier == 0
    # M2L flops per list-2 pair = 8 (p+1)^2 nd with p = 14, nd = 1 (real charges)
    assert o["m2l_pairs"] > 0
    assert o["work"][3] == pytest.approx(8 * 15 ** 2 * o["m2l_pairs"], rel=1e-12)


def test_nterms_override_is_validated():
    """Orders outside the supported range are rejected before any device work
    with ier = 13 (Laplace: 2..200; Helmholtz: 1..2000)."""
    rng = np.random.default_rng(5)
    src = rng.random((2, 500))
    q = rng.standard_normal(500) + 0j
    for bad in ([1], [0] * 5, [201], [18, 18, -3]):
        _, ier = F._fmm("l", 1e-6, None, src, q, None, None, None, 1, 0, 1,
                        opts={"nterms": bad})
        assert ier == 13, (bad, ier)
    _, ier = F._fmm("h", 1e-6, 10.0, src, q, None, None, None, 1, 0, 1,
                    opts={"nterms": [2001]})
    assert ier == 13
